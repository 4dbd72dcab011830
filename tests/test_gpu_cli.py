"""End to end through the reference-shaped API on the GPU: calculate_vext's NPZ
channels (src/string_vext/vext.py:21-43) against the oracle, the CLI's output
files (vext.py:46-80), and NativeReusePotential's two routes (native.py:26-68)."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CHANNELS = {"orientation_K": np.float32, "marginal_K": np.float32, "site_K": np.float32,
            "axes": np.float32, "cell_matrix": np.float64, "atom_fractional": np.float64,
            "framework_sigma_A": np.float64, "framework_epsilon_K": np.float64,
            "grid": np.float32}


@pytest.fixture(scope="module")
def case():
    from string_vext_amd import synthetic
    return synthetic.small_field_case()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_calculate_vext_channels_match_oracle(oracle, case, precision):
    from string_vext_amd.vext import calculate_vext
    size, k = 5, 64
    out = calculate_vext(case, size, k, precision=precision)
    ref = oracle.vext_fields(case, size, k)
    assert set(out) == set(CHANNELS)
    for name, dtype in CHANNELS.items():
        assert out[name].dtype == dtype, name
    assert out["orientation_K"].shape == (k, size, size, size)
    assert out["site_K"].shape == (2, size, size, size)
    assert out["marginal_K"].shape == out["grid"].shape == (size, size, size)
    np.testing.assert_allclose(out["orientation_K"], ref["orientation_K"], rtol=1e-6, atol=1e-3)
    np.testing.assert_allclose(out["marginal_K"], ref["marginal_K"], rtol=1e-6, atol=1e-3)
    np.testing.assert_allclose(out["site_K"], ref["site_K"], rtol=1e-6, atol=1e-3)
    np.testing.assert_allclose(out["grid"], ref["grid"], rtol=1e-5, atol=1e-6)
    np.testing.assert_array_equal(out["axes"], ref["axes"].astype(np.float32))


def test_cli_writes_npz_and_provenance_once(case, tmp_path):
    from string_vext_amd.input import write_string_input
    from string_vext_amd.vext import main
    inp = tmp_path / "input.dat"
    write_string_input(case, inp)
    out = tmp_path / "run" / "vext.npz"
    argv = ["--input", str(inp), "--output", str(out), "--size", "4", "--orientations", "64",
            "--precision", "mixed"]
    main(argv)
    with np.load(out) as npz:           # written by this repository's own code
        assert set(npz.files) == set(CHANNELS)
        assert npz["orientation_K"].shape == (64, 4, 4, 4)
    record = json.loads(out.with_suffix(".json").read_text())
    for key in ("input_sha256", "output_sha256", "kernel_sha256", "size", "orientations",
                "energy_mode", "cell_lengths_A", "cell_angles_deg"):
        assert key in record
    assert record["size"] == 4 and record["precision"] == "mixed"
    # hashed while written: the digest of the file as it lies on disk
    import hashlib
    assert record["output_sha256"] == hashlib.sha256(out.read_bytes()).hexdigest()
    with pytest.raises(FileExistsError):
        main(argv)


def test_native_reuse_potential_both_routes(oracle, case):
    """The batched route and the reference's 512-center CSR route through the
    drop-in symbol give the same energies."""
    from string_vext_amd import _lib
    from string_vext_amd.native import NativeReusePotential
    centers, axes = oracle.regular_pose_grid(6, 64)
    ref = oracle.reuse_energy(case, centers, axes)
    for batched in (True, False):
        pot = NativeReusePotential(case, _lib.LIB_PATH, batched=batched)
        e = pot.evaluate(centers, axes)
        assert (np.abs(e - ref) / np.maximum(np.abs(ref), 1.0)).max() < 1e-12
