"""The boundary ships as a package whose ``string-vext`` console script (the
reference's script name, its pyproject.toml [project.scripts]; here in
setup.cfg, which setuptools 59 reads) resolves to this
package's CLI, which keeps the reference's refusal to overwrite an output."""
import importlib
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _scripts():
    """console_scripts of setup.cfg (name = module:attr per line)."""
    import configparser
    cfg = configparser.ConfigParser()
    cfg.read(ROOT / "setup.cfg")
    lines = cfg["options.entry_points"]["console_scripts"].strip().splitlines()
    return dict((s.strip() for s in line.split("=", 1)) for line in lines)


def test_console_script_resolves_to_the_cli():
    target = _scripts()["string-vext"]
    module, _, attr = target.partition(":")
    main = getattr(importlib.import_module(module), attr)
    from string_vext_amd.vext import main as cli
    assert main is cli


def test_console_script_refuses_an_existing_output(tmp_path):
    module, _, attr = _scripts()["string-vext"].partition(":")
    main = getattr(importlib.import_module(module), attr)
    out = tmp_path / "vext.npz"
    out.write_bytes(b"kept")
    with pytest.raises(FileExistsError, match="already exists"):
        main(["--input", str(tmp_path / "input.dat"), "--output", str(out)])
    assert out.read_bytes() == b"kept"
    record = tmp_path / "other.json"
    record.write_text("{}")
    with pytest.raises(FileExistsError):
        main(["--input", str(tmp_path / "input.dat"), "--output", str(tmp_path / "other.npz")])


def test_npz_is_published_whole_and_never_overwritten(tmp_path):
    import numpy as np
    from string_vext_amd.vext import _publish_npz
    import hashlib
    out = tmp_path / "sub" / "vext.npz"
    digest = _publish_npz(out, {"a": np.arange(3.0)})
    assert digest == hashlib.sha256(out.read_bytes()).hexdigest()
    with np.load(out) as z:
        assert z["a"].tolist() == [0.0, 1.0, 2.0]
    with pytest.raises(FileExistsError):
        _publish_npz(out, {"a": np.zeros(2)})
    with np.load(out) as z:
        assert z["a"].tolist() == [0.0, 1.0, 2.0]
    assert sorted(p.name for p in out.parent.iterdir()) == ["vext.npz"]   # no staging files left


def test_non_fp64_precision_needs_the_batched_library():
    """A --kernel library that only exports evaluate_reused_neighbors (here the
    reference kernel.cpp built from source) evaluates FP64; asking it for
    another precision is refused before any evaluation."""
    kernel = ROOT / "oracle" / "_ref" / "kernel.so"
    if not kernel.exists():
        pytest.skip("oracle/_ref/kernel.so is built only where the reference tree is present")
    from string_vext_amd import synthetic
    from string_vext_amd.vext import calculate_vext
    case = synthetic.toy_case()
    for precision in ("mixed", "fp64_rcp"):
        with pytest.raises(ValueError, match="batched entry points"):
            calculate_vext(case, size=2, orientations=4, kernel=kernel, precision=precision)
