"""Ragged batches: many frameworks and guests in one (persistent) launch give
the same fields as one launch per field, and match the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ragged():
    from string_vext_amd import synthetic
    cases = synthetic.ragged_cases(n_frameworks=4, guests_per_framework=2, seed=77)
    return cases


def _jobs(cases, sizes, ks, tile_centers=64):
    from string_vext_amd.field import Plan
    from string_vext_amd.orientation import regular_pose_grid
    jobs = []
    for case, size, k in zip(cases, sizes, ks):
        plan = Plan.from_case(case)
        _, axes = regular_pose_grid(size, k)
        jobs.append((plan, plan.grid(size, axes, tile_centers=tile_centers)))
    return jobs


@pytest.mark.parametrize("persistent", [0, 8])
def test_batch_equals_single_fields(ragged, persistent):
    from string_vext_amd.field import Batch
    sizes = [6, 5, 7, 4, 6, 5, 8, 4]
    ks = [64, 64, 32, 100, 64, 16, 64, 64]
    jobs = _jobs(ragged, sizes, ks)
    temps = [c["temperature"] for c in ragged]
    batch = Batch(jobs, precision="mixed", temperatures=temps,
                  outputs=("energy", "orientation", "marginal", "grid", "site"))
    if persistent:
        batch.set_persistent(persistent)
    batch.run(sync=True)
    for j, (plan, poses) in enumerate(jobs):
        single = plan.field(poses, precision="mixed", temperature=temps[j], orientation=True,
                            marginal=True, grid=True, site=True)
        assert np.array_equal(batch.fetch("energy", j), single["energy"])
        assert np.array_equal(batch.fetch("orientation", j), single["orientation"])
        assert np.array_equal(batch.fetch("site", j), single["site"])
        assert np.array_equal(batch.fetch("marginal", j), single["marginal"])


def test_batch_matches_oracle_and_tiles_do_not_change_results(oracle, ragged):
    from string_vext_amd.field import Batch
    from string_vext_amd.orientation import regular_pose_grid
    cases = ragged[:3]
    sizes, ks = [5, 4, 5], [64, 64, 64]
    ref = [oracle.reuse_energy(c, *regular_pose_grid(s, k)) for c, s, k in zip(cases, sizes, ks)]
    for tc in (16, 64, 256):
        # mixed: 1e-5 relative with a 100 K floor (at most 1e-3 K below 100 K)
        for precision, tol, floor in (("fp64", 1e-10, 1.0), ("mixed", 1e-5, 100.0)):
            jobs = _jobs(cases, sizes, ks, tile_centers=tc)
            batch = Batch(jobs, precision=precision, outputs=("energy",))
            if precision == "mixed":
                batch.set_persistent(4)
            batch.run(sync=True)
            for j in range(len(jobs)):
                got = batch.fetch("energy", j)
                rel = np.abs(got - ref[j]) / np.maximum(np.abs(ref[j]), floor)
                assert rel.max() < tol, (tc, precision, j, rel.max())
                if precision == "mixed":
                    m = oracle.error_metrics(ref[j], got)
                    if m["accessible_poses"]:
                        assert m["accessible_max_abs_error_K"] < 1e-3


def test_autotune_times_real_batches(ragged, tmp_path, monkeypatch):
    """The autotuner on the GPU: every candidate tile size and persistent grid
    is timed (positive finite milliseconds), the choice is a candidate, and the
    field at the chosen tile size equals the oracle-checked tile-64 field."""
    from string_vext_amd import autotune
    from string_vext_amd.field import Batch
    monkeypatch.setenv("XDG_CACHE_HOME", str(tmp_path))
    res = autotune.tune(ragged[:3], size=8, orientations=16, reps=1)
    assert res["blocks_per_cu"] in autotune.BLOCK_CANDIDATES
    assert set(res["tile_centers"].values()) <= set(autotune.TILE_CANDIDATES)
    tiles = [t for t in res["trials"] if "tile_centers" in t]
    assert len(tiles) == len(autotune.TILE_CANDIDATES) * len(res["tile_centers"])
    assert all(np.isfinite(t["ms"]) and t["ms"] > 0 for t in res["trials"])
    case = ragged[0]
    tile = res["tile_centers"][autotune.density_class(case)]
    out = {}
    for tc in (64, tile):
        jobs = _jobs([case], [8], [16], tile_centers=tc)
        batch = Batch(jobs, precision="fp64", outputs=("energy",))
        batch.run(sync=True)
        out[tc] = batch.fetch("energy")
        batch.close()
    np.testing.assert_allclose(out[tile], out[64], rtol=1e-12, atol=1e-9)
