"""The autotuner's choice logic, with the timing replaced by a fixed table
(CPU): the fastest tile size per density class and the fastest persistent
grid win, choices are cached per device key, precision, class and grid shape,
and a cached choice runs no trial."""
import numpy as np
import pytest


class _FakePlan:
    def __init__(self, case):
        self.case = case

    @classmethod
    def from_case(cls, case):
        return cls(case)

    def grid(self, size, axes, tile_centers):
        return {"size": size, "k": len(axes), "tile_centers": tile_centers}


TILE_MS = {16: 5.0, 32: 3.0, 64: 2.0, 128: 4.0, 256: 6.0}
BLOCK_MS = {0: 0.0, 4: -0.5, 5: 0.2, 8: 1.0}


@pytest.fixture
def tuner(monkeypatch, tmp_path):
    from string_vext_amd import autotune
    monkeypatch.setenv("XDG_CACHE_HOME", str(tmp_path))
    monkeypatch.setattr(autotune, "Plan", _FakePlan)
    calls = []

    def fake_time(jobs, precision, blocks_per_cu, reps):
        tiles = {poses["tile_centers"] for _, poses in jobs}
        assert len(tiles) == 1
        calls.append((tiles.pop(), blocks_per_cu, len(jobs)))
        return TILE_MS[calls[-1][0]] + BLOCK_MS[blocks_per_cu]

    monkeypatch.setattr(autotune, "_time_batch", fake_time)
    return autotune, calls


def test_density_class_buckets_by_half_decade():
    from string_vext_amd import autotune, synthetic
    case = synthetic.full_field_case()                     # 1200 atoms / 16,600 A^3
    rho = len(case["frame_frac"]) / np.linalg.det(case["cell"])
    assert autotune.density_class(case) == f"rho{int(np.floor(np.log10(rho) * 2))}" == "rho-3"
    sparse = dict(case, frame_frac=np.asarray(case["frame_frac"])[:120])   # 10x sparser
    assert autotune.density_class(sparse) == "rho-5"


def test_tune_picks_the_fastest_and_caches_it(tuner):
    from string_vext_amd import synthetic
    autotune, calls = tuner
    cases = [synthetic.full_field_case(seed=synthetic.FULL_SEED + k) for k in range(4)]
    res = autotune.tune(cases, size=8, orientations=8)
    cls = autotune.density_class(cases[0])
    assert res["tile_centers"] == {cls: 64}
    assert res["blocks_per_cu"] == 4
    # every tile size on three of the class's frameworks, then every persistent
    # grid shape at the chosen tile size on the first frameworks
    assert calls[:len(TILE_MS)] == [(tc, 0, 3) for tc in autotune.TILE_CANDIDATES]
    assert calls[len(TILE_MS):] == [(64, b, 4) for b in autotune.BLOCK_CANDIDATES]
    assert autotune.cache_path().exists()
    calls.clear()
    again = autotune.tune(cases, size=8, orientations=8)
    assert calls == []
    assert again["tile_centers"] == res["tile_centers"] and again["blocks_per_cu"] == 4
    # another grid shape or precision is tuned afresh
    autotune.tune(cases, size=12, orientations=8, precision="fp64")
    assert calls and all(c[1] == 0 for c in calls)


def test_tune_without_cache_or_persistent(tuner):
    from string_vext_amd import synthetic
    autotune, calls = tuner
    res = autotune.tune([synthetic.small_field_case()], size=8, orientations=8,
                        use_cache=False, persistent=False)
    assert list(res["tile_centers"].values()) == [64]
    assert res["blocks_per_cu"] == 0
    assert all(c[1] == 0 for c in calls)
    assert not autotune.cache_path().exists()
