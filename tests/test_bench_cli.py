"""bench.py's contract pieces that need no GPU: the one-GPU refusal, the step
count check, and the error metrics of its JSON line (the reference probe's
definitions: accessible poses E < 2000 K and |E| < 1e4 K, relative errors with
a 1 K and a 100 K floor)."""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _bench(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True,
                          text=True, timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    proc = _bench("--gpus", "2")
    assert proc.returncode == 2
    assert proc.stdout == ""                  # no JSON line
    assert "one GPU" in proc.stderr


def test_zero_steps_is_refused():
    proc = _bench("--steps", "0")
    assert proc.returncode != 0 and proc.stdout == ""


def test_error_metrics_definitions():
    sys.path.insert(0, str(ROOT))
    import bench
    ref = np.array([-500.0, 0.5, 1999.0, 2500.0, -2e4, 50.0])
    cand = ref + np.array([1e-3, 2e-3, -4e-3, 1.0, 2.0, 0.0])
    m = bench.error_metrics(ref, cand)
    # accessible: -500, 0.5, 1999 and 50 (2500 is above 2000 K, -2e4 beyond 1e4 K)
    assert m["accessible_poses"] == 4
    assert m["accessible_max_abs_error_K"] == pytest.approx(4e-3)
    assert m["accessible_mae_K"] == pytest.approx((1e-3 + 2e-3 + 4e-3 + 0.0) / 4)
    assert m["accessible_rms_error_K"] == pytest.approx(np.sqrt((1e-6 + 4e-6 + 16e-6) / 4))
    # relative, 1 K floor: the 0.5 K pose reads 2e-3 / 1; 2500 K reads 1 / 2500
    assert m["all_max_relative_error_floor1K"] == pytest.approx(2e-3)
    # 100 K floor: 1 / 2500 = 4e-4 is the largest
    assert m["all_max_relative_error_floor100K"] == pytest.approx(4e-4)


def test_traffic_table_covers_every_tile_candidate():
    """The bench line's roofline.traffic comes from the PMC table of the tile
    size the autotuner picked: every candidate has an entry, each read + write
    and above the algorithmic bytes."""
    import json
    from string_vext_amd import autotune
    table = json.loads((ROOT / "profiles" / "field_mixed_60x64_traffic.json").read_text())
    for tc in autotune.TILE_CANDIDATES:
        rec = table["by_tile"][str(tc)]
        assert rec["hbm_bytes_per_launch"] == rec["read_bytes"] + rec["write_bytes"]
        assert rec["hbm_bytes_per_launch"] > table["algorithmic_write_bytes"]
