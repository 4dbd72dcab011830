"""Autotuner for the field launch: tile size and persistent grid size.

The tunables are
  * ``tile_centers`` -- pose centers per spatial tile.  Larger tiles mean
    fewer candidate lists but longer ones (each center streams and filters its
    tile's whole list); smaller tiles the opposite, plus more list-building.
  * ``blocks_per_cu`` -- size of the persistent grid of a ragged batch (0: the
    plain grid of one block per 4 work items, dispatched by the hardware).
Candidates are timed on a reduced grid of a sample of the workload's own
frameworks (HIP events on the batch stream: list build + field kernel), the
fastest wins, and results are cached per (device, density class) in
``$XDG_CACHE_HOME/string_vext_amd/autotune.json`` (or ``~/.cache/...``).
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np

from .field import Batch, Plan
from .orientation import regular_pose_grid

TILE_CANDIDATES = (16, 32, 64, 128, 256)
BLOCK_CANDIDATES = (0, 4, 5, 8)     # 0: one block per 4 work items, no queue


def cache_path() -> Path:
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return Path(base) / "string_vext_amd" / "autotune.json"


def density_class(case) -> str:
    """Candidates per center grow with atom density; bucket by decade."""
    rho = len(case["frame_frac"]) / float(np.linalg.det(np.asarray(case["cell"])))
    return f"rho{int(np.floor(np.log10(max(rho, 1e-6)) * 2))}"


def _time_batch(jobs, precision, blocks_per_cu, reps):
    """Median over `reps` runs of the whole device-side run (HIP events around
    it: list count, scan and fill, the field kernel, the epilogue and the gaps
    between them), the cost a step pays for the tile size."""
    batch = Batch(jobs, precision=precision, outputs=("orientation",))
    if blocks_per_cu:
        batch.set_persistent(blocks_per_cu)
    batch.run(sync=True)
    runs = []
    for _ in range(reps):
        batch.run(time_kernels=True, sync=True)
        runs.append(batch.timings()["run"])
    batch.close()
    return float(np.median(runs))


def tune(cases, size=24, orientations=64, precision="mixed", reps=3, use_cache=True,
         device_key="gpu0", persistent=True):
    """Pick (tile_centers, blocks_per_cu) for a set of cases.

    Returns ``{"tile_centers": {density_class: n}, "blocks_per_cu": b, "trials": [...]}``.
    """
    path = cache_path()
    cache = {}
    if use_cache and path.exists():
        try:
            cache = json.loads(path.read_text())
        except (OSError, ValueError):
            cache = {}
    _, axes = regular_pose_grid(size, orientations)
    classes = {}
    for case in cases:
        classes.setdefault(density_class(case), []).append(case)
    result = {"tile_centers": {}, "trials": []}
    for cls, members in sorted(classes.items()):
        key = f"{device_key}/{precision}/{cls}/n{size}k{orientations}"
        if key in cache:
            result["tile_centers"][cls] = cache[key]
            continue
        sample = members[:3]
        plans = [Plan.from_case(c) for c in sample]
        timing = {}
        for tc in TILE_CANDIDATES:
            jobs = [(p, p.grid(size, axes, tile_centers=tc)) for p in plans]
            timing[tc] = _time_batch(jobs, precision, 0, reps)
            result["trials"].append({"class": cls, "tile_centers": tc, "ms": timing[tc]})
        best = min(timing, key=timing.get)
        result["tile_centers"][cls] = best
        cache[key] = best
    bkey = f"{device_key}/{precision}/blocks_per_cu/n{size}k{orientations}"
    if bkey in cache:
        result["blocks_per_cu"] = cache[bkey]
    elif precision == "mixed" and persistent:
        plans = [Plan.from_case(c) for c in cases[: min(len(cases), 6)]]
        jobs = [(p, p.grid(size, axes, tile_centers=result["tile_centers"][density_class(c)]))
                for p, c in zip(plans, cases)]
        timing = {b: _time_batch(jobs, precision, b, reps) for b in BLOCK_CANDIDATES}
        for b, ms in timing.items():
            result["trials"].append({"blocks_per_cu": b, "ms": ms})
        result["blocks_per_cu"] = cache[bkey] = min(timing, key=timing.get)
    else:
        result["blocks_per_cu"] = 0
    if use_cache:
        try:
            path.parent.mkdir(parents=True, exist_ok=True)
            path.write_text(json.dumps(cache, indent=1, sort_keys=True))
        except OSError:
            pass
    return result
