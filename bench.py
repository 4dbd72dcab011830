"""Throughput of the guest-framework LJ Vext field on one MI355X.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]

Workload: the full-size single Vext field -- a 60^3 grid x 64 orientations
(13.824 M poses) over a seeded 1200-atom triclinic framework (37 x 37 x 14 A,
gamma = 120 deg, 4 LJ types), two-site guest (eps 98 K, sigma 3.75 A, bond
1.54 A), cutoff 12.8 A.  One step = one whole field: device candidate lists,
the mixed-precision field kernel (FP32 pair arithmetic, FP64 accumulation,
FP64 near-contact terms) and the fused epilogue (orientation_K, marginal,
asinh grid, site field), with the framework tables already resident in HBM.

value = exact in-cutoff pair-evals per field / seconds per step.  The same
run times the plain FP64 form, reports the mixed form's max/RMS error in K
against it, the field kernel's fraction of the FP32 vector peak (HIP events
around the kernel on its own stream), and the reference CPU kernel
(kernel.cpp compiled from source, oracle/_ref) on a bounded sample of the
same workload on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "pair-evals/s, 60³×64 Vext, mixed-FP32 vs FP64 GPU at ≤1e-3 K max error"
FLOP_PER_PAIR = 18          # forward pair-eval, FMA = 2, max/select/rcp/add = 1
PEAK_FP32_TFLOPS = 157.3    # MI355X FP32 vector (AMD published)
PEAK_FP64_TFLOPS = 78.6
SIZE, ORIENTATIONS = 60, 64


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_field(plan, poses, precision, steps, warmup, temperature):
    """Seconds per field for `steps` back-to-back fields (synchronized before and
    after the timed region), then the per-kernel durations from HIP events on
    the batch's stream over a second pass of the same steps."""
    from string_vext_amd.field import Batch
    batch = Batch([(plan, poses)], precision=precision, temperatures=temperature)
    for _ in range(warmup):
        batch.run(time_kernels=False, sync=True)
    batch.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.run(time_kernels=False, sync=False)
    batch.sync()
    wall = time.perf_counter() - t0
    field_ms, run_ms = [], []
    for _ in range(steps):
        batch.run(time_kernels=True, sync=True)
        t = batch.timings()
        field_ms.append(t["field"])
        run_ms.append(t["run"])
    timings = batch.timings()
    batch.close()
    return wall / steps, float(np.mean(field_ms)), float(np.mean(run_ms)), timings


def energies(plan, poses, precision):
    from string_vext_amd.field import Batch
    batch = Batch([(plan, poses)], precision=precision, outputs=("energy",))
    batch.run(sync=True)
    e = batch.fetch("energy")
    batch.close()
    return e


def error_metrics(reference, candidate):
    diff = np.abs(reference - candidate)
    acc = (reference < 2000.0) & (np.abs(reference) < 1e4)
    rel = diff / np.maximum(np.abs(reference), 1.0)
    return {"accessible_poses": int(acc.sum()),
            "accessible_max_abs_error_K": float(diff[acc].max()),
            "accessible_rms_error_K": float(np.sqrt((diff[acc] ** 2).mean())),
            "all_max_relative_error_floor1K": float(rel.max())}


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, centers, axes, sample):
    """Reference kernel.cpp (built from source into oracle/_ref) or, if absent, the
    oracle's C restatement, single-threaded on `sample` grid centers."""
    from oracle import vext_oracle as O
    kind = "reference" if O.reference_kernel() is not None else "port"
    fn = O.reference_kernel() if kind == "reference" else O.lib().oracle_reuse
    setup = O.ReuseSetup(case)
    idx = np.linspace(0, len(centers) - 1, sample).round().astype(int)
    cart = np.mod(centers[idx], 1) @ setup.cell
    unit = axes / np.linalg.norm(axes, axis=1, keepdims=True)
    query = compute = 0.0
    pairs = 0
    for start in range(0, len(cart), 512):
        chunk = np.ascontiguousarray(cart[start:start + 512])
        t = time.perf_counter()
        starts, indices = O.csr(setup.tree, chunk, setup.radius)
        query += time.perf_counter() - t
        t = time.perf_counter()
        O.run_reuse_kernel(fn, chunk, unit, setup, starts, indices)
        compute += time.perf_counter() - t
        pairs += O.lib().oracle_count_pairs(
            len(chunk), len(unit), setup.cutoff ** 2, O._ptr(chunk), O._ptr(np.ascontiguousarray(unit)),
            O._ptr(np.ascontiguousarray(setup.images)), O._ptr(np.ascontiguousarray(setup.offsets)),
            O._ptr(starts), O._ptr(indices))
    return {"value": pairs / compute, "unit": "pair-evals/s", "cores": 1, "kind": kind,
            "sample": f"{sample} of the {len(centers)} grid centers (evenly spaced) x {len(axes)} "
                      f"orientations, {pairs} in-cutoff pairs; kernel {compute:.2f} s, cKDTree "
                      f"candidate query {query:.2f} s; {cpu_model()}",
            "query_seconds": query, "compute_seconds": compute}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fp64-steps", type=int, default=5)
    ap.add_argument("--cpu-sample", type=int, default=16384,
                    help="grid centers in the CPU baseline sample (0 skips it)")
    ap.add_argument("--no-check", action="store_true", help="skip the mixed-vs-FP64 error pass")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run a plain elementwise copy of MIB MiB (FP32 in, FP32 out) so that "
                         "PMC byte counters can be scaled against a kernel of known traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); every pose is independent, "
              "so N GPUs would each evaluate a slab of centers -- see the sharding note in the "
              "design document.", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1:
        sys.exit("--steps must be >= 1")

    from string_vext_amd import synthetic
    from string_vext_amd.field import Plan
    from string_vext_amd.orientation import regular_pose_grid

    case = synthetic.full_field_case()
    centers, axes = regular_pose_grid(SIZE, ORIENTATIONS)
    plan = Plan.from_case(case)
    poses = plan.grid(SIZE, axes)
    pairs = plan.census(poses)
    log(f"census: {pairs} in-cutoff pair-evals per field ({pairs / (SIZE**3 * ORIENTATIONS):.1f} per pose)")

    step_s, field_ms, run_ms, t_mixed = time_field(plan, poses, "mixed", args.steps, args.warmup,
                                                   case["temperature"])
    log(f"mixed: {step_s * 1e3:.3f} ms/step, field kernel {field_ms:.3f} ms, phases {t_mixed}")
    f64_steps = max(1, args.fp64_steps)
    step64, field64, run64, t64 = time_field(plan, poses, "fp64", f64_steps, 1, case["temperature"])
    log(f"fp64: {step64 * 1e3:.3f} ms/step, field kernel {field64:.3f} ms")

    errors = None
    if not args.no_check:
        e64 = energies(plan, poses, "fp64")
        emx = energies(plan, poses, "mixed")
        errors = error_metrics(e64, emx)
        log(f"mixed vs fp64: {errors}")

    achieved = pairs * FLOP_PER_PAIR / (field_ms * 1e-3) / 1e12
    traffic = None
    prof = ROOT / "profiles" / "field_mixed_60x64_traffic.json"
    if prof.exists():
        traffic = json.loads(prof.read_text()).get("hbm_bytes_per_launch")

    cpu = None
    if args.cpu_sample > 0:
        log(f"cpu baseline on {args.cpu_sample} centers ...")
        cpu = cpu_baseline(case, centers, axes, args.cpu_sample)
        log(f"cpu: {cpu['value']:.3e} pair-evals/s")

    if args.copy_calibration > 0:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        log(f"copy calibration: {n * 4} bytes read, {n * 4} bytes written")

    value = pairs / step_s
    line = {
        "metric": METRIC, "value": value, "unit": "pair-evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": step_s * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded framework and guest; no dataset)",
        "config": {"workload": "full-size single Vext field: 60^3 grid x 64 orientations, "
                               "1200-atom triclinic framework 37x37x14 A gamma 120, cutoff 12.8 A, "
                               "mixed FP32 pair / FP64 accumulate",
                   "grid": SIZE, "orientations": ORIENTATIONS, "framework_atoms": 1200,
                   "pair_evals_per_field": pairs, "accumulate": "f64"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "kernel": "field_mixed_kernel", "kernel_ms": field_ms,
                     "flop_per_pair": FLOP_PER_PAIR},
        "cpu_baseline": cpu and {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")},
        "fp64": {"value": pairs / step64, "ms_per_step": step64 * 1e3, "kernel_ms": field64,
                 "frac_of_fp64_peak": pairs * FLOP_PER_PAIR / (field64 * 1e-3) / 1e12 / PEAK_FP64_TFLOPS},
        "speedup_mixed_vs_fp64": step64 / step_s,
        "error_vs_fp64": errors,
        "phases_ms": t_mixed,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
