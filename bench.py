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

The default line also carries summaries of the other workloads, each with its
pair census, time per step, roofline fraction and error against FP64:
"ragged" (100 frameworks x 2 guests in one launch), "poses" (8 x 65,536
differentiable poses, energy and gradient through torch.autograd), "strings" (64
strings x 96 images, 2000 HIP-graph iterations) and "e2e" (the string-vext
CLI from input.dat to the NPZ and its provenance JSON).  --workload NAME
prints one workload's full line instead.
"""
from __future__ import annotations

import argparse
import contextlib
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
FLOP_PER_GRAD_PAIR = 28     # forward terms + dphi/d(d^2) and its 6 gradient FMAs
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
    rel100 = diff / np.maximum(np.abs(reference), 100.0)
    return {"accessible_poses": int(acc.sum()),
            "accessible_max_abs_error_K": float(diff[acc].max()),
            "accessible_p99_abs_error_K": float(np.quantile(diff[acc], 0.99)),
            "accessible_mae_K": float(diff[acc].mean()),
            "accessible_rms_error_K": float(np.sqrt((diff[acc] ** 2).mean())),
            "all_max_relative_error_floor1K": float(rel.max()),
            "all_max_relative_error_floor100K": float(rel100.max())}


def porous_field_audit(tile_centers):
    """The reference's full-grid audit (accessible-pose and marginal errors,
    tests/cof/reference_project/graph_vext/exact_neighbor_reuse_fullgrid.py)
    applied to mixed against FP64 on a porous variant of the full-size
    framework: 400 atoms in the same cell with the same LJ types, so that
    about 390,000 of the 13.8 M poses are accessible (107 on the 1200-atom
    framework)."""
    from string_vext_amd import synthetic
    from string_vext_amd.field import Plan
    from string_vext_amd.orientation import regular_pose_grid
    case = synthetic.full_field_case(n_atoms=400)
    _, axes = regular_pose_grid(SIZE, ORIENTATIONS)
    plan = Plan.from_case(case)
    poses = plan.grid(SIZE, axes, tile_centers=tile_centers)
    out = {prec: plan.field(poses, precision=prec, temperature=case["temperature"], marginal=True)
           for prec in ("fp64", "mixed")}
    res = error_metrics(out["fp64"]["energy"], out["mixed"]["energy"])
    m64 = out["fp64"]["marginal"].astype(np.float64)
    dm = np.abs(out["mixed"]["marginal"].astype(np.float64) - m64)
    acc = (m64 < 2000.0) & (np.abs(m64) < 1e4)
    res.update({"workload": "60^3 x 64 field, 400-atom framework in the 37x37x14 A gamma 120 cell",
                "accessible_marginal_centers": int(acc.sum()),
                "marginal_max_abs_error_K": float(dm[acc].max()) if acc.any() else None,
                "marginal_p99_abs_error_K": float(np.quantile(dm[acc], 0.99)) if acc.any() else None})
    return res


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, centers, axes, sample, keep_output=False):
    """Reference kernel.cpp (built from source into oracle/_ref) or, if absent, the
    oracle's C restatement, single-threaded on `sample` grid centers (with
    keep_output, their energies are returned under "energy")."""
    from oracle import vext_oracle as O
    kind = "reference" if O.reference_kernel() is not None else "port"
    fn = O.reference_kernel() if kind == "reference" else O.lib().oracle_reuse
    setup = O.ReuseSetup(case)
    idx = np.linspace(0, len(centers) - 1, sample).round().astype(int)
    cart = np.mod(centers[idx], 1) @ setup.cell
    unit = axes / np.linalg.norm(axes, axis=1, keepdims=True)
    query = compute = 0.0
    pairs = 0
    out = []
    for start in range(0, len(cart), 512):
        chunk = np.ascontiguousarray(cart[start:start + 512])
        t = time.perf_counter()
        starts, indices = O.csr(setup.tree, chunk, setup.radius)
        query += time.perf_counter() - t
        t = time.perf_counter()
        e = O.run_reuse_kernel(fn, chunk, unit, setup, starts, indices)
        compute += time.perf_counter() - t
        if keep_output:
            out.append(e)
        pairs += O.lib().oracle_count_pairs(
            len(chunk), len(unit), setup.cutoff ** 2, O._ptr(chunk), O._ptr(np.ascontiguousarray(unit)),
            O._ptr(np.ascontiguousarray(setup.images)), O._ptr(np.ascontiguousarray(setup.offsets)),
            O._ptr(starts), O._ptr(indices))
    res = {"value": pairs / compute, "unit": "pair-evals/s", "cores": 1, "kind": kind,
           "sample": f"{sample} of the {len(centers)} grid centers (evenly spaced) x {len(axes)} "
                     f"orientations, {pairs} in-cutoff pairs; kernel {compute:.2f} s, cKDTree "
                     f"candidate query {query:.2f} s; {cpu_model()}",
           "query_seconds": query, "compute_seconds": compute}
    if keep_output:
        res["energy"] = np.concatenate(out)
    return res


def small_field_report(threads):
    """The 20^3 x 64 field of the 300-atom framework (25 x 25 x 12 A, gamma 120): GPU
    mixed and FP64 against the reference CPU kernel's whole field on the same inputs,
    and that kernel's single-core and all-core rates."""
    from string_vext_amd import synthetic
    from string_vext_amd.field import Plan
    from string_vext_amd.orientation import regular_pose_grid
    case = synthetic.small_field_case()
    centers, axes = regular_pose_grid(20, 64)
    plan = Plan.from_case(case)
    poses = plan.grid(20, axes)
    pairs = plan.census(poses)
    s_mx, k_mx, _, _ = time_field(plan, poses, "mixed", 20, 3, case["temperature"])
    s_64, k_64, _, _ = time_field(plan, poses, "fp64", 10, 2, case["temperature"])
    cpu1 = cpu_baseline(case, centers, axes, len(centers), keep_output=True)
    cpun = cpu_baseline_threads(case, centers, axes, len(centers), threads)
    ref = cpu1.pop("energy")
    e64 = energies(plan, poses, "fp64")
    emx = energies(plan, poses, "mixed")
    rel64 = np.abs(e64 - ref) / np.maximum(np.abs(ref), 1.0)
    return {"workload": "20^3 x 64 field, 300-atom framework 25x25x12 A gamma 120, cutoff 12.8 A",
            "pair_evals": int(pairs),
            "gpu_mixed": {"value": pairs / s_mx, "ms_per_field": s_mx * 1e3, "kernel_ms": k_mx},
            "gpu_fp64": {"value": pairs / s_64, "ms_per_field": s_64 * 1e3, "kernel_ms": k_64},
            "cpu_1_core": {k: cpu1[k] for k in ("value", "cores", "kind", "sample")},
            "cpu_all_cores": cpun,
            "fp64_vs_reference_cpu_max_relative_error_floor1K": float(rel64.max()),
            "mixed_vs_reference_cpu": error_metrics(ref, emx)}


def cpu_baseline_threads(case, centers, axes, sample, threads):
    """The same kernel on all host cores: centers split into 64-center pieces run
    by `threads` Python threads (ctypes releases the GIL during each call); the
    CSR for every piece is built first and is not part of the timed region."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import vext_oracle as O
    kind = "reference" if O.reference_kernel() is not None else "port"
    fn = O.reference_kernel() if kind == "reference" else O.lib().oracle_reuse
    setup = O.ReuseSetup(case)
    idx = np.linspace(0, len(centers) - 1, sample).round().astype(int)
    cart = np.mod(centers[idx], 1) @ setup.cell
    unit = axes / np.linalg.norm(axes, axis=1, keepdims=True)
    pieces = []
    pairs = 0
    for start in range(0, len(cart), 64):
        chunk = np.ascontiguousarray(cart[start:start + 64])
        starts, indices = O.csr(setup.tree, chunk, setup.radius)
        pieces.append((chunk, starts, indices))
        pairs += O.lib().oracle_count_pairs(
            len(chunk), len(unit), setup.cutoff ** 2, O._ptr(chunk), O._ptr(np.ascontiguousarray(unit)),
            O._ptr(np.ascontiguousarray(setup.images)), O._ptr(np.ascontiguousarray(setup.offsets)),
            O._ptr(starts), O._ptr(indices))
    t = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as pool:
        list(pool.map(lambda p: O.run_reuse_kernel(fn, p[0], unit, setup, p[1], p[2]), pieces))
    wall = time.perf_counter() - t
    return {"value": pairs / wall, "unit": "pair-evals/s", "cores": threads, "kind": kind,
            "sample": f"{sample} grid centers x {len(axes)} orientations in 64-center pieces on "
                      f"{threads} threads, {pairs} in-cutoff pairs, kernel wall time {wall:.2f} s; "
                      f"{cpu_model()}"}


def gpu_error_metrics(ref, cand):
    """error_metrics on device tensors (accessible: E < 2000 K and |E| < 1e4 K)."""
    import torch
    diff = (ref - cand).abs()
    acc = (ref < 2000.0) & (ref.abs() < 1e4)
    rel = diff / ref.abs().clamp(min=1.0)
    n_acc = int(acc.sum())
    return {"accessible_poses": n_acc,
            "accessible_max_abs_error_K": float(diff[acc].max()) if n_acc else None,
            "accessible_rms_error_K": float(diff[acc].pow(2).mean().sqrt()) if n_acc else None,
            "all_max_relative_error_floor1K": float(rel.max())}


def ragged_line(args):
    """100 frameworks x 2 guests, 60^3 x 64 fields each, one persistent launch."""
    import torch
    from string_vext_amd import autotune, synthetic
    from string_vext_amd.field import Batch, Plan
    from string_vext_amd.orientation import regular_pose_grid
    t0 = time.perf_counter()
    cases = synthetic.ragged_cases()
    tuned = autotune.tune(cases, size=SIZE)   # trials on the fields' own 60^3 grid
    log(f"autotune: {tuned['tile_centers']} blocks/CU {tuned['blocks_per_cu']} "
        f"({time.perf_counter() - t0:.1f} s)")
    _, axes = regular_pose_grid(SIZE, ORIENTATIONS)
    cost = [len(c["frame_frac"]) / np.linalg.det(c["cell"]) * (c["cutoff"] + c["bond_length"] / 2) ** 3
            for c in cases]
    order = np.argsort(cost)[::-1]                  # longest jobs first
    plans = [Plan.from_case(cases[i]) for i in order]
    jobs = [(p, p.grid(SIZE, axes, tile_centers=tuned["tile_centers"][autotune.density_class(cases[i])]))
            for p, i in zip(plans, order)]
    temps = [cases[i]["temperature"] for i in order]
    log(f"{len(jobs)} fields set up ({time.perf_counter() - t0:.1f} s)")
    batch = Batch(jobs, precision="mixed", temperatures=temps)
    if args.blocks_per_cu is not None:
        tuned["blocks_per_cu"] = args.blocks_per_cu
    batch.set_persistent(tuned["blocks_per_cu"])
    pairs = int(batch.census().sum())
    log(f"census: {pairs} in-cutoff pair-evals over {len(jobs)} fields")
    for _ in range(args.warmup):
        batch.run(sync=True)
    batch.sync()
    t = time.perf_counter()
    for _ in range(args.steps):
        batch.run(sync=False)
    batch.sync()
    step_s = (time.perf_counter() - t) / args.steps
    batch.run(time_kernels=True, sync=True)
    phases = batch.timings()
    batch.close()
    # per-field error of the mixed form against the FP64 form (chunks, metrics on the GPU)
    per_field = []
    if not args.no_check:
        chunk = 10
        for start in range(0, len(jobs), chunk):
            part = jobs[start:start + chunk]
            outs = {}
            for prec in ("fp64", "mixed"):
                tens = [torch.empty(ps.n_centers * ps.n_axes, dtype=torch.float64, device="cuda")
                        for _, ps in part]
                torch.cuda.synchronize()
                b = Batch(part, precision=prec, outputs=(), external={"energy": [x.data_ptr() for x in tens]})
                b.run(sync=True)
                b.close()
                outs[prec] = tens
            for j in range(len(part)):
                m = gpu_error_metrics(outs["fp64"][j], outs["mixed"][j])
                m["field"] = int(order[start + j])
                per_field.append(m)
            del outs
    acc = [m["accessible_max_abs_error_K"] for m in per_field if m["accessible_max_abs_error_K"] is not None]
    achieved = pairs * FLOP_PER_PAIR / (phases["field"] * 1e-3) / 1e12
    return {"metric": "pair-evals/s, ragged batch of 100 frameworks x 2 guests, 60^3x64 fields, mixed",
            "value": pairs / step_s, "unit": "pair-evals/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": step_s * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "ragged batch: 100 frameworks (200-4000 atoms log-uniform, "
                                   "10-60 A, 60-120 deg) x 2 guests, 60^3 x 64 each, one persistent launch",
                       "fields": len(jobs), "pair_evals": pairs, "autotune": tuned["tile_centers"],
                       "blocks_per_cu": tuned["blocks_per_cu"]},
            "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                         "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                         "kernel": "field_mixed_kernel (all fields, one launch)",
                         "kernel_ms": phases["field"], "flop_per_pair": FLOP_PER_PAIR},
            "phases_ms": phases,
            "error_vs_fp64": {"fields_checked": len(per_field),
                              "max_accessible_max_abs_error_K": max(acc) if acc else None,
                              "fields_with_accessible_poses": len(acc),
                              "max_relative_error_floor1K": max((m["all_max_relative_error_floor1K"] for m in per_field), default=None)},
            "per_field": per_field}


def poses_line(args):
    """8 full-size frameworks x 65,536 random poses: forward + backward through
    torch.autograd, mixed and FP64, checked against FP64 and central differences;
    then one timed forward + backward per framework for the site kernels' roofline."""
    import torch
    from string_vext_amd import synthetic
    from string_vext_amd.autograd import _grid_for, pose_energy
    from string_vext_amd.field import Plan
    rng = np.random.default_rng(20261105)
    n_poses = 65536
    results = {}
    data = []
    for k in range(8):
        case = synthetic.full_field_case(seed=synthetic.FULL_SEED + k)
        plan = Plan.from_case(case)
        frac = rng.random((n_poses, 3))
        u = rng.normal(size=(n_poses, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        c = torch.tensor(frac @ case["cell"], device="cuda")
        data.append((plan, c, torch.tensor(u, device="cuda")))
    # pose_energy's default route while autograd records: energy and gradient
    # from one site pass in the forward, which the backward only scales;
    # "mixed_two_pass" (fused=False) recomputes the pair terms in the backward
    modes = {"mixed": ("mixed", None), "fp64": ("fp64", None), "mixed_two_pass": ("mixed", False)}
    for name, (prec, fused) in modes.items():
        times, outs = [], []
        for rep in range(args.warmup + args.steps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            step_out = []
            for plan, c, u in data:
                cc = c.clone().requires_grad_(True)
                uu = u.clone().requires_grad_(True)
                e = pose_energy(plan, cc, uu, prec, fused=fused)
                e.sum().backward()
                step_out.append((e.detach(), cc.grad, uu.grad))
            torch.cuda.synchronize()
            if rep >= args.warmup:
                times.append(time.perf_counter() - t)
            outs = step_out
        results[name] = (float(np.mean(times)), outs)
    # site-kernel durations (HIP events on torch's stream) and the pair census
    kernels = {name: {"binning": 0.0, "forward": 0.0, "backward": 0.0, "combine": 0.0}
               for name in modes}
    pairs = 0
    for name, (prec, fused) in modes.items():
        for plan, c, u in data:
            grid = _grid_for(plan, len(c))
            grid.set_timing(True)
            cc = c.clone().requires_grad_(True)
            uu = u.clone().requires_grad_(True)
            e = pose_energy(plan, cc, uu, prec, fused=fused)
            fwd = grid.timings()
            if name == "mixed":
                pairs += grid.census()
            e.sum().backward()
            bwd = grid.timings() if fused is False else {"site_kernel": 0.0, "combine": 0.0}
            grid.set_timing(False)
            kt = kernels[name]
            kt["binning"] += fwd["binning"]
            kt["forward"] += fwd["site_kernel"]
            kt["backward"] += bwd["site_kernel"]
            kt["combine"] += fwd["combine"] + bwd["combine"]
    err_e, err_g = [], []
    for (e64, gc64, gu64), (emx, gcmx, gumx) in zip(results["fp64"][1], results["mixed"][1]):
        err_e.append(gpu_error_metrics(e64, emx))
        scale = torch.maximum(gc64.abs().amax(dim=1, keepdim=True), torch.ones_like(gc64[:, :1]))
        err_g.append(float(((gcmx - gc64).abs() / scale).max()))
    # central differences (h = 1e-4 A) on a sample of accessible poses, FP64
    plan, c, u = data[0]
    e0 = results["fp64"][1][0][0]
    idx = torch.nonzero(e0 < 2000).flatten()[:64]
    h = 1e-4
    fd_err = 0.0
    for k in range(3):
        step = torch.zeros(3, dtype=torch.float64, device="cuda")
        step[k] = h
        ep = pose_energy(plan, c[idx] + step, u[idx], "fp64")
        em = pose_energy(plan, c[idx] - step, u[idx], "fp64")
        fd = (ep - em) / (2 * h)
        an = results["fp64"][1][0][1][idx, k]
        fd_err = max(fd_err, float(((fd - an).abs() / an.abs().clamp(min=1.0)).max()))
    n = 8 * n_poses
    mk, tk = kernels["mixed"], kernels["mixed_two_pass"]
    fused_tf = pairs * FLOP_PER_GRAD_PAIR / (mk["forward"] * 1e-3) / 1e12
    fwd_tf = pairs * FLOP_PER_PAIR / (tk["forward"] * 1e-3) / 1e12
    bwd_tf = pairs * FLOP_PER_GRAD_PAIR / (tk["backward"] * 1e-3) / 1e12
    route_diff = 0.0
    for (e1, gc1, gu1), (e2, gc2, gu2) in zip(results["mixed"][1], results["mixed_two_pass"][1]):
        route_diff = max(route_diff, float((e1 - e2).abs().max()), float((gc1 - gc2).abs().max()),
                         float((gu1 - gu2).abs().max()))
    return {"metric": "poses/s, energy + dE/d(center, director) via torch.autograd, 8 x 65,536 poses",
            "value": n / results["mixed"][0], "unit": "poses/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": results["mixed"][0] * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic", "config": {"workload": "8 full-size frameworks (1200 atoms, 37x37x14 A) "
                                                         "x 65,536 uniform random poses, forward + backward "
                                                         "(pose_energy's default route: energy and "
                                                         "gradient in one site pass)",
                                            "pair_evals": pairs, "pairs_per_pose": pairs / n},
            "roofline": {"bound": "valu", "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                         "kernel": "site_kernel<mixed, energy + gradient>", "achieved": fused_tf,
                         "frac": fused_tf / PEAK_FP32_TFLOPS, "flop_per_pair": FLOP_PER_GRAD_PAIR,
                         "kernel_ms_per_framework": mk["forward"] / len(data),
                         "two_pass": {
                             "forward": {"kernel": "site_kernel<mixed, energy>", "achieved": fwd_tf,
                                         "frac": fwd_tf / PEAK_FP32_TFLOPS,
                                         "flop_per_pair": FLOP_PER_PAIR,
                                         "kernel_ms_per_framework": tk["forward"] / len(data)},
                             "backward": {"kernel": "site_kernel<mixed, gradient>", "achieved": bwd_tf,
                                          "frac": bwd_tf / PEAK_FP32_TFLOPS,
                                          "flop_per_pair": FLOP_PER_GRAD_PAIR,
                                          "kernel_ms_per_framework": tk["backward"] / len(data)}}},
            "phases_ms_per_framework": {prec: {k: v / len(data) for k, v in kt.items()}
                                        for prec, kt in kernels.items()},
            "two_pass": {"value": n / results["mixed_two_pass"][0],
                         "ms_per_step": results["mixed_two_pass"][0] * 1e3,
                         "max_abs_diff_vs_default_route": route_diff},
            "fp64": {"value": n / results["fp64"][0], "ms_per_step": results["fp64"][0] * 1e3},
            "speedup_mixed_vs_fp64": results["fp64"][0] / results["mixed"][0],
            "energy_error_vs_fp64": {
                "max_accessible_max_abs_error_K": max(m["accessible_max_abs_error_K"] or 0 for m in err_e),
                "max_relative_error_floor1K": max(m["all_max_relative_error_floor1K"] for m in err_e)},
            "gradient_error_vs_fp64_rel": max(err_g),
            "central_difference_rel_error_h1e-4": fd_err}


def _string_census(plans, path):
    """In-cutoff (site, image) pairs of one energy pass over the images of `path`."""
    import torch
    from string_vext_amd.field import PoseGrid
    from string_vext_amd.strings import director
    pairs = 0
    for plan in dict.fromkeys(plans):
        mine = np.array([p is plan for p in plans])
        flat = path[mine].reshape(-1, 6)
        c = torch.tensor(np.mod(flat[:, :3], 1) @ plan.cell, device="cuda")
        u = torch.tensor(director(flat[:, 4], flat[:, 5]), device="cuda")
        e = torch.empty(len(flat), dtype=torch.float64, device="cuda")
        grid = PoseGrid(plan, len(flat))
        grid.eval(len(flat), c.data_ptr(), u.data_ptr(), "fp64", energy=e.data_ptr())
        pairs += grid.census()
        grid.close()
    return pairs


def strings_line(args):
    """64 strings x 96 images, `iterations` iterations in a HIP graph, on the
    full-size frameworks and on a porous variant (400 atoms, same cell and
    types): FP64 and mixed runs, a second FP64 run from a start moved by 1e-15
    (the iteration's own sensitivity), the FP64 path re-evaluated in mixed
    precision, and one iteration of each precision from the FP64 run's state."""
    from string_vext_amd import synthetic
    from string_vext_amd.field import Plan, PoseList
    from string_vext_amd.strings import StringSystem, barrier, director, straight_strings
    iters = args.iterations
    report = {}
    for label, n_atoms in (("full", 1200), ("porous", 400)):
        plans, init, dirs = [], [], []
        for k in range(8):
            plan = Plan.from_case(synthetic.full_field_case(seed=synthetic.FULL_SEED + k, n_atoms=n_atoms))
            for j in range(8):
                d = 1 + (k * 8 + j) % 3
                plans.append(plan)
                dirs.append(d)
                init.append(straight_strings(plan, d, 1, 96, seed=1000 * k + j)[0])
        init = np.stack(init)
        dirs = np.array(dirs)
        moved = init.copy()
        moved[:, 1:-1, :3] += 1e-15
        out = {}
        for name, prec, start in (("fp64", "fp64", init), ("mixed", "mixed", init),
                                  ("fp64_moved", "fp64", moved)):
            system = StringSystem(plans, start, dirs, precision=prec)
            system.run(1, graph=True)                 # capture + first iteration
            t = time.perf_counter()
            system.run(iters - 1, graph=True)
            dt = time.perf_counter() - t
            out[name] = (dt / (iters - 1), system.state())
            system.close()
            log(f"strings {label} {name}: {dt / (iters - 1) * 1e3:.3f} ms/iteration")
        p64, e64, it64, c64 = out["fp64"][1]
        emx, cmx = out["mixed"][1][1], out["mixed"][1][3]
        emv = out["fp64_moved"][1][1]
        b64 = barrier(e64)
        same = np.empty_like(e64)
        for s_ in range(len(p64)):
            pl = PoseList(plans[s_], np.mod(p64[s_, :, :3], 1) @ plans[s_].cell,
                          director(p64[s_, :, 4], p64[s_, :, 5]))
            same[s_] = pl.energy_host("mixed")
            pl.close()
        shadow = {}
        for prec in ("fp64", "mixed"):
            system = StringSystem(plans, p64, dirs, precision=prec)
            system.run(1, graph=True)
            shadow[prec] = system.state()
            system.close()
        dstep = np.abs(shadow["mixed"][0] - shadow["fp64"][0])
        acc = e64.max(axis=1) < 2000.0
        pairs = _string_census(plans, p64)

        def spread(other):
            d = np.abs(barrier(other) - b64)[acc]
            return {"max_K": float(d.max()), "median_K": float(np.median(d))} if acc.any() else None

        ms64, msmx = out["fp64"][0] * 1e3, out["mixed"][0] * 1e3
        # whole-iteration rate: an energy+gradient pass and an energy pass over
        # every image, at the final state's pair census
        flop = pairs * (FLOP_PER_GRAD_PAIR + FLOP_PER_PAIR)
        report[label] = {
            "ms_per_iteration_fp64": ms64, "ms_per_iteration_mixed": msmx,
            "pair_evals_per_pass": pairs,
            "iteration_rate": {"fp64_TFLOP/s": flop / (ms64 * 1e-3) / 1e12,
                               "frac_of_fp64_peak": flop / (ms64 * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                               "mixed_TFLOP/s": flop / (msmx * 1e-3) / 1e12},
            "strings_with_accessible_paths": int(acc.sum()),
            "same_path_max_profile_diff_K": float(np.abs(same - e64)[acc].max()) if acc.any() else None,
            "same_path_max_barrier_diff_K": float(np.abs(barrier(same) - b64)[acc].max()) if acc.any() else None,
            "one_iteration_max_coord_diff": float(dstep[..., :3].max()),
            "one_iteration_max_angle_diff_rad": float(dstep[..., 3:].max()),
            "independent_barrier_spread_mixed_vs_fp64": spread(emx),
            "independent_barrier_spread_fp64_moved_1e-15_vs_fp64": spread(emv),
            "iterations_fp64": sorted(set(it64.tolist())), "converged_fp64": int(c64.sum()),
            "converged_mixed": int(cmx.sum()),
            "barriers_fp64_K": [round(float(v), 6) for v in b64[acc]],
            "barriers_mixed_K": [round(float(v), 6) for v in barrier(emx)[acc]]}
    # The FP64 form is the faster one here (FP64 issue runs at the FP32 rate and
    # the image sweep is short), so it is the reported value; the mixed form's
    # time and its errors are reported beside it.
    return {"metric": "string iterations/s, 64 strings x 96 images, hipGraph",
            "value": 1e3 / report["full"]["ms_per_iteration_fp64"], "unit": "iterations/s",
            "n_gpus": 1, "steps": iters, "warmup": 1,
            "ms_per_step": report["full"]["ms_per_iteration_fp64"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"64 strings x 96 images (8 frameworks x 8 strings, straight "
                                   f"fractional paths along a/b/c), {iters} iterations, one iteration "
                                   f"per HIP graph replay"},
            "frameworks": report}


def dropin_line(args):
    """The reference's own call sequence with this library as its --kernel: the
    full 60^3 x 64 field in 512-center batches, cKDTree CSR on the host, one
    evaluate_reused_neighbors call per batch (FP64, host buffers, tables sent
    with every call).  Compared with the batched FP64 field on the same poses."""
    from string_vext_amd import _lib, synthetic
    from string_vext_amd.field import Plan
    from string_vext_amd.native import NativeReusePotential
    from string_vext_amd.orientation import regular_pose_grid
    case = synthetic.full_field_case()
    centers, axes = regular_pose_grid(SIZE, ORIENTATIONS)
    pot = NativeReusePotential(case, _lib.LIB_PATH, batched=False)
    pot.evaluate(centers[:512], axes)                      # warm-up batch
    walls, queries, computes, out = [], [], [], None
    for _ in range(args.steps):
        pot.query_seconds = pot.compute_seconds = 0.0
        t = time.perf_counter()
        out = pot.evaluate(centers, axes)
        walls.append(time.perf_counter() - t)
        queries.append(pot.query_seconds)
        computes.append(pot.compute_seconds)
    plan = Plan.from_case(case)
    poses = plan.grid(SIZE, axes)
    pairs = plan.census(poses)
    ref = plan.field(poses, precision="fp64")["energy"]
    rel = float((np.abs(out - ref) / np.maximum(np.abs(ref), 1.0)).max())
    wall, query, compute = (float(np.mean(v)) for v in (walls, queries, computes))
    return {"metric": "pair-evals/s, 60^3 x 64 Vext through the reference's call sequence "
                      "(evaluate_reused_neighbors per 512-center batch, host cKDTree CSR)",
            "value": pairs / wall, "unit": "pair-evals/s", "n_gpus": 1, "steps": args.steps,
            "warmup": 1, "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "full-size single Vext field: 60^3 grid x 64 orientations, "
                                   "1200-atom triclinic framework, drop-in C ABI route",
                       "batches": (len(centers) + 511) // 512},
            "query_seconds": query, "call_seconds": compute,
            "calls_value": pairs / compute,
            "max_relative_vs_batched_fp64_floor1K": rel}


def e2e_line(args):
    """The `string-vext` CLI end to end (reference src/string_vext/vext.py:46-80):
    parse input.dat, the 60^3 x 64 field with its marginal, grid and site
    channels, the compressed NPZ and the provenance JSON with SHA256 digests,
    for the 1200-atom framework written as an input.dat.  The reference
    publishes 22.263 s for two gases (about 11 s per field) with its CPU kernel
    on a 1568-atom framework (tests/cof/validation/vext_endtoend_speed_2materials.json)."""
    import tempfile
    from string_vext_amd import synthetic, vext
    from string_vext_amd.input import parse_string_input, write_string_input
    from string_vext_amd.npz import write_npz
    case = synthetic.full_field_case()
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        src = Path(tmp) / "input.dat"
        write_string_input(case, src)
        # warm-up: device plan setup, kernels loaded
        vext.calculate_vext(parse_string_input(src), size=8, orientations=ORIENTATIONS)
        for prec in ("fp64", "mixed"):
            walls = []
            for rep in range(max(1, args.steps)):
                dst = Path(tmp) / f"vext_{prec}_{rep}.npz"
                t = time.perf_counter()
                # the CLI prints its provenance record; keep stdout for the one bench line
                with contextlib.redirect_stdout(sys.stderr):
                    vext.main(["--input", str(src), "--output", str(dst), "--size", str(SIZE),
                               "--orientations", str(ORIENTATIONS), "--precision", prec])
                walls.append(time.perf_counter() - t)
            t = time.perf_counter()
            fields = vext.calculate_vext(parse_string_input(src), SIZE, ORIENTATIONS, precision=prec)
            calc = time.perf_counter() - t
            t = time.perf_counter()
            with open(Path(tmp) / "probe.npz", "wb") as handle:
                write_npz(handle, fields)
            npz = time.perf_counter() - t
            t = time.perf_counter()
            with open(Path(tmp) / "probe_numpy.npz", "wb") as handle:
                np.savez_compressed(handle, **fields)
            npz_numpy = time.perf_counter() - t
            out[prec] = {"cli_s": float(np.mean(walls)), "calculate_vext_s": calc,
                         "npz_write_s": npz, "numpy_savez_compressed_s": npz_numpy,
                         "npz_bytes": (Path(tmp) / "probe.npz").stat().st_size,
                         "numpy_npz_bytes": (Path(tmp) / "probe_numpy.npz").stat().st_size}
    return {"metric": "seconds per string-vext call, 60^3 x 64 field, input.dat to NPZ + JSON",
            "value": out["mixed"]["cli_s"], "unit": "s", "n_gpus": 1, "steps": max(1, args.steps),
            "ms_per_step": out["mixed"]["cli_s"] * 1e3, "higher_is_better": False,
            "data": "synthetic", "config": {"workload": "string-vext CLI, 1200-atom framework "
                                                        "(37x37x14 A), --size 60 --orientations 64"},
            "by_precision": out,
            "reference_cpu_s_per_field": 22.263 / 2}


def compact(line, keep):
    """The headline line's summary of another workload's line."""
    out = {k: line[k] for k in ("metric", "value", "unit", "ms_per_step", "steps") if k in line}
    out.update({k: line[k] for k in keep if k in line})
    return out


def field_line(args):
    """The headline: the 60^3 x 64 field of the 1200-atom framework, mixed and FP64,
    with the reference CPU kernel on a sample of the same workload."""
    from string_vext_amd import synthetic
    from string_vext_amd.field import Plan
    from string_vext_amd.orientation import regular_pose_grid

    case = synthetic.full_field_case()
    centers, axes = regular_pose_grid(SIZE, ORIENTATIONS)
    from string_vext_amd import autotune
    if args.tile_centers > 0:
        tile_centers = args.tile_centers
    else:
        tuned = autotune.tune([case], size=SIZE, orientations=ORIENTATIONS, reps=5, persistent=False)
        tile_centers = tuned["tile_centers"][autotune.density_class(case)]
        log(f"autotune: tile_centers={tile_centers} trials={tuned['trials']}")
    plan = Plan.from_case(case)
    poses = plan.grid(SIZE, axes, tile_centers=tile_centers)
    pairs = plan.census(poses)
    log(f"census: {pairs} in-cutoff pair-evals per field ({pairs / (SIZE**3 * ORIENTATIONS):.1f} per pose)")

    step_s, field_ms, run_ms, t_mixed = time_field(plan, poses, "mixed", args.steps, args.warmup,
                                                   case["temperature"])
    log(f"mixed: {step_s * 1e3:.3f} ms/step, field kernel {field_ms:.3f} ms, phases {t_mixed}")
    f64_steps = max(1, args.fp64_steps)
    step64, field64, run64, t64 = time_field(plan, poses, "fp64", f64_steps, 1, case["temperature"])
    log(f"fp64: {step64 * 1e3:.3f} ms/step, field kernel {field64:.3f} ms")
    # the FP64 form with 1/d^2 from v_rcp_f64 + one Newton step (a stricter
    # denominator than the reference's IEEE division)
    step64r, field64r, _, _ = time_field(plan, poses, "fp64_rcp", f64_steps, 1, case["temperature"])
    log(f"fp64_rcp: {step64r * 1e3:.3f} ms/step, field kernel {field64r:.3f} ms")

    errors = porous = rcp_rel = None
    if not args.no_check:
        e64 = energies(plan, poses, "fp64")
        emx = energies(plan, poses, "mixed")
        errors = error_metrics(e64, emx)
        log(f"mixed vs fp64: {errors}")
        rcp_rel = float((np.abs(energies(plan, poses, "fp64_rcp") - e64)
                         / np.maximum(np.abs(e64), 1.0)).max())
        log(f"fp64_rcp vs fp64: max relative difference {rcp_rel:.3e} (1 K floor)")
        porous = porous_field_audit(tile_centers)
        log(f"porous field, mixed vs fp64: {porous}")

    achieved = pairs * FLOP_PER_PAIR / (field_ms * 1e-3) / 1e12
    # HBM bytes per launch from PMC counters, profiled at each tile size
    traffic = None
    prof = ROOT / "profiles" / "field_mixed_60x64_traffic.json"
    if prof.exists():
        rec = json.loads(prof.read_text()).get("by_tile", {}).get(str(tile_centers))
        traffic = rec and rec["hbm_bytes_per_launch"]

    cpu = cpu_all = small = None
    if args.cpu_sample > 0:
        log(f"cpu baseline on {args.cpu_sample} centers ...")
        cpu = cpu_baseline(case, centers, axes, args.cpu_sample)
        log(f"cpu: {cpu['value']:.3e} pair-evals/s")
        threads = max(1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        cpu_all = cpu_baseline_threads(case, centers, axes, min(len(centers), 65536), threads)
        log(f"cpu, {threads} threads: {cpu_all['value']:.3e} pair-evals/s")
        small = small_field_report(threads)
        log(f"20^3 field: {small}")

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
                   "pair_evals_per_field": pairs, "accumulate": "f64",
                   "tile_centers": tile_centers},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "traffic_tile_centers": tile_centers if traffic else None,
                     "kernel": "field_mixed_kernel", "kernel_ms": field_ms,
                     "flop_per_pair": FLOP_PER_PAIR},
        "cpu_baseline": cpu and {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")},
        "cpu_baseline_all_cores": cpu_all,
        "small_field": small,
        "fp64": {"value": pairs / step64, "ms_per_step": step64 * 1e3, "kernel_ms": field64,
                 "frac_of_fp64_peak": pairs * FLOP_PER_PAIR / (field64 * 1e-3) / 1e12 / PEAK_FP64_TFLOPS},
        "speedup_mixed_vs_fp64": step64 / step_s,
        "fp64_rcp": {"value": pairs / step64r, "ms_per_step": step64r * 1e3, "kernel_ms": field64r,
                     "frac_of_fp64_peak": pairs * FLOP_PER_PAIR / (field64r * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                     "max_relative_vs_fp64_floor1K": rcp_rel,
                     "speedup_mixed_vs_fp64_rcp": step64r / step_s},
        "error_vs_fp64": errors,
        "porous_field_error_vs_fp64": porous,
        "phases_ms": t_mixed,
    }
    return line


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--workload", choices=("field", "ragged", "poses", "strings", "dropin", "e2e"),
                    default="field",
                    help="field (default): the 60^3 x 64 headline, with summaries of the ragged, "
                         "poses, strings and e2e workloads; the others print one workload's full "
                         "line (dropin: the reference's call sequence through the C ABI; e2e: the "
                         "string-vext CLI from input.dat to NPZ)")
    ap.add_argument("--iterations", type=int, default=2000, help="strings workload iterations")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fp64-steps", type=int, default=5)
    ap.add_argument("--cpu-sample", type=int, default=49152,
                    help="grid centers in the single-core CPU baseline sample, about 11 s of "
                         "reference-kernel time (0 skips the CPU baselines)")
    ap.add_argument("--no-check", action="store_true", help="skip the mixed-vs-FP64 error pass")
    ap.add_argument("--headline-only", action="store_true",
                    help="field workload: skip the ragged, poses and strings summaries")
    ap.add_argument("--tile-centers", type=int, default=0,
                    help="field workload: centers per tile (0 = autotuned)")
    ap.add_argument("--blocks-per-cu", type=int, default=None,
                    help="ragged workload: persistent grid size (0 = plain grid; default autotuned)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run a plain elementwise copy of MIB MiB (FP32 in, FP32 out) so that "
                         "PMC byte counters can be scaled against a kernel of known traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): every pose is independent, "
              "so N GPUs would each evaluate their own slab of centers with no communication; "
              "this benchmark measures one GPU's share at the full field shape.", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1:
        sys.exit("--steps must be >= 1")
    if args.workload != "field":
        line = {"ragged": ragged_line, "poses": poses_line, "strings": strings_line,
                "dropin": dropin_line, "e2e": e2e_line}[args.workload](args)
        print(json.dumps(line), flush=True)
        return
    line = field_line(args)
    if not args.headline_only:
        sub = argparse.Namespace(**vars(args))
        sub.steps, sub.warmup = 3, 1
        line["ragged"] = compact(ragged_line(sub), ("config", "roofline", "error_vs_fp64"))
        log(f"ragged: {line['ragged']['value']:.3e} pair-evals/s")
        sub.steps, sub.warmup = 5, 2
        line["poses"] = compact(poses_line(sub), (
            "config", "roofline", "phases_ms_per_framework", "two_pass", "fp64", "energy_error_vs_fp64",
            "gradient_error_vs_fp64_rel", "central_difference_rel_error_h1e-4"))
        log(f"poses: {line['poses']['value']:.3e} poses/s")
        sl = strings_line(args)
        line["strings"] = compact(sl, ("config",))
        line["strings"]["frameworks"] = {
            k: {kk: vv for kk, vv in v.items() if not kk.startswith("barriers_")}
            for k, v in sl["frameworks"].items()}
        log(f"strings: {line['strings']['value']:.1f} iterations/s")
        sub.steps = 1
        line["e2e"] = compact(e2e_line(sub), ("config", "by_precision", "reference_cpu_s_per_field"))
        log(f"e2e: {line['e2e']['value']:.2f} s per string-vext call (mixed)")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
