#!/usr/bin/env python3
"""Benchmark of the visual bundle-adjustment hot path on one MI355X.

A step is one Levenberg-Marquardt iteration (accepted or rejected, Ceres
counting) of maplab's visual BA on the 500-keyframe random-walk map (200k
landmarks, ~2.08M observations, pinhole-radtan 752x480, maplab Huber loss,
2% outliers), solved on the device through the C ABI. The map is generated
on the host and resident in HBM before the timed region starts.

Prints one JSON line with the LM iterations/s (value), observation
Jacobians/s, per-kernel device times and the roofline of the dominant kernel,
and the CPU oracle (FP64 C++ port of the same path) timed on a bounded
sample of the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 vector


def kernel_table(sm, kt, n_pcg_total):
    """Per-kernel device time in the timed region and its roofline terms.
    Algorithmic bytes per launch (FP64 observation records of 20 doubles):
      linearize: N * 196 B (36 B read + 160 B record written) + L * 176 B
      cost:      N * 36 B + L * 32 B
      spmv:      nnzb * 292 B (6x6 FP64 block + column index) + Pvar * 104 B
      schur:     flops = pairs * 288 (6x3.3x6 product + W' formation)."""
    from maplab_amd import abi
    N, L, P = kt.num_observations, kt.num_landmarks, kt.num_variable_vertices
    nnzb, pairs = kt.schur_blocks, kt.schur_pairs
    rows = {}

    def add(name, kid, bytes_per_launch=None, flops_per_launch=None):
        launches = sm.kernel_launches[kid]
        if launches == 0:
            return
        ms = sm.kernel_ms[kid] / launches
        row = {"launches": launches, "total_ms": round(sm.kernel_ms[kid], 4), "avg_ms": round(ms, 5)}
        if bytes_per_launch is not None:
            row.update(bound="hbm", bytes=int(bytes_per_launch), achieved_gbs=round(bytes_per_launch / ms / 1e6, 1),
                       frac=round(bytes_per_launch / ms / 1e6 / HBM_PEAK_GBS, 4))
        if flops_per_launch is not None:
            row.update(bound="fp64", flops=int(flops_per_launch),
                       achieved_tflops=round(flops_per_launch / ms / 1e9, 3),
                       frac=round(flops_per_launch / ms / 1e9 / FP64_PEAK_TFLOPS, 4))
        rows[name] = row

    add("k_linearize", abi.KERNEL_LINEARIZE, N * 196 + L * 176)
    add("k_linearize_cost", abi.KERNEL_COST, N * 36 + L * 32)
    # Upper block triangle only: about half of the incidence pairs.
    add("k_schur_tasks", abi.KERNEL_SCHUR_ROWS, flops_per_launch=pairs * 144)
    add("k_back_substitute", abi.KERNEL_BACK_SUBSTITUTE, N * 160 + L * 216)
    add("k_pose_rows", abi.KERNEL_POSE_ROWS)
    add("k_landmark_prep", abi.KERNEL_LANDMARK_PREP, L * (144 + 24 + 48))
    launches = sm.kernel_launches[abi.KERNEL_PCG]
    if launches and n_pcg_total > 0:
        # The PCG solve of each LM step: two launches per PCG iteration
        # (S p with the direction update, one-block vector update). Per
        # iteration it streams S once and ~14 FP64 vectors of the system.
        it_ms = sm.kernel_ms[abi.KERNEL_PCG] / n_pcg_total
        it_bytes = nnzb * 292 + P * 6 * 8 * 14
        rows["pcg_solve"] = {
            "launches": launches, "total_ms": round(sm.kernel_ms[abi.KERNEL_PCG], 4),
            "avg_ms": round(sm.kernel_ms[abi.KERNEL_PCG] / launches, 5), "pcg_iterations": int(n_pcg_total),
            "per_iteration_ms": round(it_ms, 5), "bytes_per_iteration": int(it_bytes),
            "achieved_gbs": round(it_bytes / it_ms / 1e6, 1),
            "note": "events around the whole solve: both kernels, launch gaps and the host convergence polls"}
    if kt.pcg_spmv_ms > 0 and n_pcg_total > 0:
        spmv_bytes = nnzb * 292 + P * 6 * 8 * 4
        rows["k_pcg_spmv"] = {
            "launches": int(n_pcg_total), "avg_ms": round(kt.pcg_spmv_ms, 5),
            "total_ms": round(kt.pcg_spmv_ms * n_pcg_total, 4), "bound": "hbm", "bytes": int(spmv_bytes),
            "achieved_gbs": round(spmv_bytes / kt.pcg_spmv_ms / 1e6, 1),
            "frac": round(spmv_bytes / kt.pcg_spmv_ms / 1e6 / HBM_PEAK_GBS, 4),
            "timing": "avg from mlba_time_kernels events on the library stream, x PCG iterations of the timed region"}
        rows["k_pcg_update"] = {"launches": int(n_pcg_total), "avg_ms": round(kt.pcg_vector_ms, 5),
                                "total_ms": round(kt.pcg_vector_ms * n_pcg_total, 4)}
    return rows


def cpu_baseline(vmap, opts, iters, threads):
    from oracle import binding as oracle
    import ctypes
    o = type(opts)()
    ctypes.pointer(o)[0] = opts
    o.max_num_iterations = iters
    t0 = time.perf_counter()
    sm, rows, _, _, _ = oracle.solve(vmap, o, linear_solver=oracle.SCHUR_DIRECT, threads=threads)
    wall = time.perf_counter() - t0
    n_it = max(0, sm.num_iterations - 1)
    return sm, rows, n_it, wall, sm.lm_loop_seconds


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="LM iterations in the timed region")
    ap.add_argument("--warmup", type=int, default=2, help="untimed LM iterations before")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--keyframes", type=int, default=500)
    ap.add_argument("--landmarks", type=int, default=200000)
    ap.add_argument("--equidistant", action="store_true", help="equidistant distortion variant")
    ap.add_argument("--cpu-iters", type=int, default=2, help="LM iterations of the CPU oracle sample")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, nproc)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--repeats", type=int, default=20, help="launches per kernel in mlba_time_kernels")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding of the path is described in DESIGN.md, not run)", file=sys.stderr)
        return 2

    from maplab_amd import abi, problem
    from maplab_amd.solver import DeviceProblem

    kw = dict(num_keyframes=args.keyframes, num_landmarks=args.landmarks)
    if args.equidistant:
        kw.update(distortion_type=2, distortion=(-0.01, 0.02, -0.005, 0.001))
    t0 = time.perf_counter()
    vmap = problem.workload_map(2, seed=args.seed, **kw)
    gen_s = time.perf_counter() - t0
    print(f"bench: map generated in {gen_s:.2f}s: {vmap.num_vertices} keyframes, {vmap.num_landmarks} landmarks, "
          f"{vmap.num_observations} observations", file=sys.stderr, flush=True)

    opts = abi.default_options()
    dev = DeviceProblem(vmap)
    init_v, init_l = vmap.vertices.copy(), vmap.landmarks.copy()
    # Warmup (untimed), then restore the initial state.
    if args.warmup > 0:
        opts.max_num_iterations = args.warmup
        dev.solve(opts)
        dev.upload_state(init_v, init_l)
    kt = dev.time_kernels(opts, repeats=args.repeats)
    dev.upload_state(init_v, init_l)
    # Timed region: exactly `steps` LM iterations (fewer only if Ceres'
    # termination criteria stop the loop).
    opts.max_num_iterations = args.steps
    t0 = time.perf_counter()
    sm, rows = dev.solve(opts)  # synchronous: returns after the device finished
    wall = time.perf_counter() - t0
    n_iters = sm.num_iterations - 1
    its_per_s = n_iters / wall
    n_pcg = sm.total_linear_solver_iterations
    table = kernel_table(sm, kt, n_pcg)
    lin = table.get("k_linearize", {})
    obs_j_per_s = vmap.num_observations / (lin["avg_ms"] / 1e3) if lin else None
    # Dominant single kernel by device time in the timed region.
    single = {k: v for k, v in table.items() if "bound" in v and "total_ms" in v}
    dom_name = max(single, key=lambda k: single[k]["total_ms"])
    dom = single[dom_name]
    if dom["bound"] == "hbm":
        roof = {"kernel": dom_name, "bound": "hbm", "achieved": dom["achieved_gbs"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": dom["frac"], "traffic": None}
    else:
        roof = {"kernel": dom_name, "bound": "fp64", "achieved": dom["achieved_tflops"], "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": dom["frac"], "traffic": None}

    result = {
        "metric": "BA LM iters/s + obs Jacobians/s at 1e-6 cost rel err; PCG iters learned vs Jacobi",
        "value": round(its_per_s, 3),
        "unit": "LM iterations/s",
        "n_gpus": 1,
        "steps": n_iters,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / max(1, n_iters), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random-walk map generator)",
        "config": {
            "workload": (f"visual BA, random walk: {vmap.num_vertices} keyframes, {vmap.num_landmarks} landmarks, "
                         f"{vmap.num_observations} observations, pinhole-"
                         f"{'equidistant' if args.equidistant else 'radtan'} 752x480, maplab Huber loss, "
                         f"2% outliers, first vertex + one landmark fixed"),
            "keyframes": vmap.num_vertices, "landmarks": vmap.num_landmarks,
            "observations": vmap.num_observations, "lm_iterations": n_iters, "seed": args.seed,
            "linear_solver": "landmark Schur + block-Jacobi PCG (FP64, rel tol %g)" % opts.pcg_relative_tolerance,
        },
        "obs_jacobians_per_s": round(obs_j_per_s, 1) if obs_j_per_s else None,
        "pcg_iterations_block_jacobi": int(n_pcg),
        "lm_accepted": int(sm.num_successful_steps), "lm_rejected": int(sm.num_unsuccessful_steps),
        "initial_cost": sm.initial_cost, "final_cost": sm.final_cost,
        "kernels": table,
        "roofline": roof,
    }

    if not args.no_cpu_baseline:
        nproc = os.cpu_count() or 1
        threads = args.cpu_threads or min(16, nproc)
        vm_cpu = vmap.copy().normalized()
        vm_cpu.vertices, vm_cpu.landmarks = init_v.copy(), init_l.copy()
        csm, crows, cn, cwall, cloop = cpu_baseline(vm_cpu, opts, args.cpu_iters, threads)
        # GPU on the same sample for the cost parity after the same iterations.
        dev.upload_state(init_v, init_l)
        opts.max_num_iterations = args.cpu_iters
        gsm, _ = dev.solve(opts)
        rel = abs(gsm.final_cost - csm.final_cost) / abs(csm.final_cost)
        result["cpu_baseline"] = {
            "value": round(cn / cloop, 4) if cloop > 0 else None,
            "unit": "LM iterations/s",
            "cores": threads,
            "kind": "port",
            "sample": (f"FP64 C++ oracle (Schur + dense Cholesky of the reduced camera system), "
                       f"{cn} LM iterations of the same map from the same start, {threads} threads of "
                       f"{nproc} host cores, {cloop:.1f}s LM loop"),
        }
        result["cost_rel_err_vs_oracle"] = {"iterations": cn, "gpu": gsm.final_cost, "oracle": csm.final_cost,
                                            "rel_err": rel}
    dev.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
