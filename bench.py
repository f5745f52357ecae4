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
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E


def kernel_table(sm, kt, n_pcg_total):
    """Per-kernel device time in the timed region and its roofline terms.
    Algorithmic HBM bytes per launch (FP64; observation record = r, Jl, Jo =
    20 doubles; incidence = an observation coupling a variable vertex, or
    the anchor of a landmark with a variable anchor):
      linearize:          N * 196 (36 read + 160 record written) + L * 176
      cost:               N * 36 + L * 32
      schur_factors:      n_inc * 352 (160 record read, 144 W + 48 r0 written)
                          (per linearization)
      schur_blocks:       n_g * 192 (W or A_l, r0 of every factor block, read
                          once) + L * 120 (Z, g; V_LM of anchors) + staged
                          slots * 8 (slot words) + tile pairs * 2 (pair list)
                          + batches * 530 (batch word, pair offsets)
                          + segments * (1040 + 2 * 86016) (segment word,
                          thread parts, partials written and read)
                          + upper blocks * 584 (S_ij and S_ji written, index)
                          + P * 48 (rhs)
                          (Q is formed in LDS: the former per-iteration
                          incidence pass has no launches)
      pcg_spmv:           nnzb * 292 (6x6 block + column index) + P * 200 (z, p
                          read; q, p written; p'q) + P * 6 * gw * 8 (row of SW)
      pcg_update:         P * 344 (x, r read and written, p, q read, z
                          written, p'q) + P * 288 (Minv) + P * 6 * gw * 8 (SW)
                          + groups * gw * ne * 8 (E^-1 window rows)
                          + groups * (3 + ne) * 8 (partials written)
    with gw the coarse columns a row group couples to, ne the coarse columns
    and groups = ceil(P / 42).
    with n_g the (landmark, row) factor blocks. The Schur block kernel also
    does tile pairs * 216 FP64 flops (3x6' . 3x6), below its byte time at
    MI355X's FP64 peak."""
    from maplab_amd import abi
    N, L, P = kt.num_observations, kt.num_landmarks, kt.num_variable_vertices
    nnzb, pairs, n_inc, nub = kt.schur_blocks, kt.schur_pairs, kt.incidences, kt.schur_upper_blocks
    rows = {}

    def row_of(launches, total_ms, bytes_per_launch=None):
        ms = total_ms / launches
        row = {"launches": int(launches), "total_ms": round(total_ms, 4), "avg_ms": round(ms, 5)}
        if bytes_per_launch is not None:
            row.update(bound="hbm", bytes=int(bytes_per_launch), achieved_gbs=round(bytes_per_launch / ms / 1e6, 1),
                       frac=round(bytes_per_launch / ms / 1e6 / HBM_PEAK_GBS, 4))
        return row

    def add(name, kid, bytes_per_launch=None):
        launches = sm.kernel_launches[kid]
        if launches:
            rows[name] = row_of(launches, sm.kernel_ms[kid], bytes_per_launch)

    add("k_linearize", abi.KERNEL_LINEARIZE, N * 196 + L * 176)
    add("k_linearize_cost", abi.KERNEL_COST, N * 36 + L * 32)
    add("k_schur_factors", abi.KERNEL_SCHUR_FACTORS, n_inc * 352)
    tp = kt.schur_tile_pairs
    add("k_schur_blocks", abi.KERNEL_SCHUR_BLOCKS,
        kt.schur_factor_blocks * 192 + L * 120 + kt.schur_staged_blocks * 8 + tp * 2 + kt.schur_batches * 530
        + kt.schur_segments * (1040 + 2 * 86016) + nub * 584 + P * 48)
    if "k_schur_blocks" in rows:
        r = rows["k_schur_blocks"]
        r["flops"] = int(tp * 216)
        r["achieved_tflops"] = round(tp * 216 / r["avg_ms"] / 1e9, 3)
        r["staged_factor_blocks"] = int(kt.schur_staged_blocks)
    add("k_back_substitute", abi.KERNEL_BACK_SUBSTITUTE, N * 160 + L * 216)
    add("k_pose_rows", abi.KERNEL_POSE_ROWS)
    add("k_landmark_prep", abi.KERNEL_LANDMARK_PREP, L * (144 + 24 + 72))
    launches = sm.kernel_launches[abi.KERNEL_PCG]
    if launches and n_pcg_total > 0:
        it_ms = sm.kernel_ms[abi.KERNEL_PCG] / n_pcg_total
        rows["pcg_solve"] = {
            "launches": int(launches), "total_ms": round(sm.kernel_ms[abi.KERNEL_PCG], 4),
            "avg_ms": round(sm.kernel_ms[abi.KERNEL_PCG] / launches, 5), "pcg_iterations": int(n_pcg_total),
            "per_iteration_ms": round(it_ms, 5),
            "note": "events around each whole solve: coarse-space set-up, both kernels of every iteration, "
                    "launch gaps and the host convergence polls"}
    if kt.pcg_spmv_ms > 0 and n_pcg_total > 0:
        timing = ("mlba_time_kernels: events around a batch of back-to-back launches on the library stream "
                  "(launch gaps included), per launch, x PCG iterations")
        gw, ne = kt.pcg_window_columns, kt.pcg_coarse_columns
        groups = (P + 41) // 42
        spmv_b = nnzb * 292 + P * 200 + P * 6 * gw * 8
        upd_b = P * 344 + P * 288 + P * 6 * gw * 8 + groups * gw * ne * 8 + groups * (3 + ne) * 8
        for name, ms, b in (("k_pcg_spmv", kt.pcg_spmv_ms, spmv_b), ("k_pcg_update", kt.pcg_update_ms, upd_b)):
            rows[name] = row_of(n_pcg_total, ms * n_pcg_total, b)
            rows[name]["timing"] = timing
    return rows


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "lm_loop_random_walk_500kf_hbm_traffic.md")
# Bench kernel row -> the kernel whose rocprofv3 FETCH_SIZE / WRITE_SIZE row
# gives its HBM traffic (the Schur block build is dominated by its tile pass).
TRAFFIC_KERNEL = {"k_schur_blocks": "mlba::k_schur_tiles", "k_pcg_spmv": "mlba::k_pcg_spmv<false>",
                  "k_pcg_update": "mlba::k_pcg_update",
                  "k_schur_factors": "mlba::k_schur_factors<false>", "k_linearize": "mlba::k_linearize<true, false>",
                  "k_back_substitute": "mlba::k_back_substitute<false>"}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on the default workload, from the
    committed rocprofv3 counter summary (calibrated FETCH_SIZE + WRITE_SIZE);
    None when the summary has no row for it."""
    name = TRAFFIC_KERNEL.get(kernel)
    if name is None or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 7 and cells[0] == f"`{name}`":
            return int(float(cells[4]) * 1e6)
    return None


def multimap(args):
    """Aggregate LM iterations/s and observations/s of many independent maps,
    solved one after the other and concurrently (one JSON line)."""
    import numpy as np
    from maplab_amd import abi, problem
    from maplab_amd.solver import DeviceProblem
    rng = np.random.default_rng(args.seed)
    maps = []
    for i in range(args.multimap):
        kf = int(round(np.exp(rng.uniform(np.log(50), np.log(800)))))
        lm = int(round(np.exp(rng.uniform(np.log(10000), np.log(150000)))))
        kw = {} if i % 2 == 0 else dict(distortion_type=2, distortion=(-0.01, 0.02, -0.005, 0.001))
        maps.append(problem.workload_map(2, seed=args.seed + 1 + i, num_keyframes=kf, num_landmarks=lm,
                                         detection_probability=0.5, **kw))
    n_obs = sum(m.num_observations for m in maps)
    print(f"bench: {len(maps)} maps, {sum(m.num_vertices for m in maps)} keyframes, {n_obs} observations",
          file=sys.stderr, flush=True)
    opts = abi.default_options()
    opts.max_num_iterations = args.steps
    devs = [DeviceProblem(m) for m in maps]
    init = [(m.vertices.copy(), m.landmarks.copy()) for m in maps]
    if args.warmup > 0:
        w = abi.default_options()
        w.max_num_iterations = args.warmup
        DeviceProblem.solve_many(devs, w, threads=args.multimap_threads)
    for d, (v, l) in zip(devs, init):
        d.upload_state(v, l)
    t0 = time.perf_counter()
    seq = [d.solve(opts)[0] for d in devs]
    t_seq = time.perf_counter() - t0
    for d, (v, l) in zip(devs, init):
        d.upload_state(v, l)
    t0 = time.perf_counter()
    con = [sm for sm, _ in DeviceProblem.solve_many(devs, opts, threads=args.multimap_threads)]
    t_con = time.perf_counter() - t0
    iters = sum(sm.num_iterations - 1 for sm in con)
    same = all(a.final_cost == b.final_cost for a, b in zip(seq, con))
    # The same concurrent solve after each map's Schur block schedule is
    # autotuned (set-up, untimed), against the fixed default schedule above.
    tuned = None
    if not args.no_autotune:
        for d, (v, l) in zip(devs, init):
            d.upload_state(v, l)
            d.autotune(opts, repeats=3)
            d.upload_state(v, l)
        t0 = time.perf_counter()
        con_t = [sm for sm, _ in DeviceProblem.solve_many(devs, opts, threads=args.multimap_threads)]
        t_tuned = time.perf_counter() - t0
        tuned = {"wall_s": round(t_tuned, 4),
                 "lm_iterations_per_s": round(sum(sm.num_iterations - 1 for sm in con_t) / t_tuned, 2),
                 "same_results_as_fixed": all(a.final_cost == b.final_cost for a, b in zip(con, con_t))}
    for d in devs:
        d.close()
    print(json.dumps({
        "metric": "aggregate LM iterations/s over independent maps solved concurrently",
        "value": round(iters / t_con, 2), "unit": "LM iterations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "higher_is_better": True, "dtype": "f64", "data": "synthetic (seeded random-walk maps)",
        "config": {"workload": f"{len(maps)} random-walk maps, 50-800 keyframes, 10k-150k landmarks "
                               f"(log-uniform), {n_obs} observations, radtan and equidistant",
                   "threads": args.multimap_threads},
        "sequential": {"wall_s": round(t_seq, 4), "lm_iterations_per_s": round(iters / t_seq, 2)},
        "concurrent": {"wall_s": round(t_con, 4), "lm_iterations_per_s": round(iters / t_con, 2),
                       "observation_lm_iterations_per_s": round(
                           sum(m.num_observations * (sm.num_iterations - 1) for m, sm in zip(maps, con)) / t_con,
                           1)},
        "speedup": round(t_seq / t_con, 3), "same_results": bool(same),
        "concurrent_autotuned_schedule": tuned}), flush=True)
    return 0


def cpu_model():
    """Host CPU model name (first 'model name' of /proc/cpuinfo) with its
    family / model / stepping numbers (some hosts report only a generic
    name)."""
    fields = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" not in line:
                if fields:
                    break
                continue
            k, v = (x.strip() for x in line.split(":", 1))
            if k in ("model name", "cpu family", "model", "stepping") and k not in fields:
                fields[k] = v
    except OSError:
        pass
    name = fields.get("model name", "unknown")
    ids = [f"{k} {fields[k]}" for k in ("cpu family", "model", "stepping") if k in fields]
    return f"{name} ({', '.join(ids)})" if ids else name


def learned_model():
    """The committed preconditioner model, or None when it is absent."""
    from maplab_amd import learned
    return learned.Model.load() if os.path.exists(learned.DEFAULT_MODEL) else None


def learned_eval_2000(args, model):
    """PCG iterations to |r|/|b| <= 1e-6 on the reduced camera systems of the
    2000-keyframe map (1M landmarks) at three trust-region radii, with the
    learned preconditioner, plain block-Jacobi and the deflated solver."""
    from maplab_amd import abi, learned, problem
    from maplab_amd.solver import DeviceProblem
    t0 = time.perf_counter()
    m = problem.workload_map(3, seed=args.seed)
    print(f"bench: 2000-keyframe map for the preconditioner comparison ({m.num_observations} observations, "
          f"{time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
    o = abi.default_options()
    rows = {}
    with DeviceProblem(m) as dev:
        model.install(dev)
        for radius in (1e4, 1e6, 1e8):
            rows[f"{radius:g}"] = {
                name: learned.rcs_pcg_iterations(dev, o, radius, pc, 1e-6, 20000)
                for name, pc in (("learned", abi.PRECOND_LEARNED), ("block_jacobi", abi.PRECOND_BLOCK_JACOBI),
                                 ("deflated", abi.PRECOND_DEFLATED),
                                 ("learned_deflated", abi.PRECOND_LEARNED_DEFLATED))}
    tot = {k: sum(r[k] for r in rows.values()) for k in ("learned", "block_jacobi", "deflated", "learned_deflated")}
    return {"workload": f"random walk, 2000 keyframes, {m.num_landmarks} landmarks, {m.num_observations} observations",
            "tolerance": 1e-6, "per_radius": rows, "total": tot,
            "learned_vs_block_jacobi": round(tot["block_jacobi"] / max(1, tot["learned"]), 2),
            "learned_vs_deflated": round(tot["deflated"] / max(1, tot["learned"]), 2),
            "learned_deflated_vs_block_jacobi": round(tot["block_jacobi"] / max(1, tot["learned_deflated"]), 2),
            "learned_deflated_vs_deflated": round(tot["deflated"] / max(1, tot["learned_deflated"]), 2)}


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
    ap.add_argument("--cpu-iters", type=int, default=12,
                    help="LM iterations of the CPU oracle sample (about 1 s each on 16 host threads)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0: the reference's policy max(1, hw/2) with hw = min(16, nproc), and hw beside it")
    ap.add_argument("--no-learned-2000", action="store_true",
                    help="skip the preconditioner comparison on the 2000-keyframe map")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--repeats", type=int, default=20, help="launches per kernel in mlba_time_kernels")
    ap.add_argument("--multimap", type=int, default=0,
                    help="instead of the headline line: this many independent random-walk maps (sizes "
                         "log-uniform in 50-800 keyframes and 10k-150k landmarks) solved one after the other "
                         "and concurrently (mlba_solve_many)")
    ap.add_argument("--multimap-threads", type=int, default=8)
    ap.add_argument("--no-autotune", action="store_true", help="keep the default Schur block schedule")
    ap.add_argument("--pcg-partials", type=int, default=0, choices=(0, 1, 2),
                    help="PCG per-group partials summed by every SpMV work-group (1), once per update (2), "
                         "or by group count (0)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(how the path would shard over several GPUs is documented as a design and not run)", file=sys.stderr)
        return 2

    import numpy as np
    from maplab_amd import abi, problem
    from maplab_amd.solver import DeviceProblem

    if args.multimap > 0:
        return multimap(args)

    kw = dict(num_keyframes=args.keyframes, num_landmarks=args.landmarks)
    if args.equidistant:
        kw.update(distortion_type=2, distortion=(-0.01, 0.02, -0.005, 0.001))
    t0 = time.perf_counter()
    vmap = problem.workload_map(2, seed=args.seed, **kw)
    gen_s = time.perf_counter() - t0
    print(f"bench: map generated in {gen_s:.2f}s: {vmap.num_vertices} keyframes, {vmap.num_landmarks} landmarks, "
          f"{vmap.num_observations} observations", file=sys.stderr, flush=True)

    opts = abi.default_options()
    t0 = time.perf_counter()
    dev = DeviceProblem(vmap)
    create_s = time.perf_counter() - t0  # host topology build, Schur work decomposition and upload (not timed below)
    init_v, init_l = vmap.vertices.copy(), vmap.landmarks.copy()
    # Warmup (untimed), then restore the initial state.
    if args.warmup > 0:
        opts.max_num_iterations = args.warmup
        dev.solve(opts)
        dev.upload_state(init_v, init_l)
    sched = None
    if not args.no_autotune:
        # Launch schedule of the Schur block kernel (set-up, untimed; changes
        # no result bit).
        sched = dev.autotune(opts, repeats=3)
        print(f"bench: autotuned schedule (band {sched.item_band_rows}, waves {sched.item_waves}, "
              f"xcd {sched.xcd_mapping}): {sched.tuned_ms:.4f} ms vs default {sched.default_ms:.4f} ms",
              file=sys.stderr, flush=True)
    if args.pcg_partials:
        sc = dev.schedule()
        dev.set_schedule(sc.item_band_rows, sc.item_waves, sc.xcd_mapping, args.pcg_partials)
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
    # The same LM iterations with plain block-Jacobi PCG (no coarse space, no
    # recycled vectors), untimed: PCG iterations of the two preconditioners.
    dev.upload_state(init_v, init_l)
    jo = abi.default_options()
    ctypes.pointer(jo)[0] = opts
    jo.pcg_coarse_nodes, jo.pcg_ritz_vectors = -1, -1
    jsm, jrows = dev.solve(jo)
    n_pcg_jacobi = jsm.total_linear_solver_iterations
    same_log = [r.step_is_successful for r in rows] == [r.step_is_successful for r in jrows]
    # The same comparison at the metric's PCG accuracy, |r_k| / |b| <= 1e-6
    # (untimed; both solves from the same start).
    loose = {}
    for name, (nodes, ritz) in (("deflated", (0, 0)), ("block_jacobi", (-1, -1))):
        dev.upload_state(init_v, init_l)
        lo = abi.default_options()
        ctypes.pointer(lo)[0] = opts
        lo.pcg_relative_tolerance = 1e-6
        lo.pcg_coarse_nodes, lo.pcg_ritz_vectors = nodes, ritz
        lsm, _ = dev.solve(lo)
        loose[name] = int(lsm.total_linear_solver_iterations)
    # The learned preconditioner on the same LM iterations at 1e-6 (timed), its
    # LM log against block-Jacobi's at the same accuracy.
    model = learned_model()
    learned_lm = None
    if model is not None:
        from maplab_amd import learned
        dev.upload_state(init_v, init_l)
        model.install(dev)
        dev.lib.mlba_reset_preconditioner_stats(dev.handle)
        ro = abi.default_options()
        ctypes.pointer(ro)[0] = opts
        ro.pcg_relative_tolerance = 1e-6
        ro.pcg_preconditioner = abi.PRECOND_LEARNED
        t0 = time.perf_counter()
        rsm, rrows = dev.solve(ro)
        r_wall = time.perf_counter() - t0
        st = learned.stats(dev)
        dev.upload_state(init_v, init_l)
        bo = abi.default_options()
        ctypes.pointer(bo)[0] = opts
        bo.pcg_relative_tolerance = 1e-6
        bo.pcg_preconditioner = abi.PRECOND_BLOCK_JACOBI
        bsm, brows = dev.solve(bo)
        # G'G inside the hat deflation, same iterations and tolerance.
        dev.upload_state(init_v, init_l)
        dev.lib.mlba_reset_preconditioner_stats(dev.handle)
        co = abi.default_options()
        ctypes.pointer(co)[0] = opts
        co.pcg_relative_tolerance = 1e-6
        co.pcg_preconditioner = abi.PRECOND_LEARNED_DEFLATED
        t0 = time.perf_counter()
        csm, crows = dev.solve(co)
        c_wall = time.perf_counter() - t0
        cst = learned.stats(dev)
        loose["learned"] = int(rsm.total_linear_solver_iterations)
        loose["learned_deflated"] = int(csm.total_linear_solver_iterations)
        learned_lm = {
            "pcg_iterations": int(rsm.total_linear_solver_iterations),
            "block_jacobi_pcg_iterations": int(bsm.total_linear_solver_iterations),
            "ratio": round(bsm.total_linear_solver_iterations / max(1, rsm.total_linear_solver_iterations), 2),
            "lm_iterations_per_s": round((rsm.num_iterations - 1) / r_wall, 3),
            "same_lm_log_as_block_jacobi": [r.step_is_successful for r in rrows] ==
                                           [r.step_is_successful for r in brows],
            "final_cost_rel_to_block_jacobi": abs(rsm.final_cost - bsm.final_cost) / abs(bsm.final_cost),
            "solves": int(st.solves), "block_jacobi_fallbacks": int(st.fallbacks), "tolerance": 1e-6,
            "model": model.meta.get("generator"),
            "with_hat_deflation": {
                "pcg_iterations": int(csm.total_linear_solver_iterations),
                "lm_iterations_per_s": round((csm.num_iterations - 1) / c_wall, 3),
                "same_lm_log_as_block_jacobi": [r.step_is_successful for r in crows] ==
                                               [r.step_is_successful for r in brows],
                "final_cost_rel_to_block_jacobi": abs(csm.final_cost - bsm.final_cost) / abs(bsm.final_cost),
                "solves": int(cst.solves), "fallbacks": int(cst.fallbacks)}}
    # maplab's default driver (--ba_use_visual_outlier_rejection_solver):
    # chunks of 3 iterations with device outlier rejection after each,
    # timed over the same iteration count.
    dev.upload_state(init_v, init_l)
    oo = abi.default_options()
    ctypes.pointer(oo)[0] = opts
    t0 = time.perf_counter()
    osm, orows = dev.solve(oo, outlier_rejection=True)
    o_wall = time.perf_counter() - t0
    o_iters = sum(1 for r in orows if r.iteration > 0)
    outlier_run = {"lm_iterations_per_s": round(o_iters / o_wall, 3), "lm_iterations": o_iters,
                   "ms_per_step": round(1e3 * o_wall / max(1, o_iters), 3),
                   "removed_landmarks": int(osm.num_removed_landmarks), "final_cost": osm.final_cost,
                   "chunk_iterations": int(oo.reject_outliers_every_n_iters)}
    dev.set_landmark_active(np.ones(vmap.num_landmarks, dtype=np.uint8))
    # The same LM iterations with FP32 Jacobian storage (timed the same way).
    dev.upload_state(init_v, init_l)
    fo = abi.default_options()
    ctypes.pointer(fo)[0] = opts
    fo.jacobian_fp32 = 1
    t0 = time.perf_counter()
    fsm, frows = dev.solve(fo)
    f_wall = time.perf_counter() - t0
    f_iters = fsm.num_iterations - 1
    table = kernel_table(sm, kt, n_pcg)
    lin = table.get("k_linearize", {})
    obs_j_per_s = vmap.num_observations / (lin["avg_ms"] / 1e3) if lin else None
    # Dominant single kernel by device time in the timed region.
    single = {k: v for k, v in table.items() if "bound" in v and "total_ms" in v}
    dom_name = max(single, key=lambda k: single[k]["total_ms"])
    dom = single[dom_name]
    default_workload = (args.keyframes, args.landmarks, args.seed, args.equidistant) == (500, 200000, 1, False)
    roof = {"kernel": dom_name, "bound": "hbm", "achieved": dom["achieved_gbs"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": dom["frac"],
            "traffic": measured_traffic(dom_name) if default_workload else None,
            "algorithmic_bytes": dom["bytes"]}

    result = {
        "metric": "BA LM iters/s + obs Jacobians/s at 1e-6 cost rel err; PCG iters learned vs Jacobi",
        "value": round(its_per_s, 3),
        "unit": "LM iterations/s",
        "n_gpus": 1,
        "steps": n_iters,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / max(1, n_iters), 3),
        "ms_per_accepted_step": round(1e3 * wall / max(1, sm.num_successful_steps), 3),
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
            "linear_solver": ("landmark Schur + deflated block-Jacobi PCG (FP64, rel tol %g)"
                              % opts.pcg_relative_tolerance),
        },
        "obs_jacobians_per_s": round(obs_j_per_s, 1) if obs_j_per_s else None,
        "pcg_iterations": int(n_pcg),
        "pcg_preconditioner": ("block-Jacobi + deflation by piecewise-linear hat functions and up to 16 Ritz "
                               "vectors recycled from the previous solves"),
        "pcg_iterations_vs_block_jacobi": {
            "deflated": int(n_pcg), "block_jacobi": int(n_pcg_jacobi),
            "ratio": round(n_pcg_jacobi / max(1, n_pcg), 2), "same_lm_log": bool(same_log),
            "tolerance": opts.pcg_relative_tolerance},
        "pcg_iterations_vs_block_jacobi_at_1e-6": {
            **loose, "ratio": round(loose["block_jacobi"] / max(1, loose["deflated"]), 2), "tolerance": 1e-6},
        "learned_preconditioner": learned_lm,
        "outlier_rejection_solver": outlier_run,
        "fp32_jacobians": {
            "lm_iterations_per_s": round(f_iters / f_wall, 3), "ms_per_step": round(1e3 * f_wall / max(1, f_iters), 3),
            "final_cost": fsm.final_cost,
            "final_cost_rel_to_fp64": abs(fsm.final_cost - sm.final_cost) / abs(sm.final_cost),
            "note": "mlba_options.jacobian_fp32: Jacobian records in FP32, all products and sums in FP64"},
        "lm_accepted": int(sm.num_successful_steps), "lm_rejected": int(sm.num_unsuccessful_steps),
        "initial_cost": sm.initial_cost, "final_cost": sm.final_cost,
        "kernels": table,
        "roofline": roof,
        "pcg_partials": args.pcg_partials,
    }
    if sched is not None:
        result["schur_schedule"] = {
            "item_band_rows": sched.item_band_rows, "item_waves": sched.item_waves,
            "xcd_mapping": sched.xcd_mapping, "schedules_tried": sched.schedules_tried,
            "k_schur_tiles_default_ms": round(sched.default_ms, 4), "k_schur_tiles_tuned_ms": round(sched.tuned_ms, 4)}

    if model is not None and not args.no_learned_2000:
        result["learned_preconditioner_2000kf"] = learned_eval_2000(args, model)
    copy_ms = ctypes.c_double(0.0)
    if dev.lib.mlba_measure_copy(1 << 30, 20, ctypes.byref(copy_ms)) == 0 and copy_ms.value > 0:
        result["hbm_copy_gbs_measured"] = round(2 * (1 << 30) / copy_ms.value / 1e6, 1)
    result["create_s"] = round(create_s, 3)

    if not args.no_cpu_baseline:
        nproc = os.cpu_count() or 1
        hw = min(16, nproc)
        policy = args.cpu_threads or max(1, hw // 2)   # solver-options.h:37-38: max(1, hw / 2)
        vm_cpu = vmap.copy().normalized()
        vm_cpu.vertices, vm_cpu.landmarks = init_v.copy(), init_l.copy()
        csm, crows, cn, cwall, cloop = cpu_baseline(vm_cpu, opts, args.cpu_iters, policy)
        all_rate = None
        if not args.cpu_threads and hw != policy:
            _, _, an, _, aloop = cpu_baseline(vm_cpu, opts, args.cpu_iters, hw)
            all_rate = round(an / aloop, 4) if aloop > 0 else None
        # GPU on the same sample for the cost parity after the same iterations.
        dev.upload_state(init_v, init_l)
        opts.max_num_iterations = args.cpu_iters
        gsm, _ = dev.solve(opts)
        rel = abs(gsm.final_cost - csm.final_cost) / abs(csm.final_cost)
        result["cpu_baseline"] = {
            "value": round(cn / cloop, 4) if cloop > 0 else None,
            "unit": "LM iterations/s",
            "cores": policy,
            "kind": "port",
            "sample": (f"FP64 C++ oracle (Schur + dense Cholesky of the reduced camera system, the direct "
                       f"SPARSE_SCHUR-equivalent solve), {cn} LM iterations of the same map from the same start, "
                       f"{policy} threads (maplab's max(1, hw/2) policy with hw = {hw} of {nproc} visible cores), "
                       f"{cloop:.1f}s LM loop; CPU: {cpu_model()}"),
            "value_at_hw_threads": all_rate,
            "hw_threads": hw,
            "cpu_model": cpu_model(),
        }
        result["cost_rel_err_vs_oracle"] = {"iterations": cn, "gpu": gsm.final_cost, "oracle": csm.final_cost,
                                            "rel_err": rel}
    dev.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
