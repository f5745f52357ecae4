"""Learned block-Jacobi correction (include/maplab_ba_learned.h) on the device
against the CPU restatement (oracle/gnn_oracle.py): training records, the
network's forward pass and gradient, the preconditioned PCG, and its use
inside the LM loop (same LM log, fewer PCG iterations, block-Jacobi
fallback)."""
import numpy as np
import pytest

from maplab_amd import abi, learned, problem
from maplab_amd.solver import DeviceProblem
from oracle import gnn_oracle as go

pytestmark = pytest.mark.gpu


def opts(**kw):
    o = abi.default_options()
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def walk(kf=40, lm=6000, seed=3):
    return problem.workload_map(2, seed=seed, num_keyframes=kf, num_landmarks=lm)


def const_model(theta0, norm=None):
    p = np.zeros(learned.num_params(), np.float32)
    p[-2:] = theta0
    if norm is None:
        norm = np.concatenate([np.zeros(9), np.ones(9), np.zeros(3), np.ones(3)]).astype(np.float32)
    return learned.Model(p, norm)


def test_records_match_oracle_traces(oracle):
    m = walk()
    o = opts()
    radius = 1e4
    print(f"records: N {m.num_observations}", flush=True)
    with DeviceProblem(m) as dev:
        r = learned.rcs_records(dev, o, radius)
    S, b, vb = oracle.reduced_camera_system(m, o, radius)
    ref = go.traces(S)
    n = S.shape[0] // 6
    assert r.nodes == n
    np.testing.assert_array_equal(r.qc[:, 4], ref[:, 4])
    scale = np.abs(ref[:, :4]).max(axis=0)
    assert np.all(np.abs(r.qc[:, :4] - ref[:, :4]) <= 2e-4 * scale + 1e-4 * np.abs(ref[:, :4])), \
        np.abs(r.qc[:, :4] - ref[:, :4]).max(axis=0) / scale
    # Structure and the features taken from S itself.
    _, Sb, nz, _, Sh = go.normalized(S)
    deg = nz.sum(1) - 1
    lower = np.array([sum(nz[i, :i]) for i in range(n)])
    np.testing.assert_allclose(r.node_feat[:, 7], np.log1p(deg), rtol=1e-6)
    np.testing.assert_allclose(r.node_feat[:, 8], np.log1p(lower), rtol=1e-6)
    np.testing.assert_allclose(r.node_feat[:, 6], np.log([np.trace(Sb[i, i]) for i in range(n)]), rtol=1e-5)
    dst = np.repeat(np.arange(n), np.diff(r.row_ptr))
    edge = r.col != dst
    assert np.all(nz[dst, r.col])
    np.testing.assert_allclose(r.edge_feat[edge, 1], np.linalg.norm(Sh[dst[edge], r.col[edge]], axis=(1, 2)),
                               rtol=1e-4)
    assert np.all(r.edge_feat[edge, 0] > 0)
    np.testing.assert_array_equal(r.col[r.rev], dst)


def _dataset(maps, radius=1e5):
    recs = []
    for m in maps:
        with DeviceProblem(m) as dev:
            recs.append(learned.rcs_records(dev, abi.default_options(), radius))
    norm = learned.normalization(recs)
    return learned.Dataset(recs, norm), recs, norm


def test_network_forward_and_gradient_match_oracle():
    data, _, _ = _dataset([walk(30, 4000, 5), problem.workload_map(1, seed=9, num_keyframes=24, num_landmarks=900)])
    rng = np.random.default_rng(0)
    params = learned.init_params(1) + (rng.normal(size=learned.num_params()) * 0.02).astype(np.float32)
    print(f"network parity: {data.nodes} rows, {data.slots} slots", flush=True)
    theta, loss, grad = learned.evaluate(data, params)
    theta2, loss2, grad2 = learned.evaluate(data, params)
    np.testing.assert_array_equal(theta, theta2)        # fixed-order reductions: bit for bit
    np.testing.assert_array_equal(grad, grad2)
    g = go.Graph(data.row_ptr, data.col, data.rev, data.node_feat, data.edge_feat)
    ref_loss, ref_grad, ref_theta = go.loss_and_grad(params.astype(np.float64), g, data.qc.astype(np.float64))
    # BF16 activations on both sides; the residual differences are FP32
    # accumulation order and the rare BF16 rounding it flips.
    assert np.abs(theta - ref_theta).max() <= 2e-2 * np.abs(ref_theta).max()
    assert abs(loss - ref_loss) <= 1e-3 * abs(ref_loss)
    err = np.linalg.norm(grad - ref_grad) / np.linalg.norm(ref_grad)
    print(f"gradient rel err {err:.2e}", flush=True)
    assert err < 3e-2
    for name, (o, k) in go.LAYOUT.items():
        a, b = grad[o:o + k], ref_grad[o:o + k]
        assert np.linalg.norm(a - b) <= 0.1 * np.linalg.norm(b) + 1e-7, name


def test_learned_pcg_matches_oracle_iterations(oracle):
    m = walk(50, 8000, 7)
    o = opts()
    radius = 1e6
    theta0 = (0.5, -4.0)
    with DeviceProblem(m) as dev:
        const_model(theta0).install(dev)
        it_l = learned.rcs_pcg_iterations(dev, o, radius, abi.PRECOND_LEARNED, 1e-6)
        it_j = learned.rcs_pcg_iterations(dev, o, radius, abi.PRECOND_BLOCK_JACOBI, 1e-6)
    S, b, _ = oracle.reduced_camera_system(m, o, radius)
    n = S.shape[0] // 6
    G = go.kmatrix(S, np.tile(theta0, (n, 1)))
    Li = go.kmatrix(S, np.zeros((n, 2)))
    ref_l = go.pcg_iterations(S, b, lambda r: G.T @ (G @ r))
    ref_j = go.pcg_iterations(S, b, lambda r: Li.T @ (Li @ r))
    print(f"PCG to 1e-6: learned {it_l} (oracle {ref_l}), block-Jacobi {it_j} (oracle {ref_j})", flush=True)
    assert abs(it_j - ref_j) <= 1 and abs(it_l - ref_l) <= 2
    assert it_l < it_j


def test_learned_deflated_pcg_matches_oracle_iterations(oracle):
    # G'G inside the hat deflation, and the hat-deflated block-Jacobi PCG,
    # against the numpy deflated PCG over the same coarse space.
    m = walk(50, 8000, 7)
    o = opts()
    radius = 1e6
    theta0 = (0.5, -4.0)
    with DeviceProblem(m) as dev:
        const_model(theta0).install(dev)
        it_ld = learned.rcs_pcg_iterations(dev, o, radius, abi.PRECOND_LEARNED_DEFLATED, 1e-6)
        it_d = learned.rcs_pcg_iterations(dev, o, radius, abi.PRECOND_DEFLATED, 1e-6)
        it_l = learned.rcs_pcg_iterations(dev, o, radius, abi.PRECOND_LEARNED, 1e-6)
    S, b, _ = oracle.reduced_camera_system(m, o, radius)
    n = S.shape[0] // 6
    W = go.hat_basis(n, go.default_hat_nodes(n))
    assert W.shape[1] > 0
    G = go.kmatrix(S, np.tile(theta0, (n, 1)))
    Li = go.kmatrix(S, np.zeros((n, 2)))
    ref_ld, _ = go.deflated_pcg(S, b, lambda r: G.T @ (G @ r), W)
    ref_d, _ = go.deflated_pcg(S, b, lambda r: Li.T @ (Li @ r), W)
    print(f"PCG to 1e-6: learned+deflation {it_ld} (oracle {ref_ld}), deflated {it_d} (oracle {ref_d}), "
          f"learned {it_l}", flush=True)
    assert abs(it_d - ref_d) <= 1 and abs(it_ld - ref_ld) <= 2
    assert 0 < it_ld < it_d


def test_lm_with_learned_preconditioner_keeps_the_log():
    m = walk(60, 12000, 2)
    base = dict(max_num_iterations=8, pcg_relative_tolerance=1e-10, pcg_max_iterations=2000)
    out = {}
    pcs = (abi.PRECOND_DEFLATED, abi.PRECOND_BLOCK_JACOBI, abi.PRECOND_LEARNED, abi.PRECOND_LEARNED_DEFLATED)
    for pc in pcs:
        with DeviceProblem(m) as dev:
            const_model((0.5, -4.0)).install(dev)
            sm, rows = dev.solve(opts(pcg_preconditioner=pc, **base))
            out[pc] = (sm, [r.cost for r in rows], sum(r.linear_solver_iterations for r in rows),
                       learned.stats(dev))
    ref = out[abi.PRECOND_DEFLATED][1]
    for pc in pcs[1:]:
        np.testing.assert_allclose(out[pc][1], ref, rtol=1e-8)
    print(f"PCG iterations over the LM run: learned {out[abi.PRECOND_LEARNED][2]}, "
          f"block-Jacobi {out[abi.PRECOND_BLOCK_JACOBI][2]}, learned+deflation "
          f"{out[abi.PRECOND_LEARNED_DEFLATED][2]}, deflated {out[abi.PRECOND_DEFLATED][2]}", flush=True)
    for pc in (abi.PRECOND_LEARNED, abi.PRECOND_LEARNED_DEFLATED):
        st = out[pc][3]
        assert st.solves >= out[pc][0].num_iterations - 1 and st.fallbacks == 0
    assert out[abi.PRECOND_LEARNED][2] < out[abi.PRECOND_BLOCK_JACOBI][2]
    assert out[abi.PRECOND_LEARNED_DEFLATED][2] < out[abi.PRECOND_LEARNED][2]


def test_every_preconditioner_with_fp32_jacobians_keeps_the_log():
    # The four solvers with FP64 and with FP32 Jacobian storage: each run
    # follows the deflated solver's log of the same precision to 1e-7 (all
    # stop at |r| <= 1e-10 |b| with different errors in the step; the last,
    # rejected candidate's cost moved 1.1e-8). FP32 against FP64: the large
    # first step (cost 5.2e6 to 1.7e6) differs by 3.5e-6 and the logs meet
    # again (2.2e-8 at the last iteration), so the log is held to 1e-5 and
    # the final cost to the 1e-6 bar.
    m = walk(60, 12000, 2)
    base = dict(max_num_iterations=6, pcg_relative_tolerance=1e-10, pcg_max_iterations=2000)
    pcs = (abi.PRECOND_DEFLATED, abi.PRECOND_BLOCK_JACOBI, abi.PRECOND_LEARNED, abi.PRECOND_LEARNED_DEFLATED)
    logs = {}
    with DeviceProblem(m) as dev:
        const_model((0.5, -4.0)).install(dev)
        for f32 in (0, 1):
            for pc in pcs:
                dev.upload_state(m.vertices, m.landmarks)
                _, rows = dev.solve(opts(pcg_preconditioner=pc, jacobian_fp32=f32, **base))
                logs[(pc, f32)] = [r.cost for r in rows]
        assert learned.stats(dev).fallbacks == 0
    for f32 in (0, 1):
        for pc in pcs[1:]:
            np.testing.assert_allclose(logs[(pc, f32)], logs[(abi.PRECOND_DEFLATED, f32)], rtol=1e-7)
    f32, f64 = logs[(abi.PRECOND_DEFLATED, 1)], logs[(abi.PRECOND_DEFLATED, 0)]
    np.testing.assert_allclose(f32, f64, rtol=1e-5)
    assert abs(f32[-1] - f64[-1]) <= 1e-6 * f64[-1]


def test_learned_deflated_without_coarse_columns_is_the_learned_pcg():
    # 19 variable vertices: too few for the hat space (no coarse columns), so
    # the deflated PCG with G'G is the learned PCG with other reduction orders.
    m = walk(20, 2000, 6)
    assert go.default_hat_nodes(m.num_vertices - 1) == 0
    with DeviceProblem(m) as dev:
        const_model((0.5, -4.0)).install(dev)
        its = {pc: [learned.rcs_pcg_iterations(dev, opts(), r, pc, 1e-8) for r in (1e4, 1e8)]
               for pc in (abi.PRECOND_LEARNED, abi.PRECOND_LEARNED_DEFLATED)}
        logs = {}
        for pc in (abi.PRECOND_LEARNED, abi.PRECOND_LEARNED_DEFLATED):
            dev.upload_state(m.vertices, m.landmarks)
            _, rows = dev.solve(opts(pcg_preconditioner=pc, max_num_iterations=5, pcg_relative_tolerance=1e-10))
            logs[pc] = [r.cost for r in rows]
    print(f"iterations to 1e-8: {its}", flush=True)
    for a, b in zip(its[abi.PRECOND_LEARNED], its[abi.PRECOND_LEARNED_DEFLATED]):
        assert a > 0 and abs(a - b) <= 1
    np.testing.assert_allclose(logs[abi.PRECOND_LEARNED_DEFLATED], logs[abi.PRECOND_LEARNED], rtol=1e-9)


@pytest.mark.parametrize("pc", [abi.PRECOND_LEARNED, abi.PRECOND_LEARNED_DEFLATED])
def test_block_jacobi_fallback_when_the_learned_solve_does_not_converge(pc):
    m = walk(40, 4000, 4)
    o = opts(max_num_iterations=3, pcg_max_iterations=3, pcg_preconditioner=pc)
    with DeviceProblem(m) as dev:
        const_model((0.5, -4.0)).install(dev)
        dev.solve(o)
        st = learned.stats(dev)
    assert st.solves > 0 and st.fallbacks == st.solves and st.fallback_iterations > 0


@pytest.mark.parametrize("pc", [abi.PRECOND_LEARNED, abi.PRECOND_LEARNED_DEFLATED])
def test_learned_preconditioner_needs_a_model(pc):
    with DeviceProblem(walk(20, 2000, 6)) as dev:
        with pytest.raises(RuntimeError, match="model"):
            dev.solve(opts(pcg_preconditioner=pc, max_num_iterations=1))


def test_learned_deflated_with_summed_group_partials_keeps_the_log():
    # 1000 keyframes: 24 row groups, more than one LDS chunk of partials, so
    # the last work-group of k_defl_partials sums them once per iteration;
    # the solve repeats bit for bit and follows the deflated solver's LM log
    # (to 1e-7: both stop at |r| <= 1e-10 |b| on this long, ill-conditioned
    # walk, with different errors in the step; measured 3e-8 on the last).
    m = problem.workload_map(2, seed=3, num_keyframes=1000, num_landmarks=20000)
    assert (m.num_vertices + 41) // 42 > 22
    base = dict(max_num_iterations=4, pcg_relative_tolerance=1e-10, pcg_max_iterations=3000)
    logs = []
    with DeviceProblem(m) as dev:
        const_model((0.5, -4.0)).install(dev)
        init_v, init_l = m.vertices.copy(), m.landmarks.copy()
        for pc in (abi.PRECOND_LEARNED_DEFLATED, abi.PRECOND_LEARNED_DEFLATED, abi.PRECOND_DEFLATED):
            dev.upload_state(init_v, init_l)
            _, rows = dev.solve(opts(pcg_preconditioner=pc, **base))
            logs.append([(r.cost, r.linear_solver_iterations) for r in rows])
        st = learned.stats(dev)
    print(f"learned+deflation {[i for _, i in logs[0]]}, deflated {[i for _, i in logs[2]]}", flush=True)
    assert logs[0] == logs[1]
    np.testing.assert_allclose([c for c, _ in logs[0]], [c for c, _ in logs[2]], rtol=1e-7)
    assert st.fallbacks == 0


def test_training_on_the_device_lowers_the_loss():
    maps = [problem.workload_map(1, seed=40 + k, num_keyframes=12 + 6 * k, num_landmarks=800 + 300 * k)
            for k in range(4)]
    data, _, norm = _dataset(maps)
    p0 = learned.init_params(3)
    params, log = learned.train(data, p0, 60, 3e-3)
    print(f"training loss {log[0]:.5f} -> {log[-1]:.5f}", flush=True)
    assert np.isfinite(log).all() and log[-1] < log[0]
    _, loss, _ = learned.evaluate(data, params, want_grad=False)
    assert loss < log[0]


def test_committed_model_halves_the_block_jacobi_iterations():
    # The trained network (maplab_amd/models/covis_precond.json) against the
    # method without it, plain block-Jacobi, on a 300-keyframe random walk
    # (PCG to |r|/|b| <= 1e-6 at two trust-region radii), and the LM log of
    # a solve with it against block-Jacobi's.
    model = learned.Model.load()
    m = walk(300, 60000, 11)
    o = opts()
    with DeviceProblem(m) as dev:
        model.install(dev)
        lj = [learned.rcs_pcg_iterations(dev, o, r, abi.PRECOND_LEARNED, 1e-6) for r in (1e6, 1e8)]
        bj = [learned.rcs_pcg_iterations(dev, o, r, abi.PRECOND_BLOCK_JACOBI, 1e-6) for r in (1e6, 1e8)]
        print(f"committed model: learned {lj} block-Jacobi {bj}", flush=True)
        assert min(lj) > 0 and sum(lj) <= 0.6 * sum(bj)
        logs = {}
        for pc in (abi.PRECOND_BLOCK_JACOBI, abi.PRECOND_LEARNED):
            dev.upload_state(m.vertices, m.landmarks)
            _, rows = dev.solve(opts(max_num_iterations=5, pcg_relative_tolerance=1e-10, pcg_max_iterations=3000,
                                     pcg_preconditioner=pc))
            logs[pc] = [r.cost for r in rows]
        np.testing.assert_allclose(logs[abi.PRECOND_LEARNED], logs[abi.PRECOND_BLOCK_JACOBI], rtol=1e-8)
        assert learned.stats(dev).fallbacks == 0
