"""Pins the oracle's Levenberg-Marquardt loop.

* Reference known answers (algorithms/ceres-error-terms/test/
  test_visual_term_test.cc:142-226): ground-truth start stops after one
  iteration with zero cost; a landmark seen by two fixed cameras converges to
  the origin. The reference fixture's camera, pinhole with FOV-fisheye
  distortion w = 0 (test_visual_term_test.cc:26-59), is used as it is.
* Three linear solvers of the oracle (Schur + Cholesky, dense normal
  equations, Schur + PCG) agree on the LM trajectory.
* A separate numpy restatement of the Ceres LM rules (Jacobi scaling, LM
  diagonal, gain ratio, radius update) on dense normal equations reproduces
  the oracle's iteration log on a tiny map.
"""
import numpy as np
import pytest

from cases import tiny_map
from maplab_amd import abi
from maplab_amd.problem import VisualMap


def options(**kw):
    o = abi.Options()
    defaults = dict(max_num_iterations=30, use_jacobi_scaling=1, max_solver_time_in_seconds=1e6,
                    function_tolerance=1e-12, gradient_tolerance=1e-10, parameter_tolerance=1e-8,
                    initial_trust_region_radius=1e5, max_trust_region_radius=1e20,
                    min_trust_region_radius=1e-32, min_relative_decrease=1e-3, min_lm_diagonal=1e-6,
                    max_lm_diagonal=1e32, max_num_consecutive_invalid_steps=5,
                    loss_type=abi.LOSS_MAPLAB_HUBER, loss_scale_px=1.0, linear_solver_type=0,
                    pcg_max_iterations=500, pcg_relative_tolerance=1e-12, reject_outliers_every_n_iters=3,
                    use_reprojection_error=0, max_reprojection_error_px=50.0)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(o, k, v)
    return o


def fixture_map(landmark, cameras, keypoints, sigma=0.7, intr=(1.0, 1.0, 320.0, 240.0)):
    """Landmark stored in an identity base vertex, observed (local mission
    type) by constant camera vertices -- the reference test fixture."""
    P = 1 + len(cameras)
    verts = np.array([[0, 0, 0, 1.0, 0, 0, 0]] + [list(q) + list(p) for q, p in cameras])
    n = len(cameras)
    return VisualMap(
        vertices=verts, baseframes=np.array([[0, 0, 0, 1.0, 0, 0, 0]]), landmarks=np.array([landmark], float),
        camera_q_CI__C_p_CI=np.array([[0, 0, 0, 1.0, 0, 0, 0]]), camera_intrinsics=np.array([intr]),
        camera_distortion=np.zeros((1, 4)), camera_distortion_type=np.array([3], np.int32),
        camera_resolution=np.array([[640, 480]], np.int32), vertex_mission=np.zeros(P, np.int32),
        landmark_store_vertex=np.array([0], np.int32), obs_vertex=np.arange(1, P, dtype=np.int32),
        obs_landmark=np.zeros(n, np.int32), obs_camera=np.zeros(n, np.int32), obs_keypoint=np.array(keypoints, float),
        obs_sigma=np.full(n, sigma), vertex_constant=np.ones(P, np.uint8), landmark_constant=np.zeros(1, np.uint8),
    ).normalized()


def test_ground_truth_start_stops_at_iteration_zero(oracle):
    # test_visual_term_test.cc:142-174: base pose and IMU pose (0,0,-1),
    # landmark (0,0,1), intrinsics (1,1,0,0), keypoint (0,0).
    m = VisualMap(
        vertices=np.array([[0, 0, 0, 1.0, 0, 0, -1], [0, 0, 0, 1.0, 0, 0, -1]]),
        baseframes=np.array([[0, 0, 0, 1.0, 0, 0, 0]]), landmarks=np.array([[0, 0, 1.0]]),
        camera_q_CI__C_p_CI=np.array([[0, 0, 0, 1.0, 0, 0, 0]]), camera_intrinsics=np.array([[1.0, 1, 0, 0]]),
        camera_distortion=np.zeros((1, 4)), camera_distortion_type=np.array([3], np.int32),
        camera_resolution=np.array([[640, 480]], np.int32), vertex_mission=np.zeros(2, np.int32),
        landmark_store_vertex=np.array([0], np.int32), obs_vertex=np.array([1], np.int32),
        obs_landmark=np.zeros(1, np.int32), obs_camera=np.zeros(1, np.int32), obs_keypoint=np.zeros((1, 2)),
        obs_sigma=np.array([0.7]), vertex_constant=np.zeros(2, np.uint8), landmark_constant=np.zeros(1, np.uint8),
    ).normalized()
    sm, rows, v, l, _ = oracle.solve(m, options(max_num_iterations=10000, function_tolerance=1e-50,
                                                gradient_tolerance=1e-50, parameter_tolerance=1e-50))
    assert sm.final_cost == 0.0
    assert sm.num_iterations == 1
    np.testing.assert_allclose(v[:, 4:], [[0, 0, -1], [0, 0, -1]], atol=1e-15)
    np.testing.assert_allclose(l, [[0, 0, 1]], atol=1e-15)


def test_one_point_two_cameras_noisy(oracle):
    # test_visual_term_test.cc:178-226: landmark (0.2,-0.1,0.05) -> origin.
    s = np.sqrt(2) / 2
    m = fixture_map([0.2, -0.1, 0.05], [((0, 0, 0, 1.0), (0, 0, -1)), ((0, -s, 0, s), (1, 0, 0))],
                    [[320.0, 240.0], [320.0, 240.0]])
    sm, rows, v, l, _ = oracle.solve(m, options(max_num_iterations=10000, function_tolerance=1e-50,
                                                gradient_tolerance=1e-50, parameter_tolerance=1e-50))
    assert sm.final_cost < 1e-15
    np.testing.assert_allclose(l[0], 0.0, atol=1e-10)


@pytest.mark.parametrize("global_missions", [False, True])
def test_linear_solvers_agree(oracle, global_missions):
    rng = np.random.default_rng(4)
    m = tiny_map(rng, dist_type=1, n_vertices=6, n_landmarks=40, global_missions=global_missions)
    o = options(max_num_iterations=8)
    out = [oracle.solve(m, o, linear_solver=ls, threads=1) for ls in (0, 1, 2)]
    costs = [[r.cost for r in rows] for _, rows, _, _, _ in out]
    for c in costs[1:]:
        np.testing.assert_allclose(c, costs[0], rtol=1e-9)
    np.testing.assert_allclose(out[1][2], out[0][2], atol=1e-9)
    assert costs[0][-1] < 0.5 * costs[0][0]


# ---------------------------------------------------------------- numpy LM


def numpy_lm(oracle, m, o, iters):
    """Ceres LM rules restated in numpy on dense normal equations, using the
    oracle's (separately pinned) tangent Jacobians of each observation."""
    from oracle.binding import problem_evaluate
    from scipy.linalg import cho_factor, cho_solve

    import independent as ind
    P, L = m.num_vertices, m.num_landmarks
    vfree = [v for v in range(P) if not m.vertex_constant[v]]
    lfree = [l for l in range(L) if not m.landmark_constant[l]]
    vi = {v: i for i, v in enumerate(vfree)}
    li = {l: 6 * len(vfree) + 3 * i for i, l in enumerate(lfree)}
    n = 6 * len(vfree) + 3 * len(lfree)
    types = []
    for k in range(m.num_observations):
        v, l = m.obs_vertex[k], m.obs_landmark[k]
        s = m.landmark_store_vertex[l]
        types.append(0 if v == s else (1 if m.vertex_mission[v] == m.vertex_mission[s] else 2))

    def loss_a(k):
        return (10.0 if types[k] == 2 else 3.0) * m.obs_sigma[k]

    def linearize(mm):
        r, J, _ = problem_evaluate(mm, tangent=True, options=o)
        A = np.zeros((2 * mm.num_observations, n))
        cost = 0.0
        for k in range(mm.num_observations):
            s = r[k] @ r[k]
            rho = oracle.loss(o.loss_type, loss_a(k), s)
            cost += 0.5 * rho[0]
            sq = np.sqrt(rho[1])
            r[k] *= sq
            l, v, a = m.obs_landmark[k], m.obs_vertex[k], m.landmark_store_vertex[m.obs_landmark[k]]
            if l in li:
                A[2 * k:2 * k + 2, li[l]:li[l] + 3] = J[0][k] * sq
            if types[k] != 0:
                if v in vi:
                    A[2 * k:2 * k + 2, 6 * vi[v]:6 * vi[v] + 6] += J[4][k] * sq
                if a in vi:
                    A[2 * k:2 * k + 2, 6 * vi[a]:6 * vi[a] + 6] += J[1][k] * sq
        return r.reshape(-1), A, cost

    def plus(mm, delta):
        out = mm.copy()
        for v in vfree:
            d = delta[6 * vi[v]:6 * vi[v] + 6]
            out.vertices[v] = np.concatenate([ind.jpl_plus(mm.vertices[v, :4], d[:3]), mm.vertices[v, 4:] + d[3:]])
        for l in lfree:
            out.landmarks[l] = mm.landmarks[l] + delta[li[l]:li[l] + 3]
        return out.normalized()

    x = m.copy().normalized()
    r, A, cost = linearize(x)
    scale = 1.0 / (1.0 + np.sqrt((A * A).sum(0)))
    radius, dec = o.initial_trust_region_radius, 2.0
    log = [cost]
    for _ in range(iters):
        As = A * scale
        diag = np.clip((As * As).sum(0), o.min_lm_diagonal, o.max_lm_diagonal)
        y = cho_solve(cho_factor(As.T @ As + np.diag(diag / radius)), As.T @ r)
        step = -y
        Js = As @ step
        mcc = -(Js @ (r + Js / 2))
        cand = plus(x, step * scale)
        _, _, ccost = linearize(cand)
        rho = (cost - ccost) / mcc
        if rho > o.min_relative_decrease:
            x = cand
            r, A, cost = linearize(x)
            radius = min(o.max_trust_region_radius, radius / max(1 / 3, 1 - (2 * rho - 1) ** 3))
            dec = 2.0
            log.append(cost)
        else:
            radius /= dec
            dec *= 2
            log.append(ccost)
    return log


def test_numpy_restatement_reproduces_iteration_log(oracle):
    rng = np.random.default_rng(8)
    m = tiny_map(rng, dist_type=2, n_vertices=5, n_landmarks=25, noise=0.5)
    # Bigger corruption so that some steps are rejected.
    m.landmarks = m.landmarks + (m.landmark_constant == 0)[:, None] * rng.normal(size=m.landmarks.shape) * 0.3
    m.normalized()
    # Huber: every step accepted, radius grows from 1.
    o = options(max_num_iterations=12, initial_trust_region_radius=1.0, function_tolerance=0.0,
                parameter_tolerance=0.0, gradient_tolerance=0.0)
    sm, rows, _, _, _ = oracle.solve(m, o, linear_solver=1, threads=1)
    ref = numpy_lm(oracle, m, o, len(rows) - 1)
    np.testing.assert_allclose([r.cost for r in rows], ref, rtol=1e-9)
    # No loss and a strict acceptance threshold: accepted and rejected steps mix.
    o = options(max_num_iterations=12, loss_type=abi.LOSS_TRIVIAL, min_relative_decrease=0.999,
                function_tolerance=0.0, parameter_tolerance=0.0, gradient_tolerance=0.0)
    sm, rows, _, _, _ = oracle.solve(m, o, linear_solver=1, threads=1)
    flags = [r.step_is_successful for r in rows[1:]]
    assert 0 < sum(flags) < len(flags)
    ref = numpy_lm(oracle, m, o, len(rows) - 1)
    np.testing.assert_allclose([r.cost for r in rows], ref, rtol=1e-9)
