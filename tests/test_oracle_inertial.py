"""The oracle's inertial error term (RK4 IMU integration with transition and
covariance propagation, imu-integrator-inl.h:17-338, and
InertialErrorTerm::Evaluate, inertial-error-term.cc:73-286) against closed
forms, an independent ODE integration (scipy) and numerical derivatives."""
import numpy as np
import pytest
from scipy.integrate import solve_ivp
from scipy.linalg import expm

import independent as ind

G = 9.81
NOISE = [1e-3, 1e-4, 1e-2, 1e-3, G]   # gyro noise, gyro bias, acc noise, acc bias sigmas, gravity


def times(n, dt=0.005):
    return (np.arange(n) * dt * 1e9).astype(np.int64)


def state_params(q, bg, v, ba, p):
    return [np.concatenate([q, p]), np.asarray(bg, float), np.asarray(v, float), np.asarray(ba, float)]


def test_stationary_imu_keeps_the_state(oracle):
    imu = np.tile([0, 0, G, 0, 0, 0], (21, 1)).astype(float)
    fr = state_params([0, 0, 0, 1.0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [1, 2, 3])
    r, _, _, Q, xe = oracle.inertial_evaluate(imu, times(21), NOISE, fr + fr)
    np.testing.assert_allclose(xe, [0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 2, 3], atol=1e-15)
    np.testing.assert_allclose(r, 0.0, atol=1e-12)
    assert np.allclose(Q, Q.T) and np.linalg.eigvalsh(Q).min() > 0


def test_constant_yaw_rate_matches_quaternion_exponential(oracle):
    w = 0.7
    imu = np.tile([0, 0, G, 0, 0, w], (41, 1)).astype(float)
    q0 = np.array([0.1, -0.2, 0.05, 1.0]); q0 /= np.linalg.norm(q0)
    fr = state_params(q0, [0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0])
    _, _, _, _, xe = oracle.inertial_evaluate(imu, times(41), NOISE, fr + fr)
    Om = np.array([[0, w, 0, 0], [-w, 0, 0, 0], [0, 0, 0, w], [0, 0, -w, 0]], float)  # gyroOmegaJPL(0, 0, w)
    q_ref = expm(0.5 * Om * 0.2) @ q0
    np.testing.assert_allclose(xe[:4], q_ref, atol=1e-10)


def test_constant_world_acceleration_is_exact(oracle):
    a = np.array([0.3, -0.2, 0.1])
    imu = np.tile(np.concatenate([a + [0, 0, G], [0, 0, 0]]), (31, 1))
    v0, p0 = np.array([1.0, 2.0, 0.0]), np.array([0.5, 0.0, -1.0])
    fr = state_params([0, 0, 0, 1.0], [0, 0, 0], v0, [0, 0, 0], p0)
    _, _, _, _, xe = oracle.inertial_evaluate(imu, times(31), NOISE, fr + fr)
    T = 0.15
    np.testing.assert_allclose(xe[7:10], v0 + a * T, atol=1e-13)
    np.testing.assert_allclose(xe[13:16], p0 + v0 * T + 0.5 * a * T * T, atol=1e-13)


def random_edge(seed, n=41):
    rng = np.random.default_rng(seed)
    t = times(n)
    ts = t * 1e-9
    imu = np.stack([0.5 * np.sin(3 * ts + 1), 0.3 * np.cos(2 * ts), G + 0.2 * np.sin(5 * ts),
                    0.2 * np.sin(4 * ts), -0.3 * np.cos(3 * ts + 0.5), 0.4 * np.sin(ts)], axis=1)
    q0 = rng.normal(size=4); q0 /= np.linalg.norm(q0); q0 *= np.sign(q0[3])
    fr = state_params(q0, rng.normal(size=3) * 0.01, rng.normal(size=3), rng.normal(size=3) * 0.02,
                      rng.normal(size=3))
    return imu, t, fr


def ode_end_state(imu, t, x0):
    """The continuous IMU kinematics with linearly interpolated readings,
    integrated by scipy's adaptive RK45 (q_dot = 0.5 Omega(w) q in JPL form,
    v_dot = R^T a - g e_z, p_dot = v)."""
    ts = t * 1e-9
    bg, ba = x0[4:7], x0[10:13]

    def reading(s):
        return np.array([np.interp(s, ts, imu[:, k]) for k in range(6)])

    def f(s, y):
        r = reading(s)
        a, w = r[:3] - ba, r[3:] - bg
        q = y[:4] / np.linalg.norm(y[:4])
        Om = np.array([[0, w[2], -w[1], w[0]], [-w[2], 0, w[0], w[1]], [w[1], -w[0], 0, w[2]],
                       [-w[0], -w[1], -w[2], 0]])
        R = ind.R_jpl(q)
        return np.concatenate([0.5 * Om @ q, R.T @ a - [0, 0, G], y[4:7]])

    y0 = np.concatenate([x0[:4], x0[7:10], x0[13:16]])
    sol = solve_ivp(f, (ts[0], ts[-1]), y0, rtol=1e-12, atol=1e-13, max_step=0.005)
    return sol.y[:, -1]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_end_state_matches_independent_ode_integration(oracle, seed):
    imu, t, fr = random_edge(seed)
    _, _, _, _, xe = oracle.inertial_evaluate(imu, t, NOISE, fr + fr)
    x0 = np.concatenate([fr[0][:4], fr[1], fr[2], fr[3], fr[0][4:]])
    y = ode_end_state(imu, t, x0)
    np.testing.assert_allclose(xe[:4] / np.linalg.norm(xe[:4]), y[:4] / np.linalg.norm(y[:4]), atol=1e-9)
    np.testing.assert_allclose(xe[7:10], y[4:7], atol=1e-9)
    np.testing.assert_allclose(xe[13:16], y[7:10], atol=1e-9)


def numdiff(f, x, h):
    cols = []
    for i in range(len(x)):
        e = np.zeros(len(x)); e[i] = h
        cols.append((8 * f(x + e) + f(x - 2 * e) - f(x + 2 * e) - 8 * f(x - e)) / (12 * h))
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("seed", [0, 1])
def test_jacobians_match_numerical_derivatives(oracle, seed):
    # At a consistent edge (to-state = integrated end state) the tangent
    # Jacobians equal the derivative of the unwhitened residual L r with the
    # whitening L held fixed, as Ceres sees them.
    imu, t, fr = random_edge(seed)
    _, _, _, Q, xe = oracle.inertial_evaluate(imu, t, NOISE, fr + fr)
    to = state_params(xe[:4] / np.linalg.norm(xe[:4]), xe[4:7], xe[7:10], xe[10:13], xe[13:16])
    r, jac, _, Q, _ = oracle.inertial_evaluate(imu, t, NOISE, fr + to)
    L = np.linalg.cholesky(Q)
    np.testing.assert_allclose(L @ r, 0.0, atol=1e-9)

    def unwhitened(params):
        rr, _, _, _, _ = oracle.inertial_evaluate(imu, t, NOISE, params, want_jac=False)
        return L @ rr

    for side in (0, 1):
        base = fr + to
        # pose: rotation through JPL Plus (tangent 3), position plain
        q = base[4 * side][:4]

        def f_rot(d):
            p = list(base); p[4 * side] = np.concatenate([ind.jpl_plus(q, d), base[4 * side][4:]])
            return unwhitened(p)
        Jq_num = numdiff(f_rot, np.zeros(3), 1e-6)
        Jq = L @ jac[4 * side][:, :4] @ np.asarray(oracle.quat_jacobian(q))
        np.testing.assert_allclose(Jq, Jq_num, atol=2e-5 * np.abs(Jq_num).max())
        for b, lo in ((0, 4), (1, None), (2, None), (3, None)):
            def f_lin(d, b=b):
                p = list(base)
                if b == 0:
                    p[4 * side] = np.concatenate([base[4 * side][:4], base[4 * side][4:] + d])
                else:
                    p[4 * side + b] = base[4 * side + b] + d
                return unwhitened(p)
            J_num = numdiff(f_lin, np.zeros(3), 1e-5)
            J = L @ (jac[4 * side][:, 4:] if b == 0 else jac[4 * side + b])
            np.testing.assert_allclose(J, J_num, atol=2e-5 * max(1.0, np.abs(J_num).max()))
