"""Pins the oracle's VisualReprojectionError::Evaluate
(visual-error-term-inl.h:13-379) against the independent Hamilton-rotation
restatement in tests/independent.py: residuals, and every Jacobian block
lifted to the tangent space through the JPL local parameterization and
compared with a 5-point numerical derivative through JPL Plus."""
import numpy as np
import pytest

import independent as ind
from cases import HEIGHT, WIDTH, observation_case


def lift(oracle, b, J, params):
    """Ambient -> tangent with the product parameterization of the block."""
    if b in (1, 2, 3, 4):
        Jq = oracle.quat_jacobian(params[b][:4])
        return np.concatenate([J[:, :4] @ Jq, J[:, 4:]], axis=1)
    if b == 5:
        return J @ oracle.quat_jacobian(params[5])
    return J


@pytest.mark.parametrize("error_type", [ind.LOCAL_KEYFRAME, ind.LOCAL_MISSION, ind.GLOBAL])
@pytest.mark.parametrize("dist_type", [0, 1, 2, 3])
def test_residual_and_tangent_jacobians(oracle, error_type, dist_type):
    rng = np.random.default_rng(100 + 10 * error_type + dist_type)
    for _ in range(8):
        meas, sigma, params = observation_case(rng, error_type, dist_type, bf_translation=False)
        failed, r, J = oracle.evaluate_visual(meas, sigma, error_type, dist_type, WIDTH, HEIGHT, params)
        assert not failed
        r_ind, _ = ind.residual(meas, sigma, error_type, dist_type, params)
        np.testing.assert_allclose(r, r_ind, rtol=1e-12, atol=1e-9)
        num = ind.tangent_jacobians(meas, sigma, error_type, dist_type, params)
        for b, Jn in num.items():
            Jt = lift(oracle, b, J[b], params)
            scale = max(1.0, np.abs(Jn).max())
            np.testing.assert_allclose(Jt / scale, Jn / scale, atol=2e-7, err_msg=f"block {b}")


def test_global_baseframe_blocks_without_translation(oracle):
    # With zero baseframe translations the reference's [p_G]x forms
    # (visual-error-term-inl.h:208-209) are the true derivatives.
    rng = np.random.default_rng(5)
    for _ in range(5):
        meas, sigma, params = observation_case(rng, ind.GLOBAL, 1, bf_translation=False)
        _, _, J = oracle.evaluate_visual(meas, sigma, ind.GLOBAL, 1, WIDTH, HEIGHT, params)
        pose_plus = lambda x, d: np.concatenate([ind.jpl_plus(x[:4], d[:3]), x[4:] + d[3:]])
        for b in (2, 3):
            def f(d, b=b):
                pp = [p.copy() for p in params]
                pp[b] = pose_plus(pp[b], d)
                return ind.residual(meas, sigma, ind.GLOBAL, 1, pp)[0]
            Jn = ind.numdiff(f, 6)
            np.testing.assert_allclose(lift(oracle, b, J[b], params), Jn, rtol=1e-6, atol=1e-6)


def test_keyframe_landmark_jacobian_is_projection_times_extrinsics(oracle):
    # visual-error-term-inl.h:225-227: J_pC/p_B = R_C_I for the local keyframe type.
    rng = np.random.default_rng(9)
    meas, sigma, params = observation_case(rng, ind.LOCAL_KEYFRAME, 1)
    _, r, J = oracle.evaluate_visual(meas, sigma, 0, 1, WIDTH, HEIGHT, params)
    p_C = ind.p_C_of(0, *params[:7])
    st, kp, jp, _, _ = oracle.project(1, params[7], params[8], WIDTH, HEIGHT, p_C)
    np.testing.assert_allclose(J[0], jp @ oracle.rotation_jpl(params[5]) / sigma, rtol=1e-13, atol=1e-12)
    # Unused (dummy) blocks are not touched by the local keyframe type.
    assert not np.any(J[1]) and not np.any(J[4])


def test_ambient_rotation_blocks_relation(oracle):
    # ambient = tangent * 4 Jq^T / |q|^2 because Jq^T Jq = |q|^2 / 4 I.
    rng = np.random.default_rng(11)
    meas, sigma, params = observation_case(rng, ind.LOCAL_MISSION, 2)
    _, _, J = oracle.evaluate_visual(meas, sigma, 1, 2, WIDTH, HEIGHT, params)
    for b in (1, 4):
        Jq = oracle.quat_jacobian(params[b][:4])
        np.testing.assert_allclose(Jq.T @ Jq, 0.25 * np.eye(3), atol=1e-15)
        tan = J[b][:, :4] @ Jq
        np.testing.assert_allclose(J[b][:, :4], tan @ (4 * Jq.T), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("p_C_z, failed", [(0.04, True), (0.06, False), (-1.0, True)])
def test_projection_failure_zeroes_residual_and_jacobians(oracle, p_C_z, failed):
    # visual-error-term-inl.h:159-178: BEHIND, INVALID, z < 0.05 or |kp|^2 > 1e10.
    params = [np.array([0.001, 0.002, p_C_z]), np.array([0, 0, 0, 1.0, 0, 0, 0]), np.array([0, 0, 0, 1.0, 0, 0, 0]),
              np.array([0, 0, 0, 1.0, 0, 0, 0]), np.array([0, 0, 0, 1.0, 0, 0, 0]), np.array([0, 0, 0, 1.0]),
              np.zeros(3), np.array([460.0, 460, 376, 240]), np.zeros(4)]
    f, r, J = oracle.evaluate_visual([100.0, 100.0], 0.5, 0, 0, WIDTH, HEIGHT, params)
    assert f == failed
    if failed:
        assert not np.any(r) and not any(np.any(j) for j in J)
    else:
        assert np.any(r)
