"""Calibration refinement with the map held constant, on the oracle: the
reference's own calibration tests (test_visual_term_test.cc:231-321,
:450-505) as known answers, and the gradient the LM descends checked
against numerical differentiation of the cost."""
import numpy as np
import pytest

import calib_cases as cc
from cases import tiny_map
from maplab_amd import abi


def test_intrinsics_converge_to_ground_truth(oracle):
    m, fix, gt = cc.intrinsics_case()
    sm, rows, (q, i, d, x) = oracle.solve_calibration(m, cc.kat_options(), **fix)
    assert sm.final_cost < 1e-15
    np.testing.assert_allclose(i[0], gt, atol=1e-5)
    np.testing.assert_array_equal(q, m.camera_q_CI__C_p_CI)
    np.testing.assert_array_equal(d, m.camera_distortion)


def test_fisheye_parameter_converges_to_the_matlab_value(oracle):
    m, fix, gt = cc.distortion_case()
    sm, rows, (q, i, d, x) = oracle.solve_calibration(m, cc.kat_options(), **fix)
    assert sm.final_cost < 1e-15
    assert abs(d[0, 0] - gt) < 1e-5
    np.testing.assert_array_equal(i, m.camera_intrinsics)


def test_camera_to_imu_pose_converges(oracle):
    m, fix, gt = cc.cam_to_imu_case()
    sm, rows, (q, i, d, x) = oracle.solve_calibration(m, cc.kat_options(), **fix)
    assert sm.final_cost < 1e-15
    np.testing.assert_allclose(q[0, 4:], 0.0, atol=1e-10)
    qq = q[0, :4] * np.sign(q[0, 3])
    np.testing.assert_allclose(qq, gt, atol=1e-10)


def test_first_step_gradient_matches_numdiff_of_the_cost(oracle):
    # The iteration-0 gradient max norm the LM reports is |g|_inf of the
    # tangent gradient for Euclidean blocks; with only intrinsics and
    # distortion free it equals the numerical gradient of the (Huber) cost.
    rng = np.random.default_rng(3)
    m = tiny_map(rng, dist_type=1, n_vertices=5, n_landmarks=40)
    m.vertex_constant[:] = 1
    m.landmark_constant[:] = 1
    m.camera_intrinsics[0] += [3.0, -2.0, 1.5, -1.0]
    m.normalized()
    o = abi.default_options()
    o.max_num_iterations = 0
    sm, rows, _ = oracle.solve_calibration(m, o, fix_intrinsics=False, fix_distortion=False)

    def cost(p):
        mm = m.copy()
        mm.camera_intrinsics[0] = p[:4]
        mm.camera_distortion[0] = p[4:]
        return oracle.problem_evaluate(mm.normalized(), options=o)[2]

    p0 = np.concatenate([m.camera_intrinsics[0], m.camera_distortion[0]])
    h = np.array([1e-4] * 4 + [1e-6] * 4)
    grad = np.array([(cost(p0 + h * e) - cost(p0 - h * e)) / (2 * h[k]) for k, e in enumerate(np.eye(8))])
    assert abs(rows[0].gradient_max_norm - np.abs(grad).max()) <= 1e-6 * np.abs(grad).max()
    assert abs(rows[0].cost - cost(p0)) <= 1e-12 * rows[0].cost


def test_variable_map_blocks_are_rejected(oracle):
    rng = np.random.default_rng(4)
    m = tiny_map(rng, dist_type=1, n_vertices=5, n_landmarks=20)  # vertices 1.. and landmarks variable
    assert oracle.solve_calibration(m, abi.default_options(), fix_intrinsics=False) is None
    m.vertex_constant[:] = 1
    m.landmark_constant[:] = 1
    assert oracle.solve_calibration(m, abi.default_options()) is None  # nothing free
