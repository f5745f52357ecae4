"""Calibration fixtures restated from the reference's visual-term tests
(algorithms/ceres-error-terms/test/test_visual_term_test.cc): pinhole
camera 640 x 480 with FOV-fisheye distortion, sigma 0.7 px, a landmark base
vertex at the identity, all map blocks constant, no loss. Quaternions are
given as Eigen's coefficient order (x, y, z, w), as the test passes
`coeffs()` to the error term."""
from __future__ import annotations

import numpy as np

from maplab_amd import abi
from maplab_amd.problem import VisualMap

S2 = np.sqrt(2.0) / 2.0


def fixture(landmarks, observers, keypoints, intrinsics, w=0.0, q_CI=(0, 0, 0, 1.0), p_CI=(0, 0, 0)):
    """observers: (q xyzw, position) of each observation's vertex; landmark k
    is observed by observation k (one landmark per observation when
    len(landmarks) == len(keypoints), else landmark 0 for all)."""
    n = len(keypoints)
    verts = [[0, 0, 0, 1.0, 0, 0, 0]] + [list(q) + list(p) for q, p in observers]
    lm = np.asarray(landmarks, dtype=np.float64).reshape(-1, 3)
    obs_l = np.arange(n) if lm.shape[0] == n else np.zeros(n)
    return VisualMap(
        vertices=np.array(verts, dtype=np.float64), baseframes=np.array([[0, 0, 0, 1.0, 0, 0, 0]]),
        landmarks=lm, camera_q_CI__C_p_CI=np.array([list(q_CI) + list(p_CI)], dtype=np.float64),
        camera_intrinsics=np.array([intrinsics], dtype=np.float64),
        camera_distortion=np.array([[w, 0, 0, 0]], dtype=np.float64),
        camera_distortion_type=np.array([abi.DIST_FISHEYE], np.int32),
        camera_resolution=np.array([[640, 480]], np.int32), vertex_mission=np.zeros(len(verts), np.int32),
        landmark_store_vertex=np.zeros(lm.shape[0], np.int32), obs_vertex=np.arange(1, n + 1, dtype=np.int32),
        obs_landmark=np.asarray(obs_l, np.int32), obs_camera=np.zeros(n, np.int32),
        obs_keypoint=np.array(keypoints, dtype=np.float64), obs_sigma=np.full(n, 0.7),
        vertex_constant=np.ones(len(verts), np.uint8), landmark_constant=np.ones(lm.shape[0], np.uint8),
    ).normalized()


def intrinsics_case():
    # :231-276: start (fu + 18, fv - 90, cu - 10, cv + 13) -> (100, 100, 320, 240).
    m = fixture([0, 0, 0], [((0, 0, 0, 1.0), (-1, 0.5, -1)), ((0, -S2, 0, S2), (1, -0.3, 1.2))],
                [[420.0, 190.0], [200.0, 270.0]], [118.0, 10.0, 310.0, 253.0])
    return m, dict(fix_intrinsics=False), (100.0, 100.0, 320.0, 240.0)


def distortion_case():
    # :281-321: FOV w from 0.95 to the Matlab keypoint's w = 1.
    m = fixture([0, 0, 0], [((0, 0, 0, 1.0), (-1, 0.5, -1))], [[399.139364932153, 200.430317533923]],
                [100.0, 100.0, 320.0, 240.0], w=0.95)
    return m, dict(fix_distortion=False), 1.0


def cam_to_imu_case():
    # :450-505: extrinsics from a perturbed start to identity translation and
    # Hamilton (w, x, y, z) = (sqrt(1/2), 0, -sqrt(1/2), 0).
    off = 0.5
    m = fixture([[0, 0, -off], [0, 0, off], [0, off, 0]], [((0, 0, 0, 1.0), (1, 0, 0))] * 3,
                [[-100 * off + 320, 240], [100 * off + 320, 240], [320, 100 * off + 240]],
                [100.0, 100.0, 320.0, 240.0],
                q_CI=(0.013627671275108364, -0.7236496655350155, -0.0489853308190429, 0.6882920059910812),
                p_CI=(0.2, 0.1, -0.3))
    return m, dict(fix_rotation=False, fix_translation=False), (0.0, -S2, 0.0, S2)


def kat_options():
    # solveProblem (:126-137): 1e4 iterations, tolerances 1e-50, no loss.
    o = abi.default_options()
    o.max_num_iterations = 10000
    o.function_tolerance = o.gradient_tolerance = o.parameter_tolerance = 1e-50
    o.loss_type = abi.LOSS_TRIVIAL
    return o
