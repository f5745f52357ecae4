"""VIMapOptimizer::optimize (mlba_vi_map_optimize: device admission, host
selection and gauge, device solve with outlier rejection) against the same
pipeline on the CPU oracle."""
import numpy as np
import pytest

from cases import tiny_map
from maplab_amd import abi, optimizer
from maplab_amd.problem import VisualMap
from test_problem_selection import concat

pytestmark = pytest.mark.gpu


def restrict(m: VisualMap, sel) -> VisualMap:
    """The map with only the selected observations and the selection's constant blocks."""
    keep = sel.obs_in_problem
    return VisualMap(
        vertices=m.vertices.copy(), baseframes=m.baseframes, landmarks=m.landmarks.copy(),
        camera_q_CI__C_p_CI=m.camera_q_CI__C_p_CI, camera_intrinsics=m.camera_intrinsics,
        camera_distortion=m.camera_distortion, camera_distortion_type=m.camera_distortion_type,
        camera_resolution=m.camera_resolution, vertex_mission=m.vertex_mission,
        landmark_store_vertex=m.landmark_store_vertex, obs_vertex=m.obs_vertex[keep],
        obs_landmark=m.obs_landmark[keep], obs_camera=m.obs_camera[keep], obs_keypoint=m.obs_keypoint[keep],
        obs_sigma=m.obs_sigma[keep], vertex_constant=sel.vertex_constant.astype(np.uint8),
        landmark_constant=sel.landmark_constant.astype(np.uint8)).normalized()


BEHIND = [5, 9, 80, 95]


def same_log(rows, ref_rows):
    """Same accept / reject sequence and accepted costs to 1e-9. Candidate
    costs of rejected (very long) steps amplify the PCG's 1e-10 relative
    residual against the oracle's direct solve: 1e-7 there."""
    assert [r.step_is_successful for r in rows] == [r.step_is_successful for r in ref_rows]
    acc = [k for k, r in enumerate(ref_rows) if r.step_is_successful or k == 0]
    np.testing.assert_allclose([rows[k].cost for k in acc], [ref_rows[k].cost for k in acc], rtol=1e-9)
    np.testing.assert_allclose([r.cost for r in rows], [r.cost for r in ref_rows], rtol=1e-7)


def two_cluster_map():
    rng = np.random.default_rng(11)
    a = tiny_map(rng, dist_type=1, n_vertices=8, n_landmarks=70, global_missions=True)
    b = tiny_map(rng, dist_type=2, n_vertices=6, n_landmarks=60)
    m = concat(a, b)
    # Landmarks pushed behind their store vertex: the admission test rejects
    # them unless their stored quality is GOOD, and then the outlier solver does.
    for l in BEHIND:
        m.landmarks[l, 2] = -2.0
    return m.normalized()


@pytest.mark.parametrize("outlier_rejection", [True, False])
def test_vi_map_optimizer_matches_oracle_pipeline(oracle, outlier_rejection):
    m = two_cluster_map()
    o = optimizer.default_vi_problem_options()
    o.solver.max_num_iterations = 7
    o.enable_visual_outlier_rejection = int(outlier_rejection)
    o.landmark_quality.min_observers = 3
    print(f"VIMapOptimizer: outlier rejection {outlier_rejection} N {m.num_observations}", flush=True)
    stored = np.zeros(m.num_landmarks, dtype=np.uint8)
    stored[BEHIND] = abi.QUALITY_GOOD
    res = optimizer.VIMapOptimizer(o).optimize(m, landmark_quality=stored)
    good = oracle.landmark_quality(m, o.landmark_quality)
    assert not good[BEHIND].any()  # the geometric test rejects them
    good[BEHIND] = True            # the stored GOOD quality admits them
    sel = optimizer.select_problem(m, good, o)
    assert len(sel.gauge_vertices) == 2
    ref, ref_rows, rv, rl, ract = oracle.solve(restrict(m, sel), o.solver, outlier_rejection=outlier_rejection,
                                               threads=1)
    assert res.summary.num_iterations == ref.num_iterations
    same_log(res.iterations, ref_rows)
    lin = sel.landmark_in_problem
    if outlier_rejection:
        assert res.summary.num_removed_landmarks == ref.num_removed_landmarks >= 4
        bad = lin & (ract == 0)
        assert (res.landmark_quality == abi.QUALITY_BAD).tolist() == bad.tolist()
        keep = lin & (ract == 1)
    else:
        keep = lin
        assert (res.landmark_quality == stored).all()
    np.testing.assert_allclose(res.landmarks[keep], rl[keep], atol=1e-8)
    np.testing.assert_allclose(res.vertices, rv, atol=1e-8)
    # Blocks outside the problem are copied back unchanged.
    np.testing.assert_array_equal(res.landmarks[~lin], m.landmarks[~lin])


def test_stored_quality_overrides_geometry(oracle):
    # Landmarks with a stored GOOD / BAD quality skip the geometric test
    # (isLandmarkWellConstrained, landmark-quality-metrics.cc:81-101).
    m = two_cluster_map()
    o = optimizer.default_vi_problem_options()
    o.solver.max_num_iterations = 2
    o.enable_visual_outlier_rejection = 0
    q = np.full(m.num_landmarks, abi.QUALITY_GOOD, dtype=np.uint8)
    q[::7] = abi.QUALITY_BAD
    q[0] = abi.QUALITY_GOOD  # keep a landmark of the first vertex for the gauge
    res = optimizer.VIMapOptimizer(o).optimize(m, landmark_quality=q)
    sel = optimizer.select_problem(m, q == abi.QUALITY_GOOD, o)
    ref, ref_rows, rv, rl, _ = oracle.solve(restrict(m, sel), o.solver, threads=1)
    same_log(res.iterations, ref_rows)
