"""optv through the console plugin equals VIMapOptimizer::optimize with the
same flags (the plugin builds ViProblemOptions from the reference flag names)."""
import numpy as np
import pytest

from cases import tiny_map
from maplab_amd import optimizer, plugin

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("line", [
    "optv --ba_num_iterations=5 --ba_use_visual_outlier_rejection_solver=false --elq_min_observers=3",
    "optvi --ba_include_inertial=false --ba_num_iterations 4 --elq_min_observers 3 --ba_fix_landmark_positions",
])
def test_optv_equals_vimap_optimizer(line):
    m = tiny_map(np.random.default_rng(5), dist_type=1, n_vertices=8, n_landmarks=80)
    m.vertex_constant[:] = 0
    m.landmark_constant[:] = 0
    print(f"plugin: {line}", flush=True)
    c = plugin.MapConsole(m)
    rc, msg = c.run(line)
    assert rc == plugin.CMD_SUCCESS, msg
    v, l = c.state
    o = optimizer.default_vi_problem_options()
    o.landmark_quality.min_observers = 3
    if "optvi" in line:
        o.solver.max_num_iterations = 4
        o.fix_landmark_positions = 1
    else:
        o.solver.max_num_iterations = 5
        o.enable_visual_outlier_rejection = 0
    res = optimizer.VIMapOptimizer(o).optimize(m)
    assert c.summary.num_iterations == res.summary.num_iterations
    np.testing.assert_array_equal(v, res.vertices)
    np.testing.assert_array_equal(l, res.landmarks)
    c.close()


def test_optv_refines_intrinsics_with_the_map_fixed():
    # --ba_fix_ncamera_intrinsics=false with the vertices and landmarks
    # fixed: the calibration blocks are refined and written back; with a
    # variable map the command fails with the reason.
    m = tiny_map(np.random.default_rng(8), dist_type=1, n_vertices=6, n_landmarks=60)
    m.vertices[:], m.landmarks[:] = m.true_vertices, m.true_landmarks
    true_intr = m.camera_intrinsics.copy()
    m.camera_intrinsics[0] += [5.0, -4.0, 3.0, -2.0]
    m.normalized()
    c = plugin.MapConsole(m)
    rc, msg = c.run("optv --ba_fix_ncamera_intrinsics=false --elq_min_observers=3")
    assert rc == plugin.CMD_UNKNOWN_ERROR and "held constant" in msg, msg
    rc, msg = c.run("optv --ba_fix_ncamera_intrinsics=false --ba_fix_vertices --ba_fix_landmark_positions "
                    "--elq_min_observers=3 --ba_num_iterations=20")
    assert rc == plugin.CMD_SUCCESS, msg
    q, intr, dist, xi = c.cameras
    print(f"refined intrinsics {intr[0]} (true {true_intr[0]}), distortion {dist[0]}", flush=True)
    # Started 5 px off; the distortion is refined with the intrinsics (the
    # flag covers both), which trades off against the focal lengths on a
    # 6-keyframe map with 0.5 px noise.
    assert np.abs(intr[0] - true_intr[0]).max() < 1.0
    assert c.summary.final_cost < 0.1 * c.summary.initial_cost
    np.testing.assert_array_equal(c.state[0], m.vertices)
    c.close()
