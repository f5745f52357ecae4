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
