"""`optv` on a maplab map folder on the device (maplab_amd/vimap.py +
VIMapOptimizer through the C ABI): the optimized map written back equals
optimizing the same problem from its arrays."""
import numpy as np
import pytest

from maplab_amd import problem, vimap
from maplab_amd.optimizer import VIMapOptimizer, default_vi_problem_options

pytestmark = pytest.mark.gpu


def test_optv_on_a_map_folder_matches_optimizing_the_arrays(tmp_path):
    m = problem.workload_map(2, seed=4, num_keyframes=40, num_landmarks=6000)
    folder = str(tmp_path / "map")
    vimap.write_vimap(m, folder)
    o = default_vi_problem_options()
    o.solver.max_num_iterations = 6
    res = vimap.optimize_map_folder(folder, out_folder=str(tmp_path / "out"), options=o)
    # The same problem from the arrays read back (landmark order of the folder).
    r, rec = vimap.load_vimap(folder)
    q = np.asarray([lm.quality for lm in rec.landmark_protos], dtype=np.uint8)
    ref = VIMapOptimizer(o).optimize(r, landmark_quality=q)
    print(f"folder final {res.summary.final_cost:.12e} arrays {ref.summary.final_cost:.12e}", flush=True)
    assert res.summary.final_cost == ref.summary.final_cost
    out, _ = vimap.load_vimap(str(tmp_path / "out"))
    np.testing.assert_array_equal(out.vertices, ref.vertices)
    np.testing.assert_array_equal(out.landmarks, ref.landmarks)
    assert res.summary.final_cost < 0.5 * res.summary.initial_cost
