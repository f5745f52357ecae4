"""The Schur block kernel's launch schedule (dispatch order, items per
work-group, XCD mapping), the PCG's way of summing its per-group partials
and the autotuner change no result: solves under any schedule repeat the
default schedule's iteration log and state bit for bit."""
import numpy as np
import pytest

from maplab_amd import abi, problem
from maplab_amd.solver import DeviceProblem

pytestmark = pytest.mark.gpu


def solve(m, schedule=None, autotune=False):
    o = abi.default_options()
    o.max_num_iterations = 6
    with DeviceProblem(m) as dev:
        tuned = None
        if autotune:
            tuned = dev.autotune(o, repeats=2)
        elif schedule is not None:
            dev.set_schedule(*schedule)
        sc = dev.schedule()
        sm, rows = dev.solve(o)
        v, l = dev.download_state()
    return sm, [(r.cost, r.step_is_successful, r.linear_solver_iterations) for r in rows], v, l, sc, tuned


@pytest.fixture(scope="module")
def walk():
    return problem.workload_map(2, seed=4, num_keyframes=60, num_landmarks=12000)


@pytest.fixture(scope="module")
def reference(walk):
    return solve(walk)


@pytest.mark.parametrize("schedule", [(4, 4, 1), (1, 2, 1), (32, 8, 0), (7, 4, 1), (0, 1, 0, 1), (0, 1, 0, 2)])
def test_explicit_schedules_repeat_the_default_solve(walk, reference, schedule):
    sm, rows, v, l, sc, _ = solve(walk, schedule=schedule)
    print(f"schedule {schedule}: {sm.num_iterations} rows, final {sm.final_cost:.12e}", flush=True)
    assert (sc.item_band_rows, sc.item_waves, sc.xcd_mapping) == schedule[:3]
    assert sc.pcg_partials == (schedule[3] if len(schedule) > 3 else 0)
    assert sm.final_cost == reference[0].final_cost
    assert rows == reference[1]
    np.testing.assert_array_equal(v, reference[2])
    np.testing.assert_array_equal(l, reference[3])


def test_autotune_keeps_the_fastest_schedule_and_the_results(walk, reference):
    sm, rows, v, l, sc, tuned = solve(walk, autotune=True)
    print(f"autotune: tried {tuned.schedules_tried}, default {tuned.default_ms:.4f} ms, "
          f"tuned {tuned.tuned_ms:.4f} ms -> ({tuned.item_band_rows}, {tuned.item_waves}, {tuned.xcd_mapping})",
          flush=True)
    assert tuned.schedules_tried == 7 * 4 * 2
    assert 0.0 < tuned.tuned_ms <= tuned.default_ms
    assert (sc.item_band_rows, sc.item_waves, sc.xcd_mapping) == (tuned.item_band_rows, tuned.item_waves,
                                                                  tuned.xcd_mapping)
    assert rows == reference[1]
    np.testing.assert_array_equal(v, reference[2])
    np.testing.assert_array_equal(l, reference[3])


def test_invalid_schedules_are_rejected(walk):
    with DeviceProblem(walk) as dev:
        for bad in [(-1, 4, 1), (4, 3, 1), (4, 16, 1), (4, 4, 2), (5000, 4, 0), (4, 4, 1, 3), (4, 4, 1, -1)]:
            with pytest.raises(RuntimeError, match="maplab_ba error -1"):
                dev.set_schedule(*bad)
        sc = dev.schedule()
        assert (sc.item_band_rows, sc.item_waves, sc.xcd_mapping, sc.pcg_partials) == (0, 1, 0, 0)
