"""Recycled deflation space of the PCG solve (Ritz vectors carried from
solve to solve) on the device, through the C ABI.

Bar: the recycled space changes only how fast each linear solve converges,
not what it converges to (relative residual 1e-10), so the LM log with and
without it is the same run (same accept/reject sequence, costs within 1e-9
relative) while the PCG iterations drop; runs repeat bit for bit. Parity of
the recycled solver with the oracle is covered by test_gpu_parity.py, which
runs with the default options (recycling on).
"""
import numpy as np
import pytest

from maplab_amd import abi, problem
from maplab_amd.solver import DeviceProblem

pytestmark = pytest.mark.gpu


def opts(**kw):
    o = abi.default_options()
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def run(m, **kw):
    with DeviceProblem(m) as dev:
        sm, rows = dev.solve(opts(**kw))
    return sm, rows


@pytest.fixture(scope="module")
def walk():
    # 500 keyframes, 200k landmarks: four accepted steps, then a chain of
    # rejected ones whose systems differ only in the LM damping.
    return problem.workload_map(2, seed=1)


def test_recycling_keeps_the_lm_log_and_cuts_pcg_iterations(walk):
    s_off, r_off = run(walk, max_num_iterations=12, pcg_ritz_vectors=-1)
    s_on, r_on = run(walk, max_num_iterations=12)
    it_off = [r.linear_solver_iterations for r in r_off]
    it_on = [r.linear_solver_iterations for r in r_on]
    print(f"PCG iterations per LM step without recycling {it_off} ({sum(it_off)}), "
          f"with {it_on} ({sum(it_on)})", flush=True)
    assert [r.step_is_successful for r in r_on] == [r.step_is_successful for r in r_off]
    assert sum(r.step_is_successful == 0 for r in r_on) >= 3  # the chain of rejected steps is there
    np.testing.assert_allclose([r.cost for r in r_on], [r.cost for r in r_off], rtol=1e-9)
    np.testing.assert_allclose(s_on.final_cost, s_off.final_cost, rtol=1e-9)
    assert sum(it_on) < 0.8 * sum(it_off)


@pytest.mark.parametrize("r", [4, 12])
def test_recycled_runs_repeat_bit_for_bit(walk, r):
    a = run(walk, max_num_iterations=8, pcg_ritz_vectors=r)[1]
    b = run(walk, max_num_iterations=8, pcg_ritz_vectors=r)[1]
    assert [x.cost for x in a] == [x.cost for x in b]
    assert [x.linear_solver_iterations for x in a] == [x.linear_solver_iterations for x in b]


def test_small_map_recycling_agrees_with_plain_deflation():
    # A 60-keyframe random walk (radtan): every iteration's cost, recycling on and off.
    m = problem.workload_map(2, seed=2, num_keyframes=60, num_landmarks=12000)
    _, a = run(m, max_num_iterations=10, pcg_ritz_vectors=-1)
    _, b = run(m, max_num_iterations=10, pcg_ritz_vectors=6)
    assert [r.step_is_successful for r in a] == [r.step_is_successful for r in b]
    np.testing.assert_allclose([r.cost for r in b], [r.cost for r in a], rtol=1e-9)
