"""Several independent maps solved concurrently (mlba_solve_many) on the
device, through the C ABI.

Bar: each map's result is the one it gets when solved alone, bit for bit
(same iteration log and final state): the handles share nothing, and every
solve is deterministic.
"""
import numpy as np
import pytest

from maplab_amd import abi, problem
from maplab_amd.solver import DeviceProblem

pytestmark = pytest.mark.gpu


def maps():
    sizes = [(40, 8000, 3), (90, 20000, 4), (60, 12000, 5), (120, 30000, 6), (50, 9000, 7)]
    out = []
    for kf, lm, seed in sizes:
        kw = {} if seed % 2 else dict(distortion_type=2, distortion=(-0.01, 0.02, -0.005, 0.001))
        out.append(problem.workload_map(2, seed=seed, num_keyframes=kf, num_landmarks=lm, **kw))
    return out


def opts(n):
    o = abi.default_options()
    o.max_num_iterations = n
    return o


@pytest.mark.parametrize("threads", [2, 5])
def test_solve_many_equals_solving_each_map_alone(threads):
    ms = maps()
    alone = []
    for m in ms:
        with DeviceProblem(m) as dev:
            sm, rows = dev.solve(opts(6))
            alone.append((sm.final_cost, [r.cost for r in rows], dev.download_state()))
    devs = [DeviceProblem(m) for m in ms]
    try:
        res = DeviceProblem.solve_many(devs, opts(6), threads=threads)
        for (sm, rows), dev, (c, costs, (v, l)) in zip(res, devs, alone):
            print(f"final cost {sm.final_cost:.12e} alone {c:.12e}", flush=True)
            assert sm.final_cost == c
            assert [r.cost for r in rows] == costs
            v2, l2 = dev.download_state()
            np.testing.assert_array_equal(v2, v)
            np.testing.assert_array_equal(l2, l)
    finally:
        for d in devs:
            d.close()


def test_solve_many_with_outlier_rejection_matches_each_alone():
    ms = maps()[:3]
    o = opts(6)
    o.reject_outliers_every_n_iters = 3
    alone = []
    for m in ms:
        with DeviceProblem(m) as dev:
            sm, _ = dev.solve(o, outlier_rejection=True)
            alone.append(sm.final_cost)
    devs = [DeviceProblem(m) for m in ms]
    try:
        res = DeviceProblem.solve_many(devs, o, threads=3, outlier_rejection=True)
        assert [sm.final_cost for sm, _ in res] == alone
    finally:
        for d in devs:
            d.close()
