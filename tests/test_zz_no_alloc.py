"""mlba_evaluate_noalloc (include/maplab_ba.h): the allocation-free variant of
mlba_evaluate returns the same residuals, Jacobians and cost bit for bit, on
two calls in a row (the second on a smaller map, so the shared staging array
holds the first call's data), and rejects a request beyond its fixed array."""
import numpy as np
import pytest

from cases import tiny_map
from maplab_amd import abi, problem
from maplab_amd.solver import DeviceProblem

pytestmark = pytest.mark.gpu


def _same(dev, tangent, blocks, with_xi=False):
    a = dev.evaluate(tangent=tangent, blocks=blocks, with_xi=with_xi)
    b = dev.evaluate(tangent=tangent, blocks=blocks, with_xi=with_xi, noalloc=True)
    np.testing.assert_array_equal(a[0], b[0])
    assert set(a[1]) == set(b[1])
    for k in a[1]:
        np.testing.assert_array_equal(a[1][k], b[1][k])
    assert a[2] == b[2]
    return b


def test_noalloc_evaluate_matches_evaluate_on_two_calls():
    big = tiny_map(np.random.default_rng(31), dist_type=2, n_vertices=8, n_landmarks=120, global_missions=True)
    small = tiny_map(np.random.default_rng(32), dist_type=1, n_vertices=5, n_landmarks=30)
    assert small.num_observations < big.num_observations
    print(f"noalloc evaluate: N {big.num_observations} then {small.num_observations}", flush=True)
    with DeviceProblem(big) as d1, DeviceProblem(small) as d2:
        r1 = _same(d1, True, range(9))
        r2 = _same(d2, False, range(9))
        assert r2[0].shape[0] < r1[0].shape[0]
        # Residuals and cost only, and a subset of the blocks.
        _same(d2, True, ())
        _same(d1, False, (0, 4, 8))


def test_noalloc_evaluate_rejects_requests_beyond_its_array():
    # ~0.5M observations: residuals fit the 64 MiB array, all tangent Jacobians do not.
    m = problem.workload_map(2, seed=1, num_keyframes=120, num_landmarks=60000)
    n = m.num_observations
    assert 8 * 2 * n <= (64 << 20) < 8 * (2 + 84) * n
    print(f"noalloc evaluate limit: N {n}", flush=True)
    with DeviceProblem(m) as dev:
        res, _, cost = dev.evaluate(tangent=True, blocks=())
        res2, _, cost2 = dev.evaluate(tangent=True, blocks=(), noalloc=True)
        np.testing.assert_array_equal(res, res2)
        assert cost == cost2
        with pytest.raises(RuntimeError, match="staging"):
            dev.evaluate(tangent=True, blocks=range(9), noalloc=True)
        # The handle still works after the rejected call.
        res3, _, _ = dev.evaluate(tangent=True, blocks=(), noalloc=True)
        np.testing.assert_array_equal(res, res3)
