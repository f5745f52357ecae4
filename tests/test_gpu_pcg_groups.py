"""PCG on maps with many row groups, through the C ABI.

Each PCG iteration sums per-group partials (r'z, r'r and the coarse-space
product of each group of 42 vertex rows) in group order. While they fit one
LDS chunk of the SpMV kernel, each of its work-groups sums them itself; past
that (more than 22 groups with the 15-node coarse space, more than 18 once 16
recycled vectors are added) the work-group that finishes the update last sums
them once for all (an integer arrival counter picks it). The schedule field
pcg_partials forces either way.

Bar: on a 1000-keyframe random walk (24 groups) and an 850-keyframe one
(21 groups: the automatic choice switches after the first solve) both ways
give the same iteration log and state bit for bit, and runs repeat bit for
bit; on the 1000-keyframe walk the LM log
agrees with the oracle's direct Schur solve to 1e-8 relative on the first
steps (the PCG solves to 1e-10).
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


@pytest.fixture(scope="module")
def walk1000():
    m = problem.workload_map(2, seed=3, num_keyframes=1000, num_landmarks=20000)
    assert (m.num_vertices + 41) // 42 > 22
    return m


def run(m, partials=0, **kw):
    with DeviceProblem(m) as dev:
        if partials:
            dev.set_schedule(4, 4, 1, partials)
        sm, rows = dev.solve(opts(**kw))
        v, l = dev.download_state()
    return sm, [(r.cost, r.step_is_successful, r.linear_solver_iterations) for r in rows], v, l


@pytest.fixture(scope="module")
def walk850():
    # 21 groups: the first solve (15-node coarse space only) sums per SpMV,
    # the later ones (16 recycled vectors added) once per update.
    m = problem.workload_map(2, seed=5, num_keyframes=850, num_landmarks=15000)
    assert 18 < (m.num_vertices + 41) // 42 <= 22
    return m


@pytest.mark.parametrize("size", [1000, 850])
def test_both_partial_sums_give_the_same_solve(request, size):
    m = request.getfixturevalue(f"walk{size}")
    auto = run(m, max_num_iterations=6)
    per_spmv = run(m, partials=1, max_num_iterations=6)
    once = run(m, partials=2, max_num_iterations=6)
    print(f"{size} keyframes: LM log {auto[1]}", flush=True)
    assert max(r[2] for r in auto[1]) < 500
    for other in (per_spmv, once, run(m, max_num_iterations=6)):
        assert other[1] == auto[1]
        assert other[0].final_cost == auto[0].final_cost
        np.testing.assert_array_equal(other[2], auto[2])
        np.testing.assert_array_equal(other[3], auto[3])


def test_both_partial_sums_without_deflation(walk1000):
    # Block-Jacobi PCG: 3 partials per group (the per-SpMV sum by default).
    kw = dict(max_num_iterations=4, pcg_coarse_nodes=-1, pcg_ritz_vectors=-1, pcg_max_iterations=20000)
    per_spmv = run(walk1000, partials=1, **kw)
    once = run(walk1000, partials=2, **kw)
    assert once[1] == per_spmv[1]
    np.testing.assert_array_equal(once[2], per_spmv[2])
    np.testing.assert_array_equal(once[3], per_spmv[3])


def test_many_groups_match_oracle(oracle, walk1000):
    o = opts(max_num_iterations=2)
    sm, rows, v, _ = run(walk1000, max_num_iterations=2)
    ref, ref_rows, rv, _, _ = oracle.solve(walk1000, o, linear_solver=oracle.SCHUR_DIRECT, threads=16)
    print(f"1000 keyframes: costs {[r[0] for r in rows]} oracle {[r.cost for r in ref_rows]}", flush=True)
    assert sm.num_iterations == ref.num_iterations
    np.testing.assert_allclose([r[0] for r in rows], [r.cost for r in ref_rows], rtol=1e-8)
    np.testing.assert_allclose(v, rv, atol=1e-6)
