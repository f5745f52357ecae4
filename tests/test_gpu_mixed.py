"""FP32 Jacobian storage (mlba_options.jacobian_fp32): the LM reaches the
FP64 oracle's final cost within the 1e-6 relative bar after the same
iterations, at a parity size and at full size, and repeats bit for bit."""
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


@pytest.mark.parametrize("distortion", [1, 2])
def test_fp32_jacobians_reach_the_oracle_cost(oracle, distortion):
    kw = {} if distortion == 1 else dict(distortion_type=2, distortion=(-0.01, 0.02, -0.005, 0.001))
    m = problem.workload_map(2, seed=2, num_keyframes=60, num_landmarks=12000, **kw)
    o = opts(max_num_iterations=10, jacobian_fp32=1)
    with DeviceProblem(m) as dev:
        sm, rows = dev.solve(o)
        v1, l1 = dev.download_state()
        dev.upload_state(m.vertices, m.landmarks)
        sm2, rows2 = dev.solve(o)
        v2, l2 = dev.download_state()
    ref, ref_rows, _, _, _ = oracle.solve(m, opts(max_num_iterations=10), linear_solver=oracle.SCHUR_DIRECT,
                                          threads=4)
    rel = abs(sm.final_cost - ref.final_cost) / ref.final_cost
    print(f"fp32 Jacobians ({distortion}): final {sm.final_cost:.12e} oracle {ref.final_cost:.12e} rel {rel:.2e}",
          flush=True)
    assert rel <= 1e-6
    assert sm.initial_cost == ref.initial_cost or abs(sm.initial_cost - ref.initial_cost) <= 1e-12 * ref.initial_cost
    np.testing.assert_array_equal(v1, v2)
    np.testing.assert_array_equal(l1, l2)
    assert [r.cost for r in rows] == [r.cost for r in rows2]


def test_fp32_jacobians_full_size_cost_tracks_fp64():
    m = problem.workload_map(2, seed=1)
    print(f"full size fp32 Jacobians: N {m.num_observations}", flush=True)
    with DeviceProblem(m) as dev:
        sm64, rows64 = dev.solve(opts(max_num_iterations=6))
        dev.upload_state(m.vertices, m.landmarks)
        sm32, rows32 = dev.solve(opts(max_num_iterations=6, jacobian_fp32=1))
    c64 = np.array([r.cost for r in rows64])
    c32 = np.array([r.cost for r in rows32])
    print("  rel per iteration", np.abs(c32 - c64) / c64, flush=True)
    assert len(c32) == len(c64)
    assert np.all(np.abs(c32 - c64) <= 1e-6 * c64)
