"""The reduced camera system the device builds (16 x 16-block tiles staged
through LDS, fixed-order parts per block, maplab_amd/csrc/ba_tiles.h)
against the oracle's Schur complement of the same linearization
(oracle_reduced_camera_system: dense normal equations, landmark blocks
eliminated one by one) entry by entry: S and its right-hand side in the
Jacobi-scaled variables of an LM iteration, on maps with two-camera rows
(two incidences of a landmark in one row), unified-projection cameras,
landmarks over several linearization tiles, and enough keyframes for many
tiles and segments. Two exports of the same state are identical to the
last bit."""
import numpy as np
import pytest

from cases import tiny_map
from maplab_amd import abi, learned, problem
from maplab_amd.solver import DeviceProblem

pytestmark = pytest.mark.gpu


def opts(**kw):
    o = abi.default_options()
    for k, v in kw.items():
        setattr(o, k, v)
    return o


CASES = {
    "two_cameras_global_missions": lambda: tiny_map(np.random.default_rng(78), dist_type=2, n_vertices=12,
                                                    n_landmarks=150, global_missions=True, n_cameras=2),
    "unified_projection": lambda: tiny_map(np.random.default_rng(5), dist_type=1, n_vertices=9, n_landmarks=120,
                                           xi=0.9),
    "landmarks_over_three_linearization_tiles": lambda: tiny_map(np.random.default_rng(78), dist_type=1,
                                                                 n_vertices=600, n_landmarks=6),
    "walk_90_keyframes": lambda: problem.workload_map(2, seed=3, num_keyframes=90, num_landmarks=9000),
    "workload1": lambda: problem.workload_map(1, seed=1),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("radius", [1e4, 3.0])
def test_reduced_camera_system_matches_oracle(oracle, case, radius):
    m = CASES[case]()
    o = opts()
    print(f"{case}: P {m.num_vertices} L {m.num_landmarks} N {m.num_observations} radius {radius}", flush=True)
    with DeviceProblem(m) as dev:
        S, b = learned.rcs_system(dev, o, radius)
        S2, b2 = learned.rcs_system(dev, o, radius)
    np.testing.assert_array_equal(S, S2)
    np.testing.assert_array_equal(b, b2)
    # S_ji is written as the transpose of S_ij; a diagonal block's two
    # triangles are summed separately, so they agree to rounding.
    n = S.shape[0] // 6
    off = np.kron(1 - np.eye(n), np.ones((6, 6))).astype(bool)
    np.testing.assert_array_equal(S[off], S.T[off])
    assert np.abs(S - S.T).max() <= 1e-14 * np.abs(S).max()
    ref_S, ref_b, _ = oracle.reduced_camera_system(m, o, radius)
    assert S.shape == ref_S.shape
    scale = np.abs(ref_S).max()
    err = np.abs(S - ref_S).max()
    print(f"  max |S| {scale:.3e}, max |dS| {err:.3e}, max |db| {np.abs(b - ref_b).max():.3e}", flush=True)
    assert err <= 1e-12 * scale
    # Structure: every block the oracle has is a block of the device's S.
    nz_ref = np.abs(ref_S.reshape(n, 6, n, 6)).max(axis=(1, 3)) > 0
    nz_dev = np.abs(S.reshape(n, 6, n, 6)).max(axis=(1, 3)) > 0
    assert np.all(nz_dev[nz_ref])
    np.testing.assert_allclose(b, ref_b, rtol=0, atol=1e-12 * np.abs(ref_b).max())
