"""The speed / bias solve with constant poses on the device path
(mlba_solve_inertial_states: device inertial residuals and Jacobians, host
LM on the block-tridiagonal normal equations) against the oracle's dense
LM, and the recovery of integrated true states."""
import numpy as np
import pytest

import inertial_cases as ic
from maplab_amd import abi, inertial

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed, consistent, fix", [(3, True, {}), (7, False, {}),
                                                   (8, False, dict(fix_gyro_bias=True)),
                                                   (9, True, dict(fix_velocity=True, fix_accel_bias=True))])
def test_states_log_matches_oracle(oracle, seed, consistent, fix):
    start = ic.perturbed(ic.chain(oracle, seed, n_vertices=10, consistent=consistent), seed + 100)
    o = abi.default_options()
    o.max_num_iterations = 20
    sm, rows, bg, v, ba = inertial.solve_states(start, o, **fix)
    ref, ref_rows, rbg, rv, rba = oracle.solve_inertial_states(start, o, **fix)
    print(f"inertial states seed {seed}: {sm.num_iterations} rows, cost {sm.initial_cost:.6e} -> "
          f"{sm.final_cost:.6e} (oracle {ref.final_cost:.6e})", flush=True)
    assert sm.num_iterations == ref.num_iterations
    assert sm.termination_type == ref.termination_type
    # Costs of a consistent chain fall to ~1e-10 from ~1e5: past 1e-9
    # relative, the comparison is absolute at the rounding of the start cost.
    np.testing.assert_allclose([r.cost for r in rows], [r.cost for r in ref_rows], rtol=1e-9,
                               atol=1e-14 * ref_rows[0].cost)
    np.testing.assert_allclose(v, rv, rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(bg, rbg, rtol=1e-8, atol=1e-12)
    np.testing.assert_allclose(ba, rba, rtol=1e-8, atol=1e-12)


def test_integrated_states_are_recovered(oracle):
    truth = ic.chain(oracle, 11, n_vertices=12)
    o = abi.default_options()
    o.max_num_iterations = 50
    sm, rows, bg, v, ba = inertial.solve_states(ic.perturbed(truth, 12), o)
    assert sm.final_cost < 1e-12
    np.testing.assert_allclose(v, truth.velocity, atol=1e-6)
    np.testing.assert_allclose(bg, truth.gyro_bias, atol=1e-7)
    np.testing.assert_allclose(ba, truth.accel_bias, atol=1e-6)


def test_non_chain_edges_are_rejected(oracle):
    ed = ic.chain(oracle, 13, n_vertices=5)
    ed.edge_to = ed.edge_to.copy()
    ed.edge_to[1] = 3
    with pytest.raises(RuntimeError, match="consecutive"):
        inertial.solve_states(ed, abi.default_options())
