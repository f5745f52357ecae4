"""The oracle's speed / bias solve with constant poses: a chain whose states
come from the oracle's own IMU integration has zero inertial residual, and
the LM recovers those states from perturbed velocities and biases."""
import numpy as np

import inertial_cases as ic
from maplab_amd import abi


def test_true_states_have_zero_residual_and_are_recovered(oracle):
    truth = ic.chain(oracle, 3)
    o = abi.default_options()
    o.max_num_iterations = 50
    sm, rows, bg, v, ba = oracle.solve_inertial_states(truth, o)
    assert sm.initial_cost < 1e-16
    start = ic.perturbed(truth, 4)
    sm, rows, bg, v, ba = oracle.solve_inertial_states(start, o)
    assert sm.initial_cost > 1e3 and sm.final_cost < 1e-12
    np.testing.assert_allclose(v, truth.velocity, atol=1e-6)
    np.testing.assert_allclose(bg, truth.gyro_bias, atol=1e-7)
    np.testing.assert_allclose(ba, truth.accel_bias, atol=1e-6)


def test_fixed_blocks_stay_put(oracle):
    start = ic.perturbed(ic.chain(oracle, 5), 6)
    o = abi.default_options()
    o.max_num_iterations = 10
    sm, rows, bg, v, ba = oracle.solve_inertial_states(start, o, fix_gyro_bias=True, fix_velocity=True)
    np.testing.assert_array_equal(bg, start.gyro_bias)
    np.testing.assert_array_equal(v, start.velocity)
    assert np.abs(ba - start.accel_bias).max() > 0
    assert sm.final_cost < sm.initial_cost
