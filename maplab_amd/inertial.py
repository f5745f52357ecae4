"""Batched inertial error term (InertialErrorTerm::Evaluate,
algorithms/ceres-error-terms/src/inertial-error-term.cc:73-286) through the
library's C ABI (include/maplab_ba_inertial.h)."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import abi


@dataclass
class InertialEdges:
    vertices: np.ndarray          # (P, 7) [q_IM | p_MI]
    gyro_bias: np.ndarray         # (P, 3)
    velocity: np.ndarray          # (P, 3)
    accel_bias: np.ndarray        # (P, 3)
    edge_from: np.ndarray         # (E,)
    edge_to: np.ndarray           # (E,)
    edge_sample_begin: np.ndarray  # (E + 1,)
    imu: np.ndarray               # (S, 6) accelerometer, gyroscope
    timestamps_ns: np.ndarray     # (S,) int64
    noise: tuple = (1e-3, 1e-4, 1e-2, 1e-3)  # gyro noise, gyro bias, acc noise, acc bias sigmas
    gravity: float = 9.81

    def _arrays(self):
        f = lambda a, t: np.ascontiguousarray(a, dtype=t)
        return [f(self.vertices, np.float64), f(self.gyro_bias, np.float64), f(self.velocity, np.float64),
                f(self.accel_bias, np.float64), f(self.edge_from, np.int32), f(self.edge_to, np.int32),
                f(self.edge_sample_begin, np.int32), f(self.imu, np.float64), f(self.timestamps_ns, np.int64)]


def _problem(edges: InertialEdges):
    arrs = edges._arrays()
    p = abi.InertialProblem()
    p.num_vertices, p.num_edges, p.num_samples = arrs[0].shape[0], arrs[4].shape[0], arrs[8].shape[0]
    (p.vertex_q_IM__M_p_MI, p.vertex_gyro_bias, p.vertex_velocity_M, p.vertex_accel_bias, p.edge_from, p.edge_to,
     p.edge_sample_begin, p.imu, p.timestamps_ns) = [a.ctypes.data for a in arrs]
    p.gyro_noise_sigma, p.gyro_bias_sigma, p.acc_noise_sigma, p.acc_bias_sigma = edges.noise
    p.gravity_magnitude = edges.gravity
    return p, arrs


def solve_states(edges: InertialEdges, options: abi.Options, fix_gyro_bias: bool = False,
                 fix_accel_bias: bool = False, fix_velocity: bool = False, capacity: int = 256):
    """mlba_solve_inertial_states: the speed and bias states with the poses
    constant. Returns (summary, rows, gyro_bias, velocity, accel_bias)."""
    lib = abi.load_library()
    p, arrs = _problem(edges)
    io = abi.InertialOptions(int(fix_gyro_bias), int(fix_accel_bias), int(fix_velocity), 0)
    P = p.num_vertices
    bg, v, ba = np.zeros((P, 3)), np.zeros((P, 3)), np.zeros((P, 3))
    its = (abi.IterationSummary * capacity)()
    sm = abi.Summary()
    sm.iterations_capacity = capacity
    sm.iterations = ctypes.cast(its, ctypes.POINTER(abi.IterationSummary))
    rc = lib.mlba_solve_inertial_states(ctypes.byref(p), ctypes.byref(options), ctypes.byref(io), bg.ctypes.data,
                                        v.ctypes.data, ba.ctypes.data, ctypes.byref(sm))
    if rc != abi.MLBA_OK:
        raise RuntimeError(f"maplab_ba inertial error {rc}: {lib.mlba_inertial_last_error().decode()}")
    return sm, [its[i] for i in range(min(sm.num_iterations, capacity))], bg, v, ba


def evaluate(edges: InertialEdges, want_jacobians: bool = True):
    """Residuals (E, 15) and ambient Jacobians [8 x (E, 15, size)]."""
    lib = abi.load_library()
    arrs = edges._arrays()
    p = abi.InertialProblem()
    p.num_vertices, p.num_edges, p.num_samples = arrs[0].shape[0], arrs[4].shape[0], arrs[8].shape[0]
    (p.vertex_q_IM__M_p_MI, p.vertex_gyro_bias, p.vertex_velocity_M, p.vertex_accel_bias, p.edge_from, p.edge_to,
     p.edge_sample_begin, p.imu, p.timestamps_ns) = [a.ctypes.data for a in arrs]
    p.gyro_noise_sigma, p.gyro_bias_sigma, p.acc_noise_sigma, p.acc_bias_sigma = edges.noise
    p.gravity_magnitude = edges.gravity
    E = p.num_edges
    res = np.zeros((E, 15))
    jac = [np.zeros((E, 15, k)) for k in abi.INERTIAL_BLOCK_SIZES] if want_jacobians else None
    ptrs = (ctypes.c_void_p * 8)(*[j.ctypes.data for j in jac]) if jac else None
    rc = lib.mlba_inertial_evaluate(ctypes.byref(p), res.ctypes.data, ptrs)
    if rc != abi.MLBA_OK:
        raise RuntimeError(f"maplab_ba inertial error {rc}: {lib.mlba_inertial_last_error().decode()}")
    return res, jac
