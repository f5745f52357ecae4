"""map_optimization::VIMapOptimizer for the visual terms
(algorithms/map-optimization/src/vi-map-optimizer.cc:34-80), through the
library's optimizer entry points (include/maplab_ba_optimizer.h): landmark
admission on the device, problem selection and gauge fixes on the host, the
solve (plain or with outlier rejection) on the device, and the copy back of
the optimized state.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import abi
from .problem import VisualMap


def default_vi_problem_options() -> abi.ViProblemOptions:
    """ViProblemOptions::initFromGFlags defaults, visual terms
    (vi-optimization-builder.cc:98-150)."""
    o = abi.ViProblemOptions()
    abi.load_library().mlba_vi_problem_options_init(ctypes.byref(o))
    return o


def _check(rc: int) -> None:
    if rc != abi.MLBA_OK:
        raise RuntimeError(f"maplab_ba optimizer error {rc}: {abi.load_library().mlba_optimizer_last_error().decode()}")


def _mask(a, n):
    return None if a is None else np.ascontiguousarray(a, dtype=np.uint8).reshape(n)


@dataclass
class ProblemSelection:
    obs_in_problem: np.ndarray
    vertex_in_problem: np.ndarray
    landmark_in_problem: np.ndarray
    vertex_constant: np.ndarray
    landmark_constant: np.ndarray
    gauge_vertices: np.ndarray
    gauge_landmarks: np.ndarray


def select_problem(vmap: VisualMap, landmark_good, options: abi.ViProblemOptions | None = None,
                   missions=None) -> ProblemSelection:
    """constructOptimizationProblem's choice of observations, blocks in the
    problem and constant blocks (host only)."""
    lib = abi.load_library()
    o = options if options is not None else default_vi_problem_options()
    d = vmap.desc()
    P, L, N, M = vmap.num_vertices, vmap.num_landmarks, vmap.num_observations, vmap.baseframes.shape[0]
    arrs = {k: np.zeros(n, dtype=np.uint8) for k, n in (("obs_in_problem", N), ("vertex_in_problem", P),
                                                       ("landmark_in_problem", L), ("vertex_constant", P),
                                                       ("landmark_constant", L))}
    gv, gl = np.zeros(max(M, 1), dtype=np.int32), np.zeros(max(M, 1), dtype=np.int32)
    sel = abi.ProblemSelection()
    for k, a in arrs.items():
        setattr(sel, k, a.ctypes.data)
    sel.gauge_vertex, sel.gauge_landmark = gv.ctypes.data, gl.ctypes.data
    good = _mask(landmark_good, L)
    ms = _mask(missions, M)
    _check(lib.mlba_select_problem(ctypes.byref(d), ctypes.byref(o), None if ms is None else ms.ctypes.data,
                                   good.ctypes.data, ctypes.byref(sel)))
    return ProblemSelection(**{k: a.astype(bool) for k, a in arrs.items()},
                            gauge_vertices=gv[:sel.num_clusters].copy(), gauge_landmarks=gl[:sel.num_clusters].copy())


@dataclass
class OptimizationResult:
    summary: abi.Summary
    iterations: list
    vertices: np.ndarray
    landmarks: np.ndarray
    landmark_quality: np.ndarray
    cameras: tuple | None = None  # (q_CI | C_p_CI, intrinsics, distortion, xi) after the solve


class VIMapOptimizer:
    """VIMapOptimizer::optimize on a structure-of-arrays map."""

    def __init__(self, options: abi.ViProblemOptions | None = None):
        self.options = options if options is not None else default_vi_problem_options()

    def optimize(self, vmap: VisualMap, missions=None, landmark_quality=None,
                 capacity: int = 1024) -> OptimizationResult:
        lib = abi.load_library()
        d = vmap.desc()
        P, L, M = vmap.num_vertices, vmap.num_landmarks, vmap.baseframes.shape[0]
        q = (np.zeros(L, dtype=np.uint8) if landmark_quality is None
             else np.ascontiguousarray(landmark_quality, dtype=np.uint8).copy())
        ms = _mask(missions, M)
        v_out = np.empty((P, 7))
        l_out = np.empty((L, 3))
        its = (abi.IterationSummary * capacity)()
        sm = abi.Summary()
        sm.iterations_capacity = capacity
        sm.iterations = ctypes.cast(its, ctypes.POINTER(abi.IterationSummary))
        C = vmap.camera_intrinsics.shape[0]
        cams = (np.empty((C, 7)), np.empty((C, 4)), np.empty((C, 4)), np.empty(C))
        _check(lib.mlba_vi_map_optimize_cameras(ctypes.byref(d), ctypes.byref(self.options),
                                                None if ms is None else ms.ctypes.data, q.ctypes.data,
                                                v_out.ctypes.data, l_out.ctypes.data,
                                                *[a.ctypes.data for a in cams], ctypes.byref(sm)))
        rows = [its[i] for i in range(min(sm.num_iterations, capacity))]
        return OptimizationResult(sm, rows, v_out, l_out, q, cams)
