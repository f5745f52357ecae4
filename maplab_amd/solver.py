"""Device-resident visual BA problem: a thin owner of an mlba_handle.

This is the Python side of the C ABI (include/maplab_ba.h); the maplab-shaped
driver on top of it is maplab_amd.optimizer.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import abi
from .problem import VisualMap


class DeviceProblem:
    """The map copied into HBM with its observation and Schur topology."""

    def __init__(self, vmap: VisualMap):
        self.lib = abi.load_library()
        self.map = vmap
        self._desc = vmap.desc()
        h = ctypes.c_void_p()
        abi.check(self.lib.mlba_create(ctypes.byref(self._desc), ctypes.byref(h)))
        self.handle = h

    def close(self) -> None:
        if self.handle:
            self.lib.mlba_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> None:
        abi.check(rc, self.handle)

    def upload_state(self, vertices=None, landmarks=None) -> None:
        v = None if vertices is None else np.ascontiguousarray(vertices, dtype=np.float64)
        l = None if landmarks is None else np.ascontiguousarray(landmarks, dtype=np.float64)
        self._check(self.lib.mlba_upload_state(self.handle, None if v is None else v.ctypes.data,
                                               None if l is None else l.ctypes.data))

    def download_state(self):
        v = np.empty((self.map.num_vertices, 7))
        l = np.empty((self.map.num_landmarks, 3))
        self._check(self.lib.mlba_download_state(self.handle, v.ctypes.data, l.ctypes.data))
        return v, l

    def set_landmark_active(self, mask) -> None:
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        self._check(self.lib.mlba_set_landmark_active(self.handle, m.ctypes.data))

    def landmark_active(self) -> np.ndarray:
        m = np.empty(self.map.num_landmarks, dtype=np.uint8)
        self._check(self.lib.mlba_get_landmark_active(self.handle, m.ctypes.data))
        return m

    def landmark_quality(self, options: abi.QualityOptions | None = None) -> np.ndarray:
        """isVisualLandmarkWellConstrained of every landmark at the current
        device state (bool array, caller landmark order)."""
        o = options if options is not None else abi.default_quality_options()
        good = np.empty(self.map.num_landmarks, dtype=np.uint8)
        self._check(self.lib.mlba_landmark_quality(self.handle, ctypes.byref(o), good.ctypes.data))
        return good.astype(bool)

    def evaluate(self, tangent: bool = False, blocks=range(9), options: abi.Options | None = None,
                 want_cost: bool = True, with_xi: bool = False, noalloc: bool = False):
        """Batched VisualReprojectionError::Evaluate: residuals (N, 2), a dict
        block -> (N, 2, size) Jacobians, and the robust cost. With with_xi,
        block 9 (N, 2, 1) is the unified-projection xi column. With noalloc,
        through mlba_evaluate_noalloc (fixed staging array, no allocation)."""
        n = self.map.num_observations
        sizes = (abi.TANGENT_BLOCK_SIZES if tangent else abi.AMBIENT_BLOCK_SIZES) + (1,)
        nb = 10 if with_xi else 9
        res = np.empty((n, 2))
        jac = {b: np.empty((n, 2, sizes[b])) for b in blocks if b < nb}
        ptrs = (ctypes.c_void_p * nb)(*[jac[b].ctypes.data if b in jac else None for b in range(nb)])
        cost = ctypes.c_double(0.0)
        opt = options if options is not None else abi.default_options()
        flags = (abi.EVAL_TANGENT if tangent else abi.EVAL_AMBIENT) | (abi.EVAL_WITH_XI if with_xi else 0)
        fn = self.lib.mlba_evaluate_noalloc if noalloc else self.lib.mlba_evaluate
        self._check(fn(self.handle, flags, ctypes.byref(opt), res.ctypes.data, ptrs,
                       ctypes.byref(cost) if want_cost else None))
        return res, jac, cost.value

    def solve(self, options: abi.Options, outlier_rejection: bool = False, capacity: int = 256):
        """ceres::Solve (or solveWithOutlierRejection) on the device state."""
        its = (abi.IterationSummary * capacity)()
        sm = abi.Summary()
        sm.iterations_capacity = capacity
        sm.iterations = ctypes.cast(its, ctypes.POINTER(abi.IterationSummary))
        fn = self.lib.mlba_solve_with_outlier_rejection if outlier_rejection else self.lib.mlba_solve
        self._check(fn(self.handle, ctypes.byref(options), ctypes.byref(sm)))
        rows = [its[i] for i in range(min(sm.num_iterations, capacity))]
        return sm, rows

    @staticmethod
    def solve_many(problems, options, threads: int = 0, outlier_rejection: bool = False, capacity: int = 256):
        """mlba_solve_many: the problems (DeviceProblem, distinct) solved
        concurrently, problem i with options[i] (or one Options for all).
        Returns [(summary, rows)] in problem order; raises on a failed solve."""
        n = len(problems)
        if isinstance(options, abi.Options):
            options = [options] * n
        lib = abi.load_library()
        its = [(abi.IterationSummary * capacity)() for _ in range(n)]
        sms = [abi.Summary() for _ in range(n)]
        for sm, it in zip(sms, its):
            sm.iterations_capacity = capacity
            sm.iterations = ctypes.cast(it, ctypes.POINTER(abi.IterationSummary))
        handles = (ctypes.c_void_p * n)(*[p.handle for p in problems])
        opts = (ctypes.POINTER(abi.Options) * n)(*[ctypes.pointer(o) for o in options])
        sums = (ctypes.POINTER(abi.Summary) * n)(*[ctypes.pointer(sm) for sm in sms])
        statuses = (ctypes.c_int32 * n)()
        rc = lib.mlba_solve_many(handles, opts, sums, statuses, n, threads, 1 if outlier_rejection else 0)
        for p, st in zip(problems, statuses):
            p._check(st)
        if rc:
            raise RuntimeError(f"mlba_solve_many failed: {rc}")
        return [(sm, [it[i] for i in range(min(sm.num_iterations, capacity))]) for sm, it in zip(sms, its)]

    def solve_calibration(self, options: abi.Options, fix_intrinsics: bool = True, fix_distortion: bool = True,
                          fix_rotation: bool = True, fix_translation: bool = True, capacity: int = 256):
        """mlba_solve_calibration: the free camera blocks refined with the map
        held constant. Returns (summary, iteration rows)."""
        co = abi.CalibrationOptions(int(fix_intrinsics), int(fix_distortion), int(fix_rotation),
                                    int(fix_translation))
        its = (abi.IterationSummary * capacity)()
        sm = abi.Summary()
        sm.iterations_capacity = capacity
        sm.iterations = ctypes.cast(its, ctypes.POINTER(abi.IterationSummary))
        self._check(self.lib.mlba_solve_calibration(self.handle, ctypes.byref(options), ctypes.byref(co),
                                                    ctypes.byref(sm)))
        return sm, [its[i] for i in range(min(sm.num_iterations, capacity))]

    def download_cameras(self):
        """(q_CI | C_p_CI (C, 7), intrinsics (C, 4), distortion (C, 4), xi (C,))."""
        C = self.map.camera_intrinsics.shape[0]
        q, i, d, x = np.zeros((C, 7)), np.zeros((C, 4)), np.zeros((C, 4)), np.zeros(C)
        self._check(self.lib.mlba_download_cameras(self.handle, q.ctypes.data, i.ctypes.data, d.ctypes.data,
                                                   x.ctypes.data))
        return q, i, d, x

    def autotune(self, options: abi.Options, repeats: int = 5) -> abi.Schedule:
        """mlba_autotune: time every dispatch schedule of the Schur block
        kernel at the current state and keep the fastest (results unchanged)."""
        sc = abi.Schedule()
        self._check(self.lib.mlba_autotune(self.handle, ctypes.byref(options), repeats, ctypes.byref(sc)))
        return sc

    def set_schedule(self, band: int, waves: int, xcd: int, pcg_partials: int = 0) -> None:
        sc = abi.Schedule(item_band_rows=band, item_waves=waves, xcd_mapping=xcd, pcg_partials=pcg_partials)
        self._check(self.lib.mlba_set_schedule(self.handle, ctypes.byref(sc)))

    def schedule(self) -> abi.Schedule:
        sc = abi.Schedule()
        self._check(self.lib.mlba_get_schedule(self.handle, ctypes.byref(sc)))
        return sc

    def time_kernels(self, options: abi.Options, repeats: int = 10) -> abi.KernelTimes:
        """Average device time of each kernel of one LM iteration (HIP events
        on the library's stream)."""
        t = abi.KernelTimes()
        self._check(self.lib.mlba_time_kernels(self.handle, ctypes.byref(options), repeats, ctypes.byref(t)))
        return t
