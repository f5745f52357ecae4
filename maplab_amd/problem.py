"""Structure-of-arrays visual map: the state and topology that maplab's
OptimizationStateBuffer and ProblemInformation hold for the visual BA terms
(algorithms/map-optimization/src/optimization-state-buffer.cc:153-306,
optimization-terms-addition.cc:22-241), as numpy arrays.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import abi


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


@dataclass
class VisualMap:
    vertices: np.ndarray            # (P, 7) [q_IM xyzw | M_p_MI]
    baseframes: np.ndarray          # (M, 7) [q_GM | G_p_GM]
    landmarks: np.ndarray           # (L, 3) p_B in the store vertex frame
    camera_q_CI__C_p_CI: np.ndarray  # (C, 7)
    camera_intrinsics: np.ndarray   # (C, 4)
    camera_distortion: np.ndarray   # (C, 4)
    camera_distortion_type: np.ndarray  # (C,) int32
    camera_resolution: np.ndarray   # (C, 2) int32
    vertex_mission: np.ndarray      # (P,) int32
    landmark_store_vertex: np.ndarray  # (L,) int32
    obs_vertex: np.ndarray          # (N,) int32
    obs_landmark: np.ndarray        # (N,) int32
    obs_camera: np.ndarray          # (N,) int32
    obs_keypoint: np.ndarray        # (N, 2)
    obs_sigma: np.ndarray           # (N,)
    vertex_constant: np.ndarray     # (P,) uint8
    landmark_constant: np.ndarray   # (L,) uint8
    true_vertices: np.ndarray | None = None
    true_landmarks: np.ndarray | None = None
    camera_model: np.ndarray | None = None  # (C,) int32 MLBA_CAMERA_*, None: all pinhole
    camera_xi: np.ndarray | None = None     # (C,) unified-projection xi, None: 0
    _desc: abi.ProblemDesc | None = field(default=None, repr=False)

    @property
    def num_vertices(self) -> int:
        return self.vertices.shape[0]

    @property
    def num_landmarks(self) -> int:
        return self.landmarks.shape[0]

    @property
    def num_observations(self) -> int:
        return self.obs_vertex.shape[0]

    def normalized(self) -> "VisualMap":
        """Make every array contiguous with the dtype the ABI expects."""
        f = lambda a, t: np.ascontiguousarray(a, dtype=t)
        self.vertices = f(self.vertices, np.float64).reshape(-1, 7)
        self.baseframes = f(self.baseframes, np.float64).reshape(-1, 7)
        self.landmarks = f(self.landmarks, np.float64).reshape(-1, 3)
        self.camera_q_CI__C_p_CI = f(self.camera_q_CI__C_p_CI, np.float64).reshape(-1, 7)
        self.camera_intrinsics = f(self.camera_intrinsics, np.float64).reshape(-1, 4)
        self.camera_distortion = f(self.camera_distortion, np.float64).reshape(-1, 4)
        self.camera_distortion_type = f(self.camera_distortion_type, np.int32)
        self.camera_resolution = f(self.camera_resolution, np.int32).reshape(-1, 2)
        self.vertex_mission = f(self.vertex_mission, np.int32)
        self.landmark_store_vertex = f(self.landmark_store_vertex, np.int32)
        self.obs_vertex = f(self.obs_vertex, np.int32)
        self.obs_landmark = f(self.obs_landmark, np.int32)
        self.obs_camera = f(self.obs_camera, np.int32)
        self.obs_keypoint = f(self.obs_keypoint, np.float64).reshape(-1, 2)
        self.obs_sigma = f(self.obs_sigma, np.float64)
        self.vertex_constant = f(self.vertex_constant, np.uint8)
        self.landmark_constant = f(self.landmark_constant, np.uint8)
        if self.camera_model is not None:
            self.camera_model = f(self.camera_model, np.int32)
        if self.camera_xi is not None:
            self.camera_xi = f(self.camera_xi, np.float64)
        return self

    def desc(self) -> abi.ProblemDesc:
        """ProblemDesc pointing into this map's arrays (kept alive by self)."""
        self.normalized()
        d = abi.ProblemDesc()
        d.num_vertices = self.num_vertices
        d.num_landmarks = self.num_landmarks
        d.num_observations = self.num_observations
        d.num_cameras = self.camera_intrinsics.shape[0]
        d.num_missions = self.baseframes.shape[0]
        d.vertex_q_IM__M_p_MI = _ptr(self.vertices)
        d.baseframe_q_GM__G_p_GM = _ptr(self.baseframes)
        d.landmark_p_B = _ptr(self.landmarks)
        d.camera_q_CI__C_p_CI = _ptr(self.camera_q_CI__C_p_CI)
        d.camera_intrinsics = _ptr(self.camera_intrinsics)
        d.camera_distortion = _ptr(self.camera_distortion)
        d.camera_distortion_type = _ptr(self.camera_distortion_type)
        d.camera_resolution = _ptr(self.camera_resolution)
        d.vertex_mission = _ptr(self.vertex_mission)
        d.landmark_store_vertex = _ptr(self.landmark_store_vertex)
        d.obs_vertex = _ptr(self.obs_vertex)
        d.obs_landmark = _ptr(self.obs_landmark)
        d.obs_camera = _ptr(self.obs_camera)
        d.obs_keypoint = _ptr(self.obs_keypoint)
        d.obs_sigma = _ptr(self.obs_sigma)
        d.vertex_constant = _ptr(self.vertex_constant)
        d.landmark_constant = _ptr(self.landmark_constant)
        d.camera_model = _ptr(self.camera_model)
        d.camera_xi = _ptr(self.camera_xi)
        self._desc = d
        return d

    def copy(self) -> "VisualMap":
        kw = {k: (v.copy() if isinstance(v, np.ndarray) else v)
              for k, v in self.__dict__.items() if k != "_desc"}
        return VisualMap(**kw)


def synth_config(workload: int, **overrides) -> abi.SynthConfig:
    """Configuration of a benchmark workload (1: circle, 2: 500 keyframes,
    3: 2000 keyframes) with field overrides."""
    lib = abi.load_synth()
    cfg = abi.SynthConfig()
    abi.check(lib.mlba_synth_config_init(ctypes.byref(cfg), workload))
    if workload in (2, 3) and "detection_probability" not in overrides:
        cfg.detection_probability = 0.25
    for k, v in overrides.items():
        if k in ("distortion", "intrinsics", "p_C_I"):
            arr = getattr(cfg, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(cfg, k, v)
    return cfg


def generate(cfg: abi.SynthConfig) -> VisualMap:
    """Run the seeded generator and copy its arrays into a VisualMap."""
    lib = abi.load_synth()
    handle = ctypes.c_void_p()
    abi.check(lib.mlba_synth_generate(ctypes.byref(cfg), ctypes.byref(handle)))
    try:
        d = abi.ProblemDesc()
        abi.check(lib.mlba_synth_describe(handle, ctypes.byref(d)))
        P, L, N, C, M = d.num_vertices, d.num_landmarks, d.num_observations, d.num_cameras, d.num_missions

        def arr(ptr, n, dtype):
            if n == 0:
                return np.zeros(0, dtype=dtype)
            ct = {np.float64: ctypes.c_double, np.int32: ctypes.c_int32, np.uint8: ctypes.c_uint8}[dtype]
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ct)), (n,)).copy()

        tv = arr(ctypes.cast(lib.mlba_synth_true_vertices(handle), ctypes.c_void_p).value, 7 * P, np.float64)
        tl = arr(ctypes.cast(lib.mlba_synth_true_landmarks(handle), ctypes.c_void_p).value, 3 * L, np.float64)
        m = VisualMap(
            vertices=arr(d.vertex_q_IM__M_p_MI, 7 * P, np.float64).reshape(P, 7),
            baseframes=arr(d.baseframe_q_GM__G_p_GM, 7 * M, np.float64).reshape(M, 7),
            landmarks=arr(d.landmark_p_B, 3 * L, np.float64).reshape(L, 3),
            camera_q_CI__C_p_CI=arr(d.camera_q_CI__C_p_CI, 7 * C, np.float64).reshape(C, 7),
            camera_intrinsics=arr(d.camera_intrinsics, 4 * C, np.float64).reshape(C, 4),
            camera_distortion=arr(d.camera_distortion, 4 * C, np.float64).reshape(C, 4),
            camera_distortion_type=arr(d.camera_distortion_type, C, np.int32),
            camera_resolution=arr(d.camera_resolution, 2 * C, np.int32).reshape(C, 2),
            vertex_mission=arr(d.vertex_mission, P, np.int32),
            landmark_store_vertex=arr(d.landmark_store_vertex, L, np.int32),
            obs_vertex=arr(d.obs_vertex, N, np.int32),
            obs_landmark=arr(d.obs_landmark, N, np.int32),
            obs_camera=arr(d.obs_camera, N, np.int32),
            obs_keypoint=arr(d.obs_keypoint, 2 * N, np.float64).reshape(N, 2),
            obs_sigma=arr(d.obs_sigma, N, np.float64),
            vertex_constant=arr(d.vertex_constant, P, np.uint8),
            landmark_constant=arr(d.landmark_constant, L, np.uint8),
            true_vertices=tv.reshape(P, 7),
            true_landmarks=tl.reshape(L, 3),
        )
        return m.normalized()
    finally:
        lib.mlba_synth_free(handle)


def workload_map(workload: int, seed: int = 1, **overrides) -> VisualMap:
    return generate(synth_config(workload, seed=seed, **overrides))
