"""TEST INFRASTRUCTURE ONLY: ctypes binding of the FP64 CPU oracle
(oracle/liboracle_ba.so). Imported by tests/, the smoke check and the
cpu_baseline leg of bench.py; never by maplab_amd/.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(_HERE, "liboracle_ba.so")

SCHUR_DIRECT, DENSE_NORMAL, SCHUR_PCG = 0, 1, 2

_lib = None


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", _HERE])


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_PATH):
        build()
    from maplab_amd import abi  # structures only: the ABI data layout
    lib = ctypes.CDLL(ORACLE_PATH)
    vp, i32, f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double
    P = ctypes.POINTER
    sig = {
        "oracle_project": (i32, [i32, vp, vp, i32, i32, vp, vp, vp, vp, vp]),
        "oracle_project_camera": (i32, [i32, f64, i32, vp, vp, i32, i32, vp, vp, vp, vp, vp, vp]),
        "oracle_distort": (None, [i32, vp, vp, vp]),
        "oracle_quat_plus": (None, [vp, vp, vp]),
        "oracle_quat_jacobian": (None, [vp, vp]),
        "oracle_rotation_jpl": (None, [vp, vp]),
        "oracle_evaluate_visual": (i32, [vp, f64, i32, i32, i32, i32, vp, vp, vp]),
        "oracle_loss": (None, [i32, f64, f64, vp]),
        "oracle_problem_evaluate": (i32, [P(abi.ProblemDesc), vp, i32, P(abi.Options), vp, vp, vp]),
        "oracle_solve": (i32, [P(abi.ProblemDesc), P(abi.Options), i32, i32, vp, vp, vp, P(abi.Summary)]),
        "oracle_solve_with_outlier_rejection": (i32, [P(abi.ProblemDesc), P(abi.Options), i32, i32, vp, vp, vp,
                                                      P(abi.Summary)]),
        "oracle_solve_calibration": (i32, [P(abi.ProblemDesc), P(abi.Options), vp, vp, vp, vp, vp, vp,
                                           P(abi.Summary)]),
        "oracle_reduced_camera_system": (i32, [P(abi.ProblemDesc), P(abi.Options), f64, i32, vp, vp, vp]),
        "oracle_landmark_quality": (i32, [P(abi.ProblemDesc), P(abi.QualityOptions), vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data


def project(dist_type, intrinsics, dist, width, height, p_C, want_jac=True):
    lib = load()
    intr = np.ascontiguousarray(intrinsics, dtype=np.float64)
    d = np.ascontiguousarray(dist if dist is not None else np.zeros(4), dtype=np.float64)
    p = np.ascontiguousarray(p_C, dtype=np.float64)
    kp = np.zeros(2)
    jp, ji, jd = np.zeros((2, 3)), np.zeros((2, 4)), np.zeros((2, 4))
    st = lib.oracle_project(dist_type, _p(intr), _p(d), width, height, _p(p), _p(kp),
                            _p(jp) if want_jac else None, _p(ji) if want_jac else None,
                            _p(jd) if want_jac else None)
    return st, kp, jp, ji, jd


def project_camera(model, xi, dist_type, intrinsics, dist, width, height, p_C, want_jac=True):
    """(status, kp, d kp/d p_C, d kp/d (fu fv cu cv), d kp/d distortion, d kp/d xi)."""
    lib = load()
    intr = np.ascontiguousarray(intrinsics, dtype=np.float64)
    d = np.ascontiguousarray(dist if dist is not None else np.zeros(4), dtype=np.float64)
    p = np.ascontiguousarray(p_C, dtype=np.float64)
    kp = np.zeros(2)
    jp, ji, jd, jx = np.zeros((2, 3)), np.zeros((2, 4)), np.zeros((2, 4)), np.zeros(2)
    w = want_jac
    st = lib.oracle_project_camera(model, xi, dist_type, _p(intr), _p(d), width, height, _p(p), _p(kp),
                                   _p(jp) if w else None, _p(ji) if w else None, _p(jd) if w else None,
                                   _p(jx) if w else None)
    return st, kp, jp, ji, jd, jx


def distort(dist_type, dist, point):
    lib = load()
    pt = np.ascontiguousarray(point, dtype=np.float64).copy()
    jd = np.zeros((2, 2))
    lib.oracle_distort(dist_type, _p(np.ascontiguousarray(dist, dtype=np.float64)), _p(pt), _p(jd))
    return pt, jd


def quat_plus(q, delta):
    lib = load()
    out = np.zeros(4)
    lib.oracle_quat_plus(_p(np.ascontiguousarray(q, dtype=np.float64)),
                         _p(np.ascontiguousarray(delta, dtype=np.float64)), _p(out))
    return out


def quat_jacobian(q):
    lib = load()
    out = np.zeros((4, 3))
    lib.oracle_quat_jacobian(_p(np.ascontiguousarray(q, dtype=np.float64)), _p(out))
    return out


def rotation_jpl(q):
    lib = load()
    out = np.zeros((3, 3))
    lib.oracle_rotation_jpl(_p(np.ascontiguousarray(q, dtype=np.float64)), _p(out))
    return out


AMBIENT = (3, 7, 7, 7, 7, 4, 3, 4, 4)


def evaluate_visual(meas, sigma, error_type, dist_type, width, height, params):
    """params: list of 9 arrays in the reference block order. Returns
    (failed, residual, [9 ambient Jacobians])."""
    lib = load()
    arrs = [np.ascontiguousarray(p, dtype=np.float64) for p in params]
    pp = (ctypes.c_void_p * 9)(*[a.ctypes.data for a in arrs])
    res = np.zeros(2)
    jac = [np.zeros((2, AMBIENT[b])) for b in range(9)]
    jp = (ctypes.c_void_p * 9)(*[j.ctypes.data for j in jac])
    failed = lib.oracle_evaluate_visual(_p(np.ascontiguousarray(meas, dtype=np.float64)), sigma, error_type,
                                        dist_type, width, height, pp, _p(res), jp)
    return bool(failed), res, jac


def loss(loss_type, a, s):
    lib = load()
    rho = np.zeros(3)
    lib.oracle_loss(loss_type, a, s, _p(rho))
    return rho


def problem_evaluate(vmap, tangent=False, options=None, active=None, with_xi=False):
    from maplab_amd import abi
    lib = load()
    d = vmap.desc()
    n = vmap.num_observations
    sizes = (abi.TANGENT_BLOCK_SIZES if tangent else abi.AMBIENT_BLOCK_SIZES) + (1,)
    nb = 10 if with_xi else 9
    res = np.zeros((n, 2))
    jac = {b: np.zeros((n, 2, sizes[b])) for b in range(nb)}
    jp = (ctypes.c_void_p * nb)(*[jac[b].ctypes.data for b in range(nb)])
    cost = ctypes.c_double(0.0)
    act = None if active is None else np.ascontiguousarray(active, dtype=np.uint8)
    flags = (1 if tangent else 0) | (abi.EVAL_WITH_XI if with_xi else 0)
    rc = lib.oracle_problem_evaluate(ctypes.byref(d), _p(act), flags,
                                     ctypes.byref(options) if options is not None else None,
                                     _p(res), jp, ctypes.byref(cost))
    assert rc == 0, rc
    return res, jac, cost.value


def solve(vmap, options, linear_solver=SCHUR_DIRECT, threads=0, outlier_rejection=False, capacity=256,
          active=None):
    from maplab_amd import abi
    lib = load()
    d = vmap.desc()
    v = np.zeros((vmap.num_vertices, 7))
    l = np.zeros((vmap.num_landmarks, 3))
    its = (abi.IterationSummary * capacity)()
    sm = abi.Summary()
    sm.iterations_capacity = capacity
    sm.iterations = ctypes.cast(its, ctypes.POINTER(abi.IterationSummary))
    if outlier_rejection:
        act_out = np.zeros(vmap.num_landmarks, dtype=np.uint8)
        rc = lib.oracle_solve_with_outlier_rejection(ctypes.byref(d), ctypes.byref(options), linear_solver, threads,
                                                     _p(v), _p(l), _p(act_out), ctypes.byref(sm))
    else:
        act_out = None
        act = None if active is None else np.ascontiguousarray(active, dtype=np.uint8)
        rc = lib.oracle_solve(ctypes.byref(d), ctypes.byref(options), linear_solver, threads, _p(act), _p(v), _p(l),
                              ctypes.byref(sm))
    assert rc == 0, rc
    rows = [its[i] for i in range(min(sm.num_iterations, capacity))]
    return sm, rows, v, l, act_out


def solve_calibration(vmap, options, fix_intrinsics=True, fix_distortion=True, fix_rotation=True,
                      fix_translation=True, capacity=256):
    """Returns (summary, rows, (q_CI | C_p_CI, intrinsics, distortion, xi)) or
    None when the oracle rejects the problem."""
    from maplab_amd import abi
    lib = load()
    d = vmap.desc()
    C = vmap.camera_intrinsics.shape[0]
    q, i, dd, x = np.zeros((C, 7)), np.zeros((C, 4)), np.zeros((C, 4)), np.zeros(C)
    fix = np.array([fix_intrinsics, fix_distortion, fix_rotation, fix_translation], dtype=np.int32)
    its = (abi.IterationSummary * capacity)()
    sm = abi.Summary()
    sm.iterations_capacity = capacity
    sm.iterations = ctypes.cast(its, ctypes.POINTER(abi.IterationSummary))
    rc = lib.oracle_solve_calibration(ctypes.byref(d), ctypes.byref(options), _p(fix), None, _p(q), _p(i), _p(dd),
                                      _p(x), ctypes.byref(sm))
    if rc != 0:
        return None
    return sm, [its[k] for k in range(min(sm.num_iterations, capacity))], (q, i, dd, x)


def reduced_camera_system(vmap, options, mu):
    lib = load()
    d = vmap.desc()
    nmax = 6 * vmap.num_vertices
    S = np.zeros(nmax * nmax)
    rhs = np.zeros(nmax)
    vb = np.zeros(vmap.num_vertices, dtype=np.int32)
    n = lib.oracle_reduced_camera_system(ctypes.byref(d), ctypes.byref(options), mu, nmax * nmax, _p(S), _p(rhs),
                                         _p(vb))
    assert n >= 0, n
    return S[: n * n].reshape(n, n), rhs[:n], vb[: n // 6]


def landmark_quality(vmap, options):
    """isVisualLandmarkWellConstrained of every landmark (bool array)."""
    lib = load()
    d = vmap.desc()
    good = np.empty(vmap.num_landmarks, dtype=np.uint8)
    rc = lib.oracle_landmark_quality(ctypes.byref(d), ctypes.byref(options), _p(good))
    assert rc == 0, rc
    return good.astype(bool)


def solve_inertial_states(edges, options, fix_gyro_bias=False, fix_accel_bias=False, fix_velocity=False,
                          capacity=256):
    """Dense-LM restatement of the speed / bias solve with constant poses.
    edges: an InertialEdges. Returns (summary, rows, gyro_bias, velocity,
    accel_bias)."""
    from maplab_amd import abi
    lib = load()
    fn = lib.oracle_solve_inertial_states
    fn.restype = ctypes.c_int32
    vp = ctypes.c_void_p
    fn.argtypes = [ctypes.c_int32, ctypes.c_int32] + [vp] * 11 + [ctypes.POINTER(abi.Options), vp, vp, vp,
                                                                  ctypes.POINTER(abi.Summary)]
    a = edges._arrays()
    P, E = a[0].shape[0], a[4].shape[0]
    noise = np.ascontiguousarray(list(edges.noise) + [edges.gravity], dtype=np.float64)
    fix = np.array([fix_gyro_bias, fix_accel_bias, fix_velocity], dtype=np.int32)
    bg, v, ba = np.zeros((P, 3)), np.zeros((P, 3)), np.zeros((P, 3))
    its = (abi.IterationSummary * capacity)()
    sm = abi.Summary()
    sm.iterations_capacity = capacity
    sm.iterations = ctypes.cast(its, ctypes.POINTER(abi.IterationSummary))
    rc = fn(P, E, *[x.ctypes.data for x in a], noise.ctypes.data, fix.ctypes.data, ctypes.byref(options),
            bg.ctypes.data, v.ctypes.data, ba.ctypes.data, ctypes.byref(sm))
    assert rc == 0, rc
    return sm, [its[i] for i in range(min(sm.num_iterations, capacity))], bg, v, ba


def inertial_evaluate(imu, t_ns, noise, params, want_jac=True):
    """InertialErrorTerm::Evaluate for one edge: returns (residual (15),
    jacobians [8] ambient 15 x {7,3,3,3,7,3,3,3}, phi_accum, Q_accum,
    end_state (16))."""
    lib = load()
    fn = lib.oracle_inertial_evaluate
    fn.restype = ctypes.c_int32
    fn.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                   ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    imu = np.ascontiguousarray(imu, dtype=np.float64).reshape(-1, 6)
    t = np.ascontiguousarray(t_ns, dtype=np.int64)
    nz = np.ascontiguousarray(noise, dtype=np.float64)
    ps = [np.ascontiguousarray(p, dtype=np.float64) for p in params]
    pp = (ctypes.c_void_p * 8)(*[p.ctypes.data for p in ps])
    r = np.zeros(15)
    sizes = (7, 3, 3, 3, 7, 3, 3, 3)
    jac = [np.zeros((15, k)) for k in sizes]
    jp = (ctypes.c_void_p * 8)(*[j.ctypes.data for j in jac])
    phi, Q, xe = np.zeros((15, 15)), np.zeros((15, 15)), np.zeros(16)
    rc = fn(imu.shape[0], imu.ctypes.data, t.ctypes.data, nz.ctypes.data, pp, r.ctypes.data,
            jp if want_jac else None, phi.ctypes.data, Q.ctypes.data, xe.ctypes.data)
    assert rc == 0, rc
    return r, jac, phi, Q, xe
