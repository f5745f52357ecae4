"""Second, independent restatement of the visual reprojection error used to
pin the oracle: it works with Hamilton rotations from scipy
(scipy.spatial.transform.Rotation, scalar-last) instead of JPL matrices, and
with textbook camera models, so a sign, index or transposition mistake in the
oracle shows up as a mismatch.

maplab semantics being restated (see the reference files cited):
  * a vertex block stores the Hamilton q_M_I coefficients, so its JPL matrix
    is R_I_M = R_Ham(q)^T (optimization-state-buffer.cc:161-193);
  * JPL Plus(q, d) = q_d (x)_JPL q, i.e. Hamilton q (x) q_d
    (quaternion-param-jpl.cc:35-68);
  * p_C = R_C_I p_I + p_C_I (visual-error-term-inl.h:121).
"""
from __future__ import annotations

import numpy as np
from scipy.spatial.transform import Rotation

LOCAL_KEYFRAME, LOCAL_MISSION, GLOBAL = 0, 1, 2


def R_jpl(q):
    """JPL rotation matrix of coefficients q = [x y z w]: the transpose of the
    Hamilton matrix of the same coefficients."""
    return Rotation.from_quat(np.asarray(q, dtype=float)).as_matrix().T


def jpl_plus(q, d):
    """q (+) d with the small-angle branch of the reference."""
    d = np.asarray(d, dtype=float)
    sq = d @ d
    if sq < (0.262 * 2) ** 2:
        qd = np.array([0.5 * d[0], 0.5 * d[1], 0.5 * d[2], 1.0])
        qd /= np.linalg.norm(qd)
    else:
        h = 0.5 * np.sqrt(sq)
        qd = np.concatenate([d / (2 * h) * np.sin(h), [np.cos(h)]])
    prod = (Rotation.from_quat(q) * Rotation.from_quat(qd)).as_quat()  # Hamilton q * qd
    # scipy canonicalizes nothing here; enforce w >= 0 like the reference.
    prod = prod / np.linalg.norm(prod)
    return prod if prod[3] >= 0 else -prod


def distort_radtan(k, x, y):
    """Brown-Conrady: radial k1, k2 and tangential p1, p2."""
    k1, k2, p1, p2 = k
    r2 = x * x + y * y
    rad = 1 + k1 * r2 + k2 * r2 * r2
    return (x * rad + 2 * p1 * x * y + p2 * (r2 + 2 * x * x),
            y * rad + 2 * p2 * x * y + p1 * (r2 + 2 * y * y))


def distort_equidistant(k, x, y):
    """Kannala-Brandt: theta_d = theta (1 + k1 t^2 + k2 t^4 + k3 t^6 + k4 t^8)."""
    r = np.hypot(x, y)
    if r < 1e-10:
        return x, y
    th = np.arctan(r)
    thd = th * (1 + k[0] * th**2 + k[1] * th**4 + k[2] * th**6 + k[3] * th**8)
    s = thd / r if r > 1e-8 else 1.0
    return x * s, y * s


def distort_fov(k, x, y):
    """FOV model (Devernay & Faugeras 2001): r_d = atan(2 r tan(w / 2)) / w,
    with the reference's limits (w^2 < 1e-5: identity; r^2 < 1e-5: the
    factor's limit 2 tan(w / 2) / w)."""
    w = k[0]
    r2 = x * x + y * y
    if w * w < 1e-5:
        f = 1.0
    elif r2 < 1e-5:
        f = 2 * np.tan(w / 2) / w
    else:
        r = np.sqrt(r2)
        f = np.arctan(2 * r * np.tan(w / 2)) / (w * r)
    return x * f, y * f


def project(dist_type, intr, dist, p_C, xi=None):
    """Pinhole, or with xi the unified (sphere) model of Geyer & Daniilidis /
    Mei & Rives: the point is projected through the unit sphere, (x, y) /
    (z + xi |p|)."""
    if xi is None:
        x, y = p_C[0] / p_C[2], p_C[1] / p_C[2]
    else:
        den = p_C[2] + xi * np.linalg.norm(p_C)
        x, y = p_C[0] / den, p_C[1] / den
    if dist_type == 1:
        x, y = distort_radtan(dist, x, y)
    elif dist_type == 2:
        x, y = distort_equidistant(dist, x, y)
    elif dist_type == 3:
        x, y = distort_fov(dist, x, y)
    return np.array([intr[0] * x + intr[2], intr[1] * y + intr[3]])


def p_C_of(error_type, p_B, anchor, bf_anchor, bf_obs, observer, q_CI, p_CI):
    """Frame chain with Hamilton rotations. anchor / observer / baseframes are
    7-vectors [q | p]; baseframe blocks store the Hamilton q_M_G (inverse of
    q_G_M), whose JPL matrix is R_G_M."""
    R_C_I = R_jpl(q_CI)
    if error_type == LOCAL_KEYFRAME:
        p_I = np.asarray(p_B, float)
    else:
        R_M_B = Rotation.from_quat(anchor[:4]).as_matrix()      # = R_B_LM^T
        p_M = R_M_B @ p_B + anchor[4:]
        if error_type == GLOBAL:
            R_G_LM = R_jpl(bf_anchor[:4])
            p_G = R_G_LM @ p_M + bf_anchor[4:]
            R_G_M = R_jpl(bf_obs[:4])
            p_M = R_G_M.T @ (p_G - bf_obs[4:])
        R_M_I = Rotation.from_quat(observer[:4]).as_matrix()
        p_I = R_M_I.T @ (p_M - observer[4:])
    return R_C_I @ p_I + p_CI


def residual(meas, sigma, error_type, dist_type, params):
    p_C = p_C_of(error_type, params[0], params[1], params[2], params[3], params[4], params[5], params[6])
    kp = project(dist_type, params[7], params[8], p_C)
    return (kp - np.asarray(meas)) / sigma, p_C


def numdiff(f, n, h=1e-5):
    """5-point central stencil (aslam numdiff-jacobian-tester.h:180-191)."""
    cols = []
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        cols.append((8 * f(e) + f(-2 * e) - f(2 * e) - 8 * f(-e)) / (12 * h))
    return np.stack(cols, axis=-1)


def tangent_jacobians(meas, sigma, error_type, dist_type, params, h=1e-5):
    """Tangent Jacobians of the residual w.r.t. landmark (3), anchor pose (6),
    observer pose (6), q_CI (3), p_CI (3), intrinsics (4), distortion (4), by
    differentiating through JPL Plus for the rotations."""
    params = [np.asarray(p, dtype=float) for p in params]

    def r_with(idx, fn):
        def f(d):
            pp = [p.copy() for p in params]
            pp[idx] = fn(pp[idx], d)
            return residual(meas, sigma, error_type, dist_type, pp)[0]
        return f

    pose_plus = lambda x, d: np.concatenate([jpl_plus(x[:4], d[:3]), x[4:] + d[3:]])
    out = {
        0: numdiff(r_with(0, lambda x, d: x + d), 3, h),
        5: numdiff(r_with(5, lambda x, d: jpl_plus(x, d)), 3, h),
        6: numdiff(r_with(6, lambda x, d: x + d), 3, h),
        7: numdiff(r_with(7, lambda x, d: x + d), 4, h),
        8: numdiff(r_with(8, lambda x, d: x + d), 4, h),
    }
    if error_type != LOCAL_KEYFRAME:
        out[1] = numdiff(r_with(1, pose_plus), 6, h)
        out[4] = numdiff(r_with(4, pose_plus), 6, h)
    return out


def landmark_quality(vmap, min_observers=4, min_angle_deg=5.0, max_dist=40.0, min_dist=0.05, max_reproj_px=-1.0):
    """isVisualLandmarkWellConstrained (landmark-quality-metrics.cc:103-171)
    restated with Hamilton rotations: camera centres c_G = R_G_M (R_M_I c_I +
    p_M_I) + p_G_M with c_I = -R_C_I^T p_C_I, the landmark in G from its store
    vertex, and the disparity angle as the largest pairwise angle between
    incidence lines (antiparallel rays count as parallel)."""
    v, bf = vmap.vertices, vmap.baseframes
    out = np.zeros(vmap.num_landmarks, dtype=bool)
    by_lm = [[] for _ in range(vmap.num_landmarks)]
    for i, l in enumerate(vmap.obs_landmark):
        by_lm[l].append(i)

    def frame(k):
        m = vmap.vertex_mission[k]
        R_M_I = Rotation.from_quat(v[k, :4]).as_matrix()
        return R_jpl(bf[m, :4]), bf[m, 4:], R_M_I, v[k, 4:]

    for l, obs in enumerate(by_lm):
        if len(obs) < min_observers:
            continue
        R_G_M, p_G_M, R_M_I, p_M_I = frame(vmap.landmark_store_vertex[l])
        p_G = R_G_M @ (R_M_I @ vmap.landmarks[l] + p_M_I) + p_G_M
        sd, dirs, bad = [], [], False
        for i in obs:
            R_G_M, p_G_M, R_M_I, p_M_I = frame(vmap.obs_vertex[i])
            c = vmap.obs_camera[i]
            R_C_I = R_jpl(vmap.camera_q_CI__C_p_CI[c, :4])
            p_C_I = vmap.camera_q_CI__C_p_CI[c, 4:]
            c_G = R_G_M @ (R_M_I @ (-R_C_I.T @ p_C_I) + p_M_I) + p_G_M
            p_C = R_C_I @ (R_M_I.T @ (R_G_M.T @ (p_G - p_G_M) - p_M_I)) + p_C_I
            ray = c_G - p_G
            d = np.linalg.norm(ray)
            sd.append(-d if p_C[2] < 0 else d)
            if d > 0:
                dirs.append(ray / d)
            if max_reproj_px > 0:
                if p_C[2] <= 1e-10:
                    bad = True
                    continue
                kp = project(vmap.camera_distortion_type[c], vmap.camera_intrinsics[c], vmap.camera_distortion[c], p_C)
                if np.sum((kp - vmap.obs_keypoint[i]) ** 2) > max_reproj_px ** 2:
                    bad = True
        if bad or min(sd) > max_dist or min(sd) < min_dist:
            continue
        angle = 0.0
        if len(dirs) >= 2:
            D = np.array(dirs)
            cosines = np.abs(D @ D.T)[np.triu_indices(len(dirs), 1)]
            angle = np.degrees(np.arccos(min(1.0, cosines.min())))
        out[l] = angle >= min_angle_deg
    return out
