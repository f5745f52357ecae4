"""Seeded single-observation configurations and tiny maps shared by the
oracle tests and the GPU parity tests."""
from __future__ import annotations

import numpy as np

import independent as ind
from maplab_amd.problem import VisualMap

CAMERAS = {
    0: np.zeros(4),
    1: np.array([-0.28, 0.07, 2e-4, 2e-4]),
    2: np.array([-0.01, 0.02, -0.005, 0.001]),
    3: np.array([0.9, 0.0, 0.0, 0.0]),  # fisheye (FOV) w
}
INTR = np.array([460.0, 460.0, 376.0, 240.0])
WIDTH, HEIGHT = 752, 480


def rand_quat(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    return q * np.sign(q[3])


def rand_pose(rng, scale=2.0):
    return np.concatenate([rand_quat(rng), rng.normal(size=3) * scale])


def observation_case(rng, error_type, dist_type, bf_translation=True):
    """Parameter blocks (reference order) of one observation whose landmark
    projects inside the image at 2-10 m depth, and a noisy measurement."""
    anchor, observer = rand_pose(rng), rand_pose(rng)
    bf_a, bf_o = rand_pose(rng), rand_pose(rng)
    if not bf_translation:
        bf_a[4:] = 0.0
        bf_o[4:] = 0.0
    if error_type != ind.GLOBAL:
        bf_a = np.array([0, 0, 0, 1.0, 0, 0, 0])
        bf_o = bf_a.copy()
    cam = np.concatenate([rand_quat(rng) * 0 + np.array([0.02, -0.01, 0.03, 1.0]) / np.linalg.norm([0.02, -0.01, 0.03, 1.0]),
                          rng.normal(size=3) * 0.05])
    z = rng.uniform(2.0, 10.0)
    u, v = rng.uniform(60, WIDTH - 60), rng.uniform(60, HEIGHT - 60)
    p_C = np.array([(u - INTR[2]) / INTR[0] * z, (v - INTR[3]) / INTR[1] * z, z])
    R_C_I = ind.R_jpl(cam[:4])
    p_I = R_C_I.T @ (p_C - cam[4:])
    if error_type == ind.LOCAL_KEYFRAME:
        p_B = p_I
    else:
        R_I_M = ind.R_jpl(observer[:4])
        p_M = R_I_M.T @ p_I + observer[4:]
        if error_type == ind.GLOBAL:
            R_G_M = ind.R_jpl(bf_o[:4])
            p_G = R_G_M @ p_M + bf_o[4:]
            R_G_LM = ind.R_jpl(bf_a[:4])
            p_M = R_G_LM.T @ (p_G - bf_a[4:])
        R_B_LM = ind.R_jpl(anchor[:4])
        p_B = R_B_LM @ (p_M - anchor[4:])
    params = [p_B, anchor, bf_a, bf_o, observer, cam[:4], cam[4:], INTR.copy(), CAMERAS[dist_type].copy()]
    r0, _ = ind.residual([0, 0], 1.0, error_type, dist_type, params)
    meas = r0 + rng.normal(size=2) * 2.0
    sigma = rng.uniform(0.3, 1.5)
    return meas, sigma, params


def tiny_map(rng, dist_type=1, n_vertices=6, n_landmarks=40, global_missions=False, noise=0.5, xi=None,
             n_cameras=1):
    """A small map whose landmarks are stored in their first observer; with
    global_missions, odd vertices live in a second mission (global error type);
    with xi, the camera is a unified-projection camera of that xi; with
    n_cameras = 2, a second camera of the rig (rotated 5 degrees about y,
    other intrinsics) observes every landmark its vertex sees as well."""
    verts = []
    for k in range(n_vertices):
        ang = 2 * np.pi * k / n_vertices
        c = np.array([3 * np.cos(ang), 3 * np.sin(ang), 0.2 * rng.normal()])
        f = -c / np.linalg.norm(c)
        up = np.array([0, 0, 1.0])
        x = np.cross(f, up); x /= np.linalg.norm(x)
        y = np.cross(f, x)
        R_M_C = np.stack([x, y, f], axis=1)
        from scipy.spatial.transform import Rotation
        q = Rotation.from_matrix(R_M_C).as_quat()
        q *= np.sign(q[3])
        verts.append(np.concatenate([q, c]))
    verts = np.array(verts)
    missions = np.array([(k % 2 if global_missions else 0) for k in range(n_vertices)], dtype=np.int32)
    bfs = np.array([[0, 0, 0, 1.0, 0, 0, 0], [0.0, 0.0, np.sin(0.05), np.cos(0.05), 0.1, -0.2, 0.05]])
    if not global_missions:
        bfs = bfs[:1]
    world = rng.uniform(-1.0, 1.0, size=(n_landmarks, 3))
    obs_v, obs_l, kps = [], [], []
    store = np.zeros(n_landmarks, dtype=np.int32)
    cam = np.array([0, 0, 0, 1.0, 0.05, 0, 0])
    th = np.deg2rad(5.0)
    cams = [cam, np.array([0, np.sin(th / 2), 0, np.cos(th / 2), -0.05, 0.01, 0.0])][:n_cameras]
    intrs = [INTR, INTR * np.array([1.02, 0.98, 1.0, 1.0])][:n_cameras]
    obs_c = []

    def to_mission(p_W, m):  # world == mission 0 frame; mission 1 via its baseframe
        if m == 0:
            return p_W
        R_G_M = ind.R_jpl(bfs[m, :4])
        return R_G_M.T @ (p_W - bfs[m, 4:])

    for l in range(n_landmarks):
        first = None
        for k in range(n_vertices):
            p_M = to_mission(world[l], missions[k])
            p_I = ind.R_jpl(verts[k, :4]) @ (p_M - verts[k, 4:])
            seen = False
            for ci, (cq, ki) in enumerate(zip(cams, intrs)):
                p_C = ind.R_jpl(cq[:4]) @ p_I + cq[4:]
                if p_C[2] < 0.5:
                    continue
                kp = ind.project(dist_type, ki, CAMERAS[dist_type], p_C, xi=xi)
                if not (0 <= kp[0] < WIDTH and 0 <= kp[1] < HEIGHT):
                    continue
                obs_v.append(k); obs_l.append(l); obs_c.append(ci); kps.append(kp + rng.normal(size=2) * noise)
                seen = True
            if seen and first is None:
                first = k
        store[l] = first if first is not None else 0
    lms = np.zeros((n_landmarks, 3))
    for l in range(n_landmarks):
        s = store[l]
        p_M = to_mission(world[l], missions[s])
        lms[l] = ind.R_jpl(verts[s, :4]) @ (p_M - verts[s, 4:])
    vconst = np.zeros(n_vertices, dtype=np.uint8); vconst[0] = 1
    lconst = np.zeros(n_landmarks, dtype=np.uint8)
    first0 = np.where(store == 0)[0]
    if len(first0):
        lconst[first0[0]] = 1
    C = len(cams)
    m = VisualMap(
        vertices=verts, baseframes=bfs, landmarks=lms, camera_q_CI__C_p_CI=np.array(cams),
        camera_intrinsics=np.array(intrs), camera_distortion=np.tile(CAMERAS[dist_type], (C, 1)),
        camera_distortion_type=np.full(C, dist_type, dtype=np.int32),
        camera_resolution=np.tile([WIDTH, HEIGHT], (C, 1)).astype(np.int32), vertex_mission=missions,
        landmark_store_vertex=store, obs_vertex=np.array(obs_v, dtype=np.int32),
        obs_landmark=np.array(obs_l, dtype=np.int32), obs_camera=np.array(obs_c, dtype=np.int32),
        obs_keypoint=np.array(kps), obs_sigma=np.full(len(obs_v), 0.5), vertex_constant=vconst,
        landmark_constant=lconst, true_vertices=verts.copy(), true_landmarks=lms.copy(),
        camera_model=None if xi is None else np.full(C, 1, dtype=np.int32),
        camera_xi=None if xi is None else np.full(C, xi))
    # corrupt
    m.vertices = m.vertices.copy()
    for k in range(1, n_vertices):
        m.vertices[k] = np.concatenate([ind.jpl_plus(m.vertices[k, :4], rng.normal(size=3) * 0.01),
                                        m.vertices[k, 4:] + rng.normal(size=3) * 0.03])
    free = lconst == 0
    m.landmarks = m.landmarks + free[:, None] * rng.normal(size=m.landmarks.shape) * 0.05
    return m.normalized()
