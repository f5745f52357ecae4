"""The oracle's landmark admission test (isVisualLandmarkWellConstrained,
landmark-quality-metrics.cc:103-171) against hand-built geometry with known
answers and against an independent Hamilton-rotation restatement."""
import numpy as np
import pytest

import independent as ind
from cases import tiny_map
from maplab_amd import abi, problem
from maplab_amd.problem import VisualMap


def quality_options(**kw):
    o = abi.default_quality_options()
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def geometry_map():
    """Landmarks with known admission outcomes under the default flags
    (4 observers, 5 deg, 0.05-40 m). Cameras look along +z of their vertex
    (identity orientation, zero extrinsics), one mission at the origin."""
    positions, landmarks = [], []

    def vertex(x, y, z):
        positions.append((x, y, z))
        return len(positions) - 1

    spread = [vertex(x, 0, 0) for x in (-3, -1, 1, 3)]             # +-16.7 deg at 10 m
    narrow = [vertex(x, 0, 0) for x in (-0.1, 0.0, 0.1, 0.2)]     # < 2 deg at 10 m
    behind = vertex(0.5, 0, 12)                                   # landmark 2 m behind it
    near = vertex(0.02, 0.0, 9.98)                                # 0.028 m from the landmark
    cases = [  # (landmark in G, observers, expected good)
        ((0, 0, 10), spread, True),
        ((0, 0, 10), spread[:3], False),                    # three observers
        ((0, 0, 50), spread, False),                        # closest observer > 40 m
        ((0, 0, 10), narrow, False),                        # disparity < 5 deg
        ((0, 0, 10), spread[:3] + [behind], False),         # negative signed distance
        ((0, 0, 10), spread[:3] + [near], False),           # closest observer < 0.05 m
        ((0.5, 0.3, 20), spread, True),                     # 17 deg, 20 m
    ]
    P = len(positions)
    verts = np.zeros((P, 7))
    verts[:, 3] = 1.0
    verts[:, 4:] = np.array(positions, dtype=float)
    intr = np.array([460.0, 460.0, 376.0, 240.0])
    obs_v, obs_l, kps, store, lms, expected = [], [], [], [], [], []
    for l, (p_G, observers, good) in enumerate(cases):
        p_G = np.array(p_G, dtype=float)
        s = observers[0]
        store.append(s)
        lms.append(p_G - verts[s, 4:])
        expected.append(good)
        for k in observers:
            p_C = p_G - verts[k, 4:]
            kp = intr[:2] * p_C[:2] / p_C[2] + intr[2:] if abs(p_C[2]) > 1e-9 else np.zeros(2)
            obs_v.append(k); obs_l.append(l); kps.append(kp)
    m = VisualMap(
        vertices=verts, baseframes=np.array([[0, 0, 0, 1.0, 0, 0, 0]]), landmarks=np.array(lms),
        camera_q_CI__C_p_CI=np.array([[0, 0, 0, 1.0, 0, 0, 0]]), camera_intrinsics=intr[None, :],
        camera_distortion=np.zeros((1, 4)), camera_distortion_type=np.array([0], dtype=np.int32),
        camera_resolution=np.array([[752, 480]], dtype=np.int32), vertex_mission=np.zeros(P, dtype=np.int32),
        landmark_store_vertex=np.array(store, dtype=np.int32), obs_vertex=np.array(obs_v, dtype=np.int32),
        obs_landmark=np.array(obs_l, dtype=np.int32), obs_camera=np.zeros(len(obs_v), dtype=np.int32),
        obs_keypoint=np.array(kps), obs_sigma=np.full(len(obs_v), 0.5),
        vertex_constant=np.zeros(P, dtype=np.uint8), landmark_constant=np.zeros(len(cases), dtype=np.uint8))
    return m.normalized(), np.array(expected)


def test_known_geometry(oracle):
    m, expected = geometry_map()
    good = oracle.landmark_quality(m, abi.default_quality_options())
    assert good.tolist() == expected.tolist()
    assert ind.landmark_quality(m).tolist() == expected.tolist()


def test_reprojection_threshold(oracle):
    # Landmark 0 with one keypoint moved by 4 px: rejected at 3 px, kept at 5 px
    # and with the check off (negative threshold, the default).
    m, _ = geometry_map()
    m.obs_keypoint[1] += np.array([4.0, 0.0])
    for px, want in ((3.0, False), (5.0, True), (-1.0, True)):
        good = oracle.landmark_quality(m, quality_options(max_reprojection_error_px=px))
        assert bool(good[0]) is want, px
        assert bool(ind.landmark_quality(m, max_reproj_px=px)[0]) is want, px


@pytest.mark.parametrize("thresholds", [
    dict(),
    dict(min_observation_angle_deg=10.0),
    dict(max_distance_from_closest_observer=8.0, min_distance_from_closest_observer=3.0),
    dict(max_reprojection_error_px=30.0),
    dict(min_observers=6),
])
def test_matches_independent_restatement(oracle, thresholds):
    # A random-walk map (local-mission chain) where each threshold rejects
    # part of the landmarks, and a two-mission map (global chain).
    walk = problem.workload_map(2, seed=3, num_keyframes=40, num_landmarks=1500)
    two_missions = tiny_map(np.random.default_rng(1), dist_type=2, n_vertices=8, n_landmarks=80,
                            global_missions=True)
    o = quality_options(**thresholds)
    for m in (walk, two_missions):
        got = oracle.landmark_quality(m, o)
        want = ind.landmark_quality(m, o.min_observers, o.min_observation_angle_deg,
                                    o.max_distance_from_closest_observer, o.min_distance_from_closest_observer,
                                    o.max_reprojection_error_px)
        assert got.tolist() == want.tolist()
    walk_good = oracle.landmark_quality(walk, o)
    assert 0 < walk_good.sum() < len(walk_good)
