"""Landmark admission (isVisualLandmarkWellConstrained) on the device through
the C ABI against the CPU oracle: identical accept / reject decisions."""
import numpy as np
import pytest

from cases import tiny_map
from maplab_amd import abi, problem
from maplab_amd.problem import VisualMap
from maplab_amd.solver import DeviceProblem
from test_oracle_quality import geometry_map, quality_options

pytestmark = pytest.mark.gpu

THRESHOLDS = [
    dict(),
    dict(min_observation_angle_deg=10.0),
    dict(max_distance_from_closest_observer=8.0, min_distance_from_closest_observer=3.0),
    dict(max_reprojection_error_px=30.0),
    dict(min_observers=6),
]


@pytest.mark.parametrize("thresholds", THRESHOLDS)
def test_landmark_quality_matches_oracle(oracle, thresholds):
    o = quality_options(**thresholds)
    maps = [problem.workload_map(2, seed=3, num_keyframes=40, num_landmarks=1500),
            tiny_map(np.random.default_rng(1), dist_type=2, n_vertices=8, n_landmarks=80, global_missions=True),
            geometry_map()[0]]
    for m in maps:
        print(f"landmark quality: {thresholds} L {m.num_landmarks} N {m.num_observations}", flush=True)
        with DeviceProblem(m) as dev:
            got = dev.landmark_quality(o)
        want = oracle.landmark_quality(m, o)
        assert got.tolist() == want.tolist()


def test_landmark_quality_known_geometry():
    m, expected = geometry_map()
    with DeviceProblem(m) as dev:
        assert dev.landmark_quality().tolist() == expected.tolist()


def many_observer_map(n_obs=300, seed=0):
    """One landmark 10 m in front of n_obs observers spread over a 5 m disc
    (more rays than the device keeps in LDS), plus one with 5 observers."""
    rng = np.random.default_rng(seed)
    P = n_obs
    verts = np.zeros((P, 7))
    verts[:, 3] = 1.0
    r = 5.0 * np.sqrt(rng.uniform(size=P))
    a = rng.uniform(0, 2 * np.pi, size=P)
    verts[:, 4] = r * np.cos(a)
    verts[:, 5] = r * np.sin(a)
    intr = np.array([460.0, 460.0, 376.0, 240.0])
    lms = np.array([[0.0, 0.0, 10.0], [0.0, 0.0, 10.0]]) - verts[0, 4:]
    obs_v = list(range(P)) + list(range(5))
    obs_l = [0] * P + [1] * 5
    kps = []
    for k, l in zip(obs_v, obs_l):
        p_C = np.array([0.0, 0.0, 10.0]) - verts[k, 4:]
        kps.append(intr[:2] * p_C[:2] / p_C[2] + intr[2:])
    m = VisualMap(
        vertices=verts, baseframes=np.array([[0, 0, 0, 1.0, 0, 0, 0]]), landmarks=lms,
        camera_q_CI__C_p_CI=np.array([[0, 0, 0, 1.0, 0, 0, 0]]), camera_intrinsics=intr[None, :],
        camera_distortion=np.zeros((1, 4)), camera_distortion_type=np.array([0], dtype=np.int32),
        camera_resolution=np.array([[752, 480]], dtype=np.int32), vertex_mission=np.zeros(P, dtype=np.int32),
        landmark_store_vertex=np.array([0, 0], dtype=np.int32), obs_vertex=np.array(obs_v, dtype=np.int32),
        obs_landmark=np.array(obs_l, dtype=np.int32), obs_camera=np.zeros(len(obs_v), dtype=np.int32),
        obs_keypoint=np.array(kps), obs_sigma=np.full(len(obs_v), 0.5),
        vertex_constant=np.zeros(P, dtype=np.uint8), landmark_constant=np.zeros(2, dtype=np.uint8))
    return m.normalized()


@pytest.mark.parametrize("angle", [5.0, 40.0, 60.0])
def test_landmark_quality_many_observers(oracle, angle):
    m = many_observer_map()
    o = quality_options(min_observation_angle_deg=angle)
    with DeviceProblem(m) as dev:
        got = dev.landmark_quality(o)
    assert got.tolist() == oracle.landmark_quality(m, o).tolist()
