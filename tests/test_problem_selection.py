"""Problem construction of the optimizer driver (mlba_select_problem): which
observations enter, the blocks in the problem and the constant blocks
(constructOptimizationProblem, vi-optimization-builder.cc:154-374), against a
restatement written from the reference's loops, plus known-answer cases."""
import numpy as np
import pytest

from cases import tiny_map
from maplab_amd import optimizer
from maplab_amd.problem import VisualMap


def concat(a: VisualMap, b: VisualMap) -> VisualMap:
    """Two maps side by side: b's vertices, landmarks and missions follow a's."""
    P, L, M = a.num_vertices, a.num_landmarks, a.baseframes.shape[0]
    cat = np.concatenate
    return VisualMap(
        vertices=cat([a.vertices, b.vertices]), baseframes=cat([a.baseframes, b.baseframes]),
        landmarks=cat([a.landmarks, b.landmarks]), camera_q_CI__C_p_CI=a.camera_q_CI__C_p_CI,
        camera_intrinsics=a.camera_intrinsics, camera_distortion=a.camera_distortion,
        camera_distortion_type=a.camera_distortion_type, camera_resolution=a.camera_resolution,
        vertex_mission=cat([a.vertex_mission, b.vertex_mission + M]),
        landmark_store_vertex=cat([a.landmark_store_vertex, b.landmark_store_vertex + P]),
        obs_vertex=cat([a.obs_vertex, b.obs_vertex + P]), obs_landmark=cat([a.obs_landmark, b.obs_landmark + L]),
        obs_camera=cat([a.obs_camera, b.obs_camera]), obs_keypoint=cat([a.obs_keypoint, b.obs_keypoint]),
        obs_sigma=cat([a.obs_sigma, b.obs_sigma]), vertex_constant=np.zeros(P + b.num_vertices, np.uint8),
        landmark_constant=np.zeros(L + b.num_landmarks, np.uint8)).normalized()


def reference_selection(m, good, missions=None, min_landmarks_per_frame=0, fix_vertices=False,
                        fix_landmarks=False):
    """addLandmarkTermsForVertices + clusterMissionByLandmarkCoobservations +
    applyGaugeFixesForInitialVertices, written with Python sets."""
    M = m.baseframes.shape[0]
    sel_m = set(range(M)) if missions is None else {k for k in range(M) if missions[k]}
    frames = {}
    for i in range(m.num_observations):
        frames.setdefault((m.obs_vertex[i], m.obs_camera[i]), []).append(i)
    obs_in = np.zeros(m.num_observations, bool)
    keyframes, lms = set(), set()
    for v in range(m.num_vertices):                      # vertices of the missions to optimize
        if m.vertex_mission[v] not in sel_m:
            continue
        for (fv, c), idx in sorted(frames.items()):
            if fv != v:
                continue
            if min_landmarks_per_frame > 0 and sum(good[m.obs_landmark[i]] for i in idx) < min_landmarks_per_frame:
                continue
            keyframes.add(v)
            for i in idx:
                l = m.obs_landmark[i]
                if m.vertex_mission[m.landmark_store_vertex[l]] in sel_m and good[l]:
                    obs_in[i] = True
                    lms.add(l)
    # clusters: missions sharing an observed landmark, merged transitively
    observed = {k: {m.obs_landmark[i] for i in range(m.num_observations) if m.vertex_mission[m.obs_vertex[i]] == k}
                for k in sel_m}
    components = []
    for k in sorted(sel_m):
        comp = {k}
        for c in list(components):
            if any(observed[k] & observed[x] for x in c):
                comp |= c
                components.remove(c)
        components.append(comp)
    vconst = np.array([fix_vertices or bool(m.vertex_constant[v]) for v in range(m.num_vertices)])
    lconst = np.array([fix_landmarks or bool(m.landmark_constant[l]) for l in range(m.num_landmarks)])
    gauges = []
    for comp in sorted(components, key=min):
        first_mission = min(comp)
        vs = [v for v in range(m.num_vertices) if m.vertex_mission[v] == first_mission and v in keyframes]
        if not vs:
            continue
        v0 = vs[0]
        ls = [l for l in range(m.num_landmarks) if m.landmark_store_vertex[l] == v0 and l in lms]
        vconst[v0] = True
        lconst[ls[0]] = True
        gauges.append((v0, ls[0]))
    vin = np.array([v in keyframes for v in range(m.num_vertices)])
    lin = np.array([l in lms for l in range(m.num_landmarks)])
    return obs_in, vin, lin, vconst, lconst, gauges


def two_disconnected_missions(seed=0):
    rng = np.random.default_rng(seed)
    return concat(tiny_map(rng, dist_type=1, n_vertices=6, n_landmarks=50),
                  tiny_map(rng, dist_type=1, n_vertices=5, n_landmarks=40))


def three_missions(seed=1):
    rng = np.random.default_rng(seed)
    return concat(tiny_map(rng, dist_type=2, n_vertices=8, n_landmarks=60, global_missions=True),
                  tiny_map(rng, dist_type=2, n_vertices=6, n_landmarks=50))


def options(**kw):
    o = optimizer.default_vi_problem_options()
    for k, v in kw.items():
        setattr(o, k, v)
    return o


@pytest.mark.parametrize("builder", [two_disconnected_missions, three_missions])
@pytest.mark.parametrize("missions", [None, "first_only", "last_only"])
@pytest.mark.parametrize("min_per_frame", [0, 12, 30])
@pytest.mark.parametrize("fix", [False, True])
def test_selection_matches_restatement(builder, missions, min_per_frame, fix):
    m = builder()
    M = m.baseframes.shape[0]
    ms = None if missions is None else np.array([k == (0 if missions == "first_only" else M - 1) for k in range(M)])
    rng = np.random.default_rng(7)
    good = rng.uniform(size=m.num_landmarks) < 0.8
    sel = optimizer.select_problem(m, good, options(min_landmarks_per_frame=min_per_frame, fix_vertices=int(fix),
                                                    fix_landmark_positions=int(fix)), ms)
    obs_in, vin, lin, vconst, lconst, gauges = reference_selection(m, good, ms, min_per_frame, fix, fix)
    assert sel.obs_in_problem.tolist() == obs_in.tolist()
    assert sel.vertex_in_problem.tolist() == vin.tolist()
    assert sel.landmark_in_problem.tolist() == lin.tolist()
    assert sel.vertex_constant.tolist() == vconst.tolist()
    assert sel.landmark_constant.tolist() == lconst.tolist()
    assert list(zip(sel.gauge_vertices.tolist(), sel.gauge_landmarks.tolist())) == gauges


def test_gauge_known_answers():
    # Two missions that never co-observe a landmark are two clusters, each
    # fixing its first vertex and the first landmark stored there.
    m = two_disconnected_missions()
    good = np.ones(m.num_landmarks, bool)
    sel = optimizer.select_problem(m, good)
    assert sel.gauge_vertices.tolist() == [0, 6]
    first0 = int(np.flatnonzero(m.landmark_store_vertex == 0)[0])
    first6 = int(np.flatnonzero(m.landmark_store_vertex == 6)[0])
    assert sel.gauge_landmarks.tolist() == [first0, first6]
    assert sel.vertex_constant.sum() == 2 and sel.landmark_constant.sum() == 2
    # A two-mission map whose missions share landmarks is one cluster.
    one = tiny_map(np.random.default_rng(3), dist_type=1, n_vertices=8, n_landmarks=60, global_missions=True)
    one.vertex_constant[:] = 0
    one.landmark_constant[:] = 0
    sel = optimizer.select_problem(one, np.ones(one.num_landmarks, bool))
    assert sel.gauge_vertices.tolist() == [0]
    # Dropping every landmark stored in vertex 0 moves nothing but is an error
    # in the reference (CHECK): the first vertex must store a landmark.
    good = one.landmark_store_vertex != 0
    with pytest.raises(RuntimeError, match="stores no landmark"):
        optimizer.select_problem(one, good)


def test_unselected_mission_contributes_nothing():
    m = two_disconnected_missions()
    sel = optimizer.select_problem(m, np.ones(m.num_landmarks, bool), missions=np.array([False, True]))
    assert not sel.obs_in_problem[m.vertex_mission[m.obs_vertex] == 0].any()
    assert not sel.vertex_in_problem[:6].any() and sel.vertex_in_problem[6:].all()
    assert sel.gauge_vertices.tolist() == [6]
