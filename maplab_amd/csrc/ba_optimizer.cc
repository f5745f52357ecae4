// ba_optimizer.cc -- host-side problem construction and driver of the visual
// BA (map_optimization::VIMapOptimizer, vi-map-optimizer.cc:34-80), built on
// the device library's C ABI (include/maplab_ba.h).
#include "../../include/maplab_ba_optimizer.h"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

// Union-find over mission indices.
struct DisjointSets {
  std::vector<int> parent;
  explicit DisjointSets(int n) : parent(n) { std::iota(parent.begin(), parent.end(), 0); }
  int find(int a) {
    while (parent[a] != a) a = parent[a] = parent[parent[a]];
    return a;
  }
  void unite(int a, int b) {
    a = find(a);
    b = find(b);
    if (a != b) parent[std::max(a, b)] = std::min(a, b);
  }
};

int check_map(const mlba_problem_desc* d) {
  if (!d) return fail(MLBA_EINVAL, "NULL map");
  const int N = d->num_observations, L = d->num_landmarks, P = d->num_vertices, M = d->num_missions;
  if (N < 0 || L < 0 || P < 0 || M < 0) return fail(MLBA_EINVAL, "negative size");
  if ((N > 0 && (!d->obs_vertex || !d->obs_landmark || !d->obs_camera)) || (P > 0 && !d->vertex_mission) ||
      (L > 0 && !d->landmark_store_vertex))
    return fail(MLBA_EINVAL, "NULL array in map");
  for (int o = 0; o < N; ++o)
    if (d->obs_vertex[o] < 0 || d->obs_vertex[o] >= P || d->obs_landmark[o] < 0 || d->obs_landmark[o] >= L)
      return fail(MLBA_EINVAL, "observation index out of range");
  for (int l = 0; l < L; ++l)
    if (d->landmark_store_vertex[l] < 0 || d->landmark_store_vertex[l] >= P)
      return fail(MLBA_EINVAL, "landmark_store_vertex out of range");
  for (int v = 0; v < P; ++v)
    if (d->vertex_mission[v] < 0 || d->vertex_mission[v] >= M) return fail(MLBA_EINVAL, "vertex_mission out of range");
  return MLBA_OK;
}

}  // namespace

extern "C" {

void mlba_vi_problem_options_init(mlba_vi_problem_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->add_visual_constraints = 1;
  o->fix_landmark_positions = 0;
  o->fix_vertices = 0;
  mlba_calibration_options_init(&o->calibration);
  o->min_landmarks_per_frame = 0;
  o->enable_visual_outlier_rejection = 1;
  mlba_options_init(&o->solver);
  mlba_quality_options_init(&o->landmark_quality);
}

const char* mlba_optimizer_last_error(void) { return g_error.c_str(); }

int32_t mlba_select_problem(const mlba_problem_desc* d, const mlba_vi_problem_options* o,
                            const uint8_t* mission_selected, const uint8_t* landmark_good,
                            mlba_problem_selection* sel) {
  int rc = check_map(d);
  if (rc) return rc;
  if (!o || !sel || (d->num_landmarks > 0 && !landmark_good) ||
      (d->num_observations > 0 && !sel->obs_in_problem) || (d->num_vertices > 0 && !sel->vertex_in_problem) ||
      (d->num_landmarks > 0 && !sel->landmark_in_problem) || (d->num_vertices > 0 && !sel->vertex_constant) ||
      (d->num_landmarks > 0 && !sel->landmark_constant))
    return fail(MLBA_EINVAL, "NULL argument");
  if (o->min_landmarks_per_frame < 0) return fail(MLBA_EINVAL, "min_landmarks_per_frame must be >= 0");
  const int N = d->num_observations, L = d->num_landmarks, P = d->num_vertices, M = d->num_missions;
  auto selected = [&](int m) { return mission_selected == nullptr || mission_selected[m] != 0; };

  // Frames (vertex, camera) with too few well-constrained landmarks are
  // skipped (optimization-terms-addition.cc:268-278); a frame counts the
  // good landmarks of its keypoints (vi-map-queries.cc:260-278).
  std::vector<int64_t> frame_keys;   // distinct (vertex, camera) frames, sorted
  std::vector<int32_t> frame_good;   // well-constrained landmark count of each frame
  auto key_of = [&](int i) { return (int64_t)d->obs_vertex[i] * std::max(d->num_cameras, 1) + d->obs_camera[i]; };
  if (o->min_landmarks_per_frame > 0) {
    std::vector<std::pair<int64_t, int>> fr(N);
    for (int i = 0; i < N; ++i) fr[i] = {key_of(i), landmark_good[d->obs_landmark[i]] ? 1 : 0};
    std::sort(fr.begin(), fr.end());
    for (const auto& f : fr) {
      if (frame_keys.empty() || frame_keys.back() != f.first) {
        frame_keys.push_back(f.first);
        frame_good.push_back(0);
      }
      frame_good.back() += f.second;
    }
  }
  auto frame_good_count = [&](int i) {
    const size_t k = std::lower_bound(frame_keys.begin(), frame_keys.end(), key_of(i)) - frame_keys.begin();
    return frame_good[k];
  };
  std::memset(sel->vertex_in_problem, 0, P);
  std::memset(sel->landmark_in_problem, 0, L);
  int n_obs = 0;
  for (int i = 0; i < N; ++i) {
    const int v = d->obs_vertex[i], l = d->obs_landmark[i];
    bool in = o->add_visual_constraints != 0 && selected(d->vertex_mission[v]);
    const bool frame_ok = in && (o->min_landmarks_per_frame == 0 || frame_good_count(i) >= o->min_landmarks_per_frame);
    if (frame_ok) sel->vertex_in_problem[v] = 1;  // keyframes_in_problem
    in = frame_ok && selected(d->vertex_mission[d->landmark_store_vertex[l]]) && landmark_good[l];
    sel->obs_in_problem[i] = in ? 1 : 0;
    if (in) {
      sel->landmark_in_problem[l] = 1;
      ++n_obs;
    }
  }
  sel->num_observations = n_obs;
  sel->num_vertices = (int)std::count(sel->vertex_in_problem, sel->vertex_in_problem + P, 1);
  sel->num_landmarks = (int)std::count(sel->landmark_in_problem, sel->landmark_in_problem + L, 1);

  // Constant blocks: the caller's, then the option-wide fixes.
  for (int v = 0; v < P; ++v)
    sel->vertex_constant[v] = (o->fix_vertices || (d->vertex_constant && d->vertex_constant[v])) ? 1 : 0;
  for (int l = 0; l < L; ++l)
    sel->landmark_constant[l] =
        (o->fix_landmark_positions || (d->landmark_constant && d->landmark_constant[l])) ? 1 : 0;

  // Mission co-observation clusters of the selected missions: missions whose
  // vertices observe a common landmark (clusterMissionByLandmarkCoobservations).
  DisjointSets ds(std::max(M, 1));
  {
    std::vector<int> first_mission(L, -1);
    for (int i = 0; i < N; ++i) {
      const int m = d->vertex_mission[d->obs_vertex[i]], l = d->obs_landmark[i];
      if (!selected(m)) continue;
      if (first_mission[l] < 0) first_mission[l] = m;
      else ds.unite(first_mission[l], m);
    }
  }
  // Visual-only gauge per cluster (vi-optimization-builder.cc:289-331: scale,
  // global position and all rotation DoF are unobservable): fix the first
  // vertex in the problem of the cluster's first mission and the first
  // landmark in the problem stored in that vertex.
  std::vector<int> first_vertex(std::max(M, 1), -1), first_landmark(std::max(P, 1), -1);
  for (int v = 0; v < P; ++v)
    if (sel->vertex_in_problem[v] && first_vertex[d->vertex_mission[v]] < 0) first_vertex[d->vertex_mission[v]] = v;
  for (int l = 0; l < L; ++l)
    if (sel->landmark_in_problem[l] && first_landmark[d->landmark_store_vertex[l]] < 0)
      first_landmark[d->landmark_store_vertex[l]] = l;
  int clusters = 0;
  for (int m = 0; m < M; ++m) {
    if (!selected(m) || ds.find(m) != m) continue;  // root = lowest mission index of the cluster
    // Clusters without visual constraints in the problem have no gauge to fix here.
    int v = first_vertex[m];
    if (v < 0) {
      for (int m2 = 0; m2 < M && v < 0; ++m2)
        if (selected(m2) && ds.find(m2) == m && first_vertex[m2] >= 0) v = -2;  // first mission has no vertex
      if (v == -2) return fail(MLBA_ESTATE, "the first mission of a cluster has no vertex in the problem");
      continue;
    }
    const int l = first_landmark[v];
    if (l < 0) return fail(MLBA_ESTATE, "the first vertex of a cluster stores no landmark in the problem");
    sel->vertex_constant[v] = 1;
    sel->landmark_constant[l] = 1;
    if (sel->gauge_vertex) sel->gauge_vertex[clusters] = v;
    if (sel->gauge_landmark) sel->gauge_landmark[clusters] = l;
    ++clusters;
  }
  sel->num_clusters = clusters;
  return MLBA_OK;
}

int32_t mlba_vi_map_optimize(const mlba_problem_desc* d, const mlba_vi_problem_options* o,
                             const uint8_t* mission_selected, uint8_t* landmark_quality, double* vertices_out,
                             double* landmarks_out, mlba_summary* summary) {
  return mlba_vi_map_optimize_cameras(d, o, mission_selected, landmark_quality, vertices_out, landmarks_out,
                                      nullptr, nullptr, nullptr, nullptr, summary);
}

int32_t mlba_vi_map_optimize_cameras(const mlba_problem_desc* d, const mlba_vi_problem_options* o,
                                     const uint8_t* mission_selected, uint8_t* landmark_quality,
                                     double* vertices_out, double* landmarks_out, double* cam_q_out,
                                     double* intrinsics_out, double* distortion_out, double* xi_out,
                                     mlba_summary* summary) {
  int rc = check_map(d);
  if (rc) return rc;
  if (!o || (d->num_vertices > 0 && !vertices_out) || (d->num_landmarks > 0 && !landmarks_out))
    return fail(MLBA_EINVAL, "NULL argument");
  const int N = d->num_observations, L = d->num_landmarks, P = d->num_vertices, M = d->num_missions;
  if (M == 0 || (mission_selected && std::count(mission_selected, mission_selected + M, 0) == M))
    return fail(MLBA_EINVAL, "Nothing to optimize.");

  // Landmark admission (isLandmarkWellConstrained, landmark-quality-metrics.cc:
  // 81-101): stored quality first, the geometric test on the device for the
  // unknown ones.
  std::vector<uint8_t> good(L, 0);
  {
    bool need_device = landmark_quality == nullptr;
    for (int l = 0; l < L && !need_device; ++l) need_device = landmark_quality[l] == MLBA_QUALITY_UNKNOWN;
    std::vector<uint8_t> geometric(L, 0);
    if (need_device && L > 0) {
      mlba_handle* h = nullptr;
      if ((rc = mlba_create(d, &h))) return fail(rc, mlba_last_error(nullptr));
      rc = mlba_landmark_quality(h, &o->landmark_quality, geometric.data());
      if (rc) {
        g_error = mlba_last_error(h);
        mlba_destroy(h);
        return rc;
      }
      mlba_destroy(h);
    }
    for (int l = 0; l < L; ++l) {
      const int q = landmark_quality ? landmark_quality[l] : MLBA_QUALITY_UNKNOWN;
      good[l] = q == MLBA_QUALITY_LOCALIZATION_SUMMARY ? 1
                : q == MLBA_QUALITY_UNKNOWN           ? geometric[l]
                                                      : (q == MLBA_QUALITY_GOOD ? 1 : 0);
    }
  }

  std::vector<uint8_t> obs_in(N), v_in(P), l_in(L), v_const(P), l_const(L);
  mlba_problem_selection sel{};
  sel.obs_in_problem = obs_in.data();
  sel.vertex_in_problem = v_in.data();
  sel.landmark_in_problem = l_in.data();
  sel.vertex_constant = v_const.data();
  sel.landmark_constant = l_const.data();
  if ((rc = mlba_select_problem(d, o, mission_selected, good.data(), &sel))) return rc;

  // The problem's observations (the map's vertex and landmark arrays stay
  // whole: indices are unchanged, blocks outside the problem keep no
  // residuals and are inactive).
  std::vector<int32_t> ov, ol, oc;
  std::vector<double> okp, osig;
  ov.reserve(sel.num_observations);
  ol.reserve(sel.num_observations);
  oc.reserve(sel.num_observations);
  okp.reserve(2 * (size_t)sel.num_observations);
  osig.reserve(sel.num_observations);
  for (int i = 0; i < N; ++i) {
    if (!obs_in[i]) continue;
    ov.push_back(d->obs_vertex[i]);
    ol.push_back(d->obs_landmark[i]);
    oc.push_back(d->obs_camera[i]);
    okp.push_back(d->obs_keypoint[2 * i]);
    okp.push_back(d->obs_keypoint[2 * i + 1]);
    osig.push_back(d->obs_sigma[i]);
  }
  mlba_problem_desc pd = *d;
  pd.num_observations = sel.num_observations;
  pd.obs_vertex = ov.data();
  pd.obs_landmark = ol.data();
  pd.obs_camera = oc.data();
  pd.obs_keypoint = okp.data();
  pd.obs_sigma = osig.data();
  pd.vertex_constant = v_const.data();
  pd.landmark_constant = l_const.data();

  mlba_handle* h = nullptr;
  if ((rc = mlba_create(&pd, &h))) return fail(rc, mlba_last_error(nullptr));
  std::vector<uint8_t> active(l_in);
  if (L > 0 && (rc = mlba_set_landmark_active(h, active.data()))) {
    g_error = mlba_last_error(h);
    mlba_destroy(h);
    return rc;
  }
  const mlba_calibration_options& co = o->calibration;
  const bool calibrate = !co.fix_intrinsics || !co.fix_distortion || !co.fix_extrinsics_rotation ||
                         !co.fix_extrinsics_translation;
  if (calibrate)
    rc = mlba_solve_calibration(h, &o->solver, &co, summary);
  else
    rc = o->enable_visual_outlier_rejection ? mlba_solve_with_outlier_rejection(h, &o->solver, summary)
                                            : mlba_solve(h, &o->solver, summary);
  if (rc == MLBA_OK) rc = mlba_download_state(h, vertices_out, landmarks_out);
  if (rc == MLBA_OK) rc = mlba_download_cameras(h, cam_q_out, intrinsics_out, distortion_out, xi_out);
  if (rc == MLBA_OK && landmark_quality && L > 0) {
    std::vector<uint8_t> after(L);
    rc = mlba_get_landmark_active(h, after.data());
    for (int l = 0; l < L && rc == MLBA_OK; ++l)
      if (l_in[l] && !after[l]) landmark_quality[l] = MLBA_QUALITY_BAD;
  }
  if (rc) g_error = mlba_last_error(h);
  mlba_destroy(h);
  return rc;
}

}  // extern "C"
