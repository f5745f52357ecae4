// synth.cc -- seeded synthetic visual maps for the benchmark workloads.
//
// Mirrors the role of maplab's test/vi-map-generator-6dof (pose graph +
// landmarks observed where they project, 6dof-pose-graph-gen.cc:164-415)
// and the seeded corruption of vi-mapping-test-app.cc:78-153, on a
// platform-stable random stream (xoshiro256** + Box-Muller).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/maplab_ba_synth.h"
#include "ba_math.h"

namespace {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) {  // splitmix64
      x += 0x9E3779B97F4A7C15ull;
      uint64_t z = x;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      s[i] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {  // xoshiro256**
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {  // Box-Muller, one draw per call
    double u1 = uniform();
    while (u1 <= 0.0) u1 = uniform();
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  uint64_t below(uint64_t n) { return next() % n; }
};

// Hamilton quaternion [x y z w] of a rotation matrix (Shepperd), w >= 0.
void quat_of_matrix(const double* R, double* q) {
  const double tr = R[0] + R[4] + R[8];
  double x, y, z, w;
  if (tr > 0) {
    const double s = std::sqrt(tr + 1.0) * 2;
    w = 0.25 * s; x = (R[7] - R[5]) / s; y = (R[2] - R[6]) / s; z = (R[3] - R[1]) / s;
  } else if (R[0] > R[4] && R[0] > R[8]) {
    const double s = std::sqrt(1.0 + R[0] - R[4] - R[8]) * 2;
    w = (R[7] - R[5]) / s; x = 0.25 * s; y = (R[1] + R[3]) / s; z = (R[2] + R[6]) / s;
  } else if (R[4] > R[8]) {
    const double s = std::sqrt(1.0 + R[4] - R[0] - R[8]) * 2;
    w = (R[2] - R[6]) / s; x = (R[1] + R[3]) / s; y = 0.25 * s; z = (R[5] + R[7]) / s;
  } else {
    const double s = std::sqrt(1.0 + R[8] - R[0] - R[4]) * 2;
    w = (R[3] - R[1]) / s; x = (R[2] + R[6]) / s; y = (R[5] + R[7]) / s; z = 0.25 * s;
  }
  const double n = std::sqrt(x * x + y * y + z * z + w * w) * (w < 0 ? -1.0 : 1.0);
  q[0] = x / n; q[1] = y / n; q[2] = z / n; q[3] = w / n;
}

// Rotation matrix of a rotation vector (Rodrigues).
void rotvec_matrix(const double* v, double* R) {
  const double th = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (th < 1e-15) {
    for (int i = 0; i < 9; ++i) R[i] = (i % 4 == 0) ? 1.0 : 0.0;
    return;
  }
  const double k[3] = {v[0] / th, v[1] / th, v[2] / th};
  const double c = std::cos(th), s = std::sin(th), t = 1 - c;
  R[0] = c + k[0] * k[0] * t;        R[1] = k[0] * k[1] * t - k[2] * s; R[2] = k[0] * k[2] * t + k[1] * s;
  R[3] = k[1] * k[0] * t + k[2] * s; R[4] = c + k[1] * k[1] * t;        R[5] = k[1] * k[2] * t - k[0] * s;
  R[6] = k[2] * k[0] * t - k[1] * s; R[7] = k[2] * k[1] * t + k[0] * s; R[8] = c + k[2] * k[2] * t;
}

}  // namespace

struct mlba_synth_map {
  mlba_synth_config cfg;
  int P = 0, L = 0, N = 0;
  std::vector<double> vertices, vertices_true, landmarks, landmarks_true;
  std::vector<double> baseframes, camera, intrinsics, distortion;
  std::vector<int32_t> dist_type, resolution, vertex_mission, store_vertex;
  std::vector<int32_t> obs_vertex, obs_landmark, obs_camera;
  std::vector<double> obs_kp, obs_sigma;
  std::vector<uint8_t> vertex_constant, landmark_constant;
};

extern "C" {

int32_t mlba_synth_config_init(mlba_synth_config* c, int32_t workload) {
  if (!c) return MLBA_EINVAL;
  std::memset(c, 0, sizeof(*c));
  // Camera of the configurations: pinhole-radtan 752x480, f = 460.
  c->intrinsics[0] = 460.0; c->intrinsics[1] = 460.0; c->intrinsics[2] = 376.0; c->intrinsics[3] = 240.0;
  c->width = 752; c->height = 480;
  c->distortion_type = mlba::kDistRadtan;
  c->distortion[0] = -0.28; c->distortion[1] = 0.07; c->distortion[2] = 2e-4; c->distortion[3] = 2e-4;
  c->p_C_I[0] = 0.05;  // +x offset of the camera from the IMU
  c->keypoint_sigma_px = 0.5;
  c->rot_noise_rad = 0.02;
  c->pos_noise_m = 0.05;
  c->landmark_noise_m = 0.1;
  c->min_depth_m = 0.1;
  c->min_obs_per_landmark = 2;
  c->max_obs_per_landmark = 60;
  c->detection_probability = 1.0;
  c->anchor_mode = MLBA_ANCHOR_FIRST_OBSERVER;
  c->fix_scale_landmark = 1;
  c->seed = 1;
  if (workload == 1) {
    c->kind = MLBA_SYNTH_CIRCLE;
    c->num_keyframes = 10;
    c->num_landmarks = 2000;
    c->box_half_m = 2.0;
    c->circle_radius_m = 1.0;
    c->max_range_m = 1e9;
    c->min_obs_per_landmark = 2;
    c->max_obs_per_landmark = 1 << 30;
  } else if (workload == 2 || workload == 3) {
    c->kind = MLBA_SYNTH_RANDOM_WALK;
    c->num_keyframes = workload == 2 ? 500 : 2000;
    c->num_landmarks = workload == 2 ? 200000 : 1000000;
    c->step_m = 0.3;
    c->yaw_drift_rad = 0.1;
    c->min_dist_m = 2.0;
    c->max_dist_m = 15.0;
    c->max_range_m = 25.0;
    c->outlier_fraction = 0.02;
  } else {
    return MLBA_EINVAL;
  }
  return MLBA_OK;
}

int32_t mlba_synth_generate(const mlba_synth_config* cfg, mlba_synth_map** out) {
  if (!cfg || !out || cfg->num_keyframes < 1 || cfg->num_landmarks < 1) return MLBA_EINVAL;
  mlba_synth_map* m = new (std::nothrow) mlba_synth_map();
  if (!m) return MLBA_ENOMEM;
  m->cfg = *cfg;
  Rng rng(cfg->seed);
  const int K = cfg->num_keyframes;
  const bool global = cfg->anchor_mode == MLBA_ANCHOR_GLOBAL;
  // Camera centres and orientations (R_M_C columns = camera x, y, z in M).
  std::vector<double> centre(3 * K), R_M_C(9 * K);
  auto set_cam = [&](int k, const double* c, const double* fwd, const double* up) {
    double x[3] = {fwd[1] * up[2] - fwd[2] * up[1], fwd[2] * up[0] - fwd[0] * up[2], fwd[0] * up[1] - fwd[1] * up[0]};
    // x = fwd x up points right for a z-up world; y = z x x points down.
    const double nx = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    for (double& t : x) t /= nx;
    const double y[3] = {fwd[1] * x[2] - fwd[2] * x[1], fwd[2] * x[0] - fwd[0] * x[2], fwd[0] * x[1] - fwd[1] * x[0]};
    for (int i = 0; i < 3; ++i) {
      centre[3 * k + i] = c[i];
      R_M_C[9 * k + i * 3 + 0] = x[i];
      R_M_C[9 * k + i * 3 + 1] = y[i];
      R_M_C[9 * k + i * 3 + 2] = fwd[i];
    }
  };
  const double up[3] = {0, 0, 1};
  if (cfg->kind == MLBA_SYNTH_CIRCLE) {
    for (int k = 0; k < K; ++k) {
      const double a = 6.283185307179586 * k / K;
      const double c[3] = {cfg->circle_radius_m * std::cos(a), cfg->circle_radius_m * std::sin(a), 0.0};
      const double f[3] = {-std::cos(a), -std::sin(a), 0.0};
      set_cam(k, c, f, up);
    }
  } else {
    double pos[3] = {0, 0, 0}, yaw = 0.0;
    for (int k = 0; k < K; ++k) {
      const double f[3] = {std::cos(yaw), std::sin(yaw), 0.0};
      set_cam(k, pos, f, up);
      pos[0] += cfg->step_m * f[0];
      pos[1] += cfg->step_m * f[1];
      yaw += cfg->yaw_drift_rad * rng.normal();
    }
  }
  // Vertex (IMU) poses: R_C_I = I, so R_I_M = R_C_M and p_M_I = c + R_M_C p_C_I.
  const int P = K + (global ? 1 : 0);
  m->P = P;
  m->vertices_true.assign(7 * P, 0.0);
  for (int k = 0; k < K; ++k) {
    const double* Rmc = &R_M_C[9 * k];
    quat_of_matrix(Rmc, &m->vertices_true[7 * k]);  // Hamilton q_M_I = JPL q_IM coefficients
    double t[3];
    mlba::mat3_vec(Rmc, cfg->p_C_I, t);
    for (int i = 0; i < 3; ++i) m->vertices_true[7 * k + 4 + i] = centre[3 * k + i] + t[i];
  }
  if (global) m->vertices_true[7 * K + 3] = 1.0;  // identity store vertex
  // Landmarks and observations.
  const double W = cfg->width, H = cfg->height;
  std::vector<double> lm_world;
  std::vector<int> ob_v, ob_l;
  std::vector<double> ob_kp;
  std::vector<int> cand;
  std::vector<double> cand_kp;
  int L = 0;
  int attempts = 0;
  const long max_attempts = 200L * cfg->num_landmarks + 1000;
  while (L < cfg->num_landmarks && attempts < max_attempts) {
    ++attempts;
    double X[3];
    if (cfg->kind == MLBA_SYNTH_CIRCLE) {
      for (int i = 0; i < 3; ++i) X[i] = (2.0 * rng.uniform() - 1.0) * cfg->box_half_m;
    } else {
      const int k = (int)rng.below(K);
      double d[3];
      double n2 = 0;
      do {
        for (int i = 0; i < 3; ++i) d[i] = 2.0 * rng.uniform() - 1.0;
        n2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      } while (n2 > 1.0 || n2 < 1e-6);
      const double a3 = std::pow(cfg->min_dist_m, 3), b3 = std::pow(cfg->max_dist_m, 3);
      const double r = std::cbrt(a3 + rng.uniform() * (b3 - a3));
      const double n = std::sqrt(n2);
      for (int i = 0; i < 3; ++i) X[i] = centre[3 * k + i] + d[i] / n * r;
      double dmin = 1e300;
      for (int j = 0; j < K; ++j) {
        double s = 0;
        for (int i = 0; i < 3; ++i) { const double t = X[i] - centre[3 * j + i]; s += t * t; }
        dmin = std::min(dmin, s);
      }
      if (dmin < cfg->min_dist_m * cfg->min_dist_m) continue;
    }
    cand.clear();
    cand_kp.clear();
    for (int k = 0; k < K; ++k) {
      double t[3], pc[3];
      for (int i = 0; i < 3; ++i) t[i] = X[i] - centre[3 * k + i];
      const double range2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
      if (range2 > cfg->max_range_m * cfg->max_range_m) continue;
      mlba::mat3t_vec(&R_M_C[9 * k], t, pc);
      if (pc[2] < cfg->min_depth_m) continue;
      double kp[2], jp[6];
      const int st = mlba::project(cfg->distortion_type, cfg->intrinsics, cfg->distortion, W, H, pc, kp, jp);
      if (st != mlba::kVisible) continue;
      if (cfg->detection_probability < 1.0 && rng.uniform() >= cfg->detection_probability) continue;
      cand.push_back(k);
      cand_kp.push_back(kp[0]);
      cand_kp.push_back(kp[1]);
    }
    if ((int)cand.size() < cfg->min_obs_per_landmark) continue;
    // Subsample to the per-landmark cap (keeps order of keyframes).
    std::vector<int> keep(cand.size());
    for (size_t i = 0; i < keep.size(); ++i) keep[i] = (int)i;
    if ((int)cand.size() > cfg->max_obs_per_landmark) {
      for (int i = 0; i < cfg->max_obs_per_landmark; ++i) {
        const int j = i + (int)rng.below(keep.size() - i);
        std::swap(keep[i], keep[j]);
      }
      keep.resize(cfg->max_obs_per_landmark);
      std::sort(keep.begin(), keep.end());
    }
    for (int i : keep) {
      ob_v.push_back(cand[i]);
      ob_l.push_back(L);
      double u = cand_kp[2 * i], v = cand_kp[2 * i + 1];
      if (cfg->outlier_fraction > 0 && rng.uniform() < cfg->outlier_fraction) {
        u = rng.uniform() * W;
        v = rng.uniform() * H;
      } else {
        u += cfg->keypoint_sigma_px * rng.normal();
        v += cfg->keypoint_sigma_px * rng.normal();
      }
      ob_kp.push_back(u);
      ob_kp.push_back(v);
    }
    for (int i = 0; i < 3; ++i) lm_world.push_back(X[i]);
    ++L;
  }
  m->L = L;
  m->N = (int)ob_v.size();
  // Store vertex and p_B (landmark in the store vertex's IMU frame).
  m->store_vertex.assign(L, global ? K : 0);
  if (!global) {
    std::vector<int> first(L, K);
    for (int i = 0; i < m->N; ++i) first[ob_l[i]] = std::min(first[ob_l[i]], ob_v[i]);
    m->store_vertex = first;
  }
  m->landmarks_true.assign(3 * L, 0.0);
  for (int l = 0; l < L; ++l) {
    const int s = m->store_vertex[l];
    double R[9], t[3];
    mlba::rot_jpl(&m->vertices_true[7 * s], R);  // R_I_M
    for (int i = 0; i < 3; ++i) t[i] = lm_world[3 * l + i] - m->vertices_true[7 * s + 4 + i];
    mlba::mat3_vec(R, t, &m->landmarks_true[3 * l]);
  }
  m->vertex_mission.assign(P, 0);
  m->baseframes.assign(7, 0.0);
  m->baseframes[3] = 1.0;
  if (global) {
    m->vertex_mission[K] = 1;
    m->baseframes.resize(14, 0.0);
    m->baseframes[10] = 1.0;
  }
  m->camera = {0, 0, 0, 1, cfg->p_C_I[0], cfg->p_C_I[1], cfg->p_C_I[2]};
  m->intrinsics.assign(cfg->intrinsics, cfg->intrinsics + 4);
  m->distortion.assign(cfg->distortion, cfg->distortion + 4);
  m->dist_type = {cfg->distortion_type};
  m->resolution = {cfg->width, cfg->height};
  m->obs_vertex = ob_v;
  m->obs_landmark = ob_l;
  m->obs_camera.assign(m->N, 0);
  m->obs_kp = ob_kp;
  m->obs_sigma.assign(m->N, cfg->keypoint_sigma_px);
  // Corrupted initial state: left-multiplied rotation-vector noise and
  // position noise on every vertex, position noise on every landmark.
  m->vertices = m->vertices_true;
  for (int k = 0; k < K; ++k) {
    double v[3] = {cfg->rot_noise_rad * rng.normal(), cfg->rot_noise_rad * rng.normal(),
                   cfg->rot_noise_rad * rng.normal()};
    double dR[9], Rmi[9], Rt[9], Rn[9];
    rotvec_matrix(v, dR);
    mlba::rot_jpl(&m->vertices_true[7 * k], Rt);  // R_I_M
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Rmi[i * 3 + j] = Rt[j * 3 + i];
    mlba::mat3_mul(dR, Rmi, Rn);
    quat_of_matrix(Rn, &m->vertices[7 * k]);
    for (int i = 0; i < 3; ++i) m->vertices[7 * k + 4 + i] += cfg->pos_noise_m * rng.normal();
  }
  m->landmarks = m->landmarks_true;
  for (int l = 0; l < L; ++l)
    for (int i = 0; i < 3; ++i) m->landmarks[3 * l + i] += cfg->landmark_noise_m * rng.normal();
  // Gauge of a visual-only problem (optimization-problem.cc:96-176): the
  // first vertex is constant (its initial value is the true one) and, for
  // the scale, the first landmark stored in it.
  m->vertex_constant.assign(P, 0);
  m->landmark_constant.assign(L, 0);
  for (int i = 0; i < 7; ++i) m->vertices[i] = m->vertices_true[i];
  m->vertex_constant[0] = 1;
  if (global) {
    for (int i = 0; i < 7; ++i) m->vertices[7 * K + i] = m->vertices_true[7 * K + i];
    m->vertex_constant[K] = 1;
  } else if (cfg->fix_scale_landmark) {
    for (int l = 0; l < L; ++l)
      if (m->store_vertex[l] == 0) {
        m->landmark_constant[l] = 1;
        for (int i = 0; i < 3; ++i) m->landmarks[3 * l + i] = m->landmarks_true[3 * l + i];
        break;
      }
  }
  *out = m;
  return MLBA_OK;
}

void mlba_synth_free(mlba_synth_map* m) { delete m; }

int32_t mlba_synth_describe(const mlba_synth_map* m, mlba_problem_desc* d) {
  if (!m || !d) return MLBA_EINVAL;
  std::memset(d, 0, sizeof(*d));
  d->num_vertices = m->P;
  d->num_landmarks = m->L;
  d->num_observations = m->N;
  d->num_cameras = 1;
  d->num_missions = (int32_t)(m->baseframes.size() / 7);
  d->vertex_q_IM__M_p_MI = m->vertices.data();
  d->baseframe_q_GM__G_p_GM = m->baseframes.data();
  d->landmark_p_B = m->landmarks.data();
  d->camera_q_CI__C_p_CI = m->camera.data();
  d->camera_intrinsics = m->intrinsics.data();
  d->camera_distortion = m->distortion.data();
  d->camera_distortion_type = m->dist_type.data();
  d->camera_resolution = m->resolution.data();
  d->vertex_mission = m->vertex_mission.data();
  d->landmark_store_vertex = m->store_vertex.data();
  d->obs_vertex = m->obs_vertex.data();
  d->obs_landmark = m->obs_landmark.data();
  d->obs_camera = m->obs_camera.data();
  d->obs_keypoint = m->obs_kp.data();
  d->obs_sigma = m->obs_sigma.data();
  d->vertex_constant = m->vertex_constant.data();
  d->landmark_constant = m->landmark_constant.data();
  return MLBA_OK;
}

const double* mlba_synth_true_vertices(const mlba_synth_map* m) { return m ? m->vertices_true.data() : nullptr; }
const double* mlba_synth_true_landmarks(const mlba_synth_map* m) { return m ? m->landmarks_true.data() : nullptr; }

}  // extern "C"
