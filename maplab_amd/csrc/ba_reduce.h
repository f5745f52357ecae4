// ba_reduce.h -- wavefront / work-group reductions (64-lane waves).
#ifndef MAPLAB_AMD_BA_REDUCE_H_
#define MAPLAB_AMD_BA_REDUCE_H_

#include <hip/hip_runtime.h>

namespace mlba {

// Wavefront reductions over DPP lane moves (CDNA cross-lane data paths; no
// LDS round trip): two quad permutations (lanes xor 1, xor 2), the half-row
// and row mirrors (8- and 16-lane sums in every lane), then the row
// broadcasts of lane 15 into rows 1 and 3 and of lane 31 into rows 2 and 3,
// after which lane 63 holds the reduction of all 64 lanes. The result is read
// from lane 63 into every lane. All 64 lanes must be active; the combination
// order is fixed, so the result is deterministic.
template <int kCtrl, int kRowMask = 0xF>
__device__ inline double dpp_move(double v, double old) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), kCtrl, kRowMask, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), kCtrl, kRowMask, 0xF, false);
  return __hiloint2double(hi, lo);
}

template <typename Op>
__device__ inline double wave_reduce(double v, double identity, Op op) {
  v = op(v, dpp_move<0xB1>(v, identity));        // quad_perm [1, 0, 3, 2]
  v = op(v, dpp_move<0x4E>(v, identity));        // quad_perm [2, 3, 0, 1]
  v = op(v, dpp_move<0x141>(v, identity));       // row_half_mirror
  v = op(v, dpp_move<0x140>(v, identity));       // row_mirror
  v = op(v, dpp_move<0x142, 0xA>(v, identity));  // row_bcast:15
  v = op(v, dpp_move<0x143, 0xC>(v, identity));  // row_bcast:31
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

__device__ inline double wave_sum(double v) {
  return wave_reduce(v, 0.0, [](double a, double b) { return a + b; });
}
__device__ inline double wave_min(double v) {
  return wave_reduce(v, __builtin_inf(), [](double a, double b) { return fmin(a, b); });
}
__device__ inline double wave_max(double v) {
  return wave_reduce(v, -__builtin_inf(), [](double a, double b) { return fmax(a, b); });
}

// FP32 wave sum over the same DPP sequence (fixed order); result in every lane.
template <int kCtrl, int kRowMask = 0xF>
__device__ inline float dpp_move_f(float v, float old) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), kCtrl, kRowMask, 0xF, false));
}
__device__ inline float wave_sum_f(float v) {
  v += dpp_move_f<0xB1>(v, 0.0f);
  v += dpp_move_f<0x4E>(v, 0.0f);
  v += dpp_move_f<0x141>(v, 0.0f);
  v += dpp_move_f<0x140>(v, 0.0f);
  v += dpp_move_f<0x142, 0xA>(v, 0.0f);
  v += dpp_move_f<0x143, 0xC>(v, 0.0f);
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Block-wide sum of one value for a 256-thread block; result valid in thread 0.
__device__ inline double block_sum256(double v, double* lds4) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds4[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) s = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
  return s;
}

}  // namespace mlba

#endif  // MAPLAB_AMD_BA_REDUCE_H_
