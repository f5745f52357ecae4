// ba_reduce.h -- wavefront / work-group reductions (64-lane waves).
#ifndef MAPLAB_AMD_BA_REDUCE_H_
#define MAPLAB_AMD_BA_REDUCE_H_

#include <hip/hip_runtime.h>

namespace mlba {

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline double wave_min(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ inline double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
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
