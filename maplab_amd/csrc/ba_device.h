// ba_device.h -- device-side data layout of the visual BA problem in HBM and
// the kernel entry points (ba_kernels.hip). Everything here is internal to
// libmaplab_ba.so; the public ABI is include/maplab_ba.h.
#ifndef MAPLAB_AMD_BA_DEVICE_H_
#define MAPLAB_AMD_BA_DEVICE_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlba {

constexpr int kObsTile = 256;        // observations per linearization tile
constexpr int kObsRecord = 20;       // doubles per observation record: r(2) Jl(2x3) Jo(2x6)

// The record with FP32 Jacobians (residual kept in FP64): 88 bytes.
struct Rec32 {
  double r[2];
  float J[18];  // Jl (2x3) then Jo (2x6)
};

// Read access to observation record x: entry k as in the FP64 layout
// (0..1 residual, 2..7 Jl, 8..19 Jo).
template <bool kF32>
struct RecView;
template <>
struct RecView<false> {
  const double* p;
  __host__ __device__ RecView(const void* base, size_t x) : p(static_cast<const double*>(base) + kObsRecord * x) {}
  __host__ __device__ double operator[](int k) const { return p[k]; }
};
template <>
struct RecView<true> {
  const Rec32* p;
  __host__ __device__ RecView(const void* base, size_t x) : p(static_cast<const Rec32*>(base) + x) {}
  __host__ __device__ double operator[](int k) const { return k < 2 ? p->r[k] : static_cast<double>(p->J[k - 2]); }
};
constexpr int kLmAcc = 18;           // V(6) V_mission(6) g(3) g_mission(3)
constexpr int kLmZ = 9;              // Z(6, upper triangle) and a copy of g(3), read together per incidence

// A batch of independent scalar reductions (k_reduce_many), 1..kMaxReduceBatch
// of them.
struct ReduceDesc {
  int32_t n, stride, offset, is_max;
  const double* parts;
  double* out;
};
constexpr int kMaxReduceBatch = 8;
struct ReduceBatch {
  ReduceDesc d[kMaxReduceBatch];
  int32_t count;
};
constexpr int kRowThreads = 256;     // threads of a reduced-camera-system row block

// Landmark flags.
constexpr uint32_t kLmActive = 1u;
constexpr uint32_t kLmConstant = 2u;
constexpr uint32_t kLmAnchorVar = 4u;   // anchor vertex variable and landmark has pose-coupled obs
constexpr uint32_t kLmSplit = 8u;       // observations span several linearization tiles

// Observation meta word: bits 0-1 error type, bit 2 observer variable,
// bits 8.. camera index.
__host__ __device__ inline int32_t obs_type(int32_t meta) { return meta & 3; }
__host__ __device__ inline bool obs_observer_var(int32_t meta) { return (meta >> 2) & 1; }
__host__ __device__ inline int32_t obs_camera(int32_t meta) { return meta >> 8; }

struct DevCamera {
  double R_C_I[9];
  double p_C_I[3];
  double intr[4];
  double dist[4];
  double width, height;
  double xi;          // unified-projection mirror parameter (0 for pinhole)
  int32_t dist_type;
  int32_t model;      // kCameraPinhole / kCameraUnifiedProjection
};

// Read-only problem topology (observation order = landmark-sorted).
struct DevTopology {
  int32_t N, L, P, Pvar, M, C;
  const double* obs_kp;        // 2N
  const double* obs_sinv;      // N
  const int32_t* obs_vertex;   // N
  const int32_t* obs_meta;     // N
  const int32_t* obs_landmark; // N
  const int32_t* obs_col;      // N: variable index of the observer when the observation couples it, else -1
  const int32_t* lm_begin;     // L+1
  const int32_t* lm_anchor;    // L
  const uint32_t* lm_flags;    // L
  const DevCamera* cams;       // C
  const double* mission_T;     // 12M: R_G_M (JPL matrix of q_GM) and p_G_M
  const int32_t* vertex_mission; // P
  const int32_t* vertex_var;   // P: variable index or -1
  const int32_t* var_vertex;   // Pvar
  // reduced camera system
  const int32_t* row_ptr;      // Pvar+1 block row pointers of S
  const int32_t* col_var;      // nnzb: column variable index
  const int32_t* diag_pos;     // Pvar: position of the diagonal block
  const int32_t* inc_ptr;      // Pvar+1 incidence lists of each row
  const int32_t* inc;          // >=0: observation index, <0: -1-landmark (anchor)
  int32_t max_row_blocks;
};

// Device copy of the reduced camera system's work decomposition (ba_tiles.h):
// segments (in dispatch order through seg_order) with their thread parts,
// batches with their slot and pair lists, and for every upper block of S
// the partial slots (segment * kTileBlocks + thread) summed into it, in order.
struct SchurTilesDev {
  int32_t n_segs;
  const int32_t* seg_order;
  const int4* segs;
  const uint32_t* seg_thread;
  const int4* batches;
  const int2* slots;
  const uint16_t* batch_off;
  const uint16_t* pairs;
  const int32_t* uslot_ptr;
  const int32_t* uslot;
};

// Loss of the whitened residual.
struct DevLoss {
  int32_t type;      // 0 Huber (maplab), 1 Cauchy, 2 trivial
  double scale_px;   // Cauchy scale in pixels
};

}  // namespace mlba

#endif  // MAPLAB_AMD_BA_DEVICE_H_
