// Shared declarations of the gfx950 kernels and the C-ABI layer.
//
// Everything here is compiled with -ffp-contract=off: the FP32 hypothesis and
// inlier arithmetic is written operation by operation (explicit fmaf where a
// fused multiply-add is meant) so that results do not depend on the compiler's
// contraction choices.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "sloam_gpu.h"

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct SloamTimer {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  double ms = 0.0;
  int64_t launches = 0;
};

struct sloam_graph;

struct sloam_ctx {
  int device = 0;
  // graphs captured on this context and not yet destroyed: destroying the
  // context releases their HIP graphs and leaves them orphaned (a later
  // sloam_graph_destroy only frees the handle)
  std::vector<sloam_graph*> graphs;
  hipStream_t stream = nullptr;
  std::string last_error;
  bool timing = false;
  std::map<std::string, SloamTimer> timers;
  std::vector<hipEvent_t> event_pool;
  // grow-on-demand scratch buffers, one per slot
  std::vector<std::pair<void*, size_t>> scratch;
  // per slot: number of live captured graphs (sloam_graph, front-end graphs)
  // whose kernels hold the slot's address; a held slot never moves
  std::vector<int> scratch_pins;
  // slots requested since the current capture began (bit per slot)
  uint32_t capture_slots = 0;
  // sloam_graph_begin .. sloam_graph_end: the stream is being captured (no
  // synchronisation or allocation may happen); timing is off meanwhile
  bool capturing = false;
  bool timing_before_capture = false;
};

// Scratch buffer `slot` (0..31) of at least `bytes` bytes (reallocated when
// too small; the previous contents are not kept).  Callers own a slot for the
// duration of one entry point.  A slot used inside a capture is held by the
// captured graph until the graph is destroyed; a request that would move a
// held slot fails with SLOAM_ERR_INVALID_ARG instead of freeing memory the
// graph still writes.
int sloam_ctx_scratch(sloam_ctx* ctx, int slot, size_t bytes, void** out);
// Adds `delta` (+1 / -1) to the hold count of every slot in the bit set.
void sloam_ctx_pin_slots(sloam_ctx* ctx, uint32_t slots, int delta);

// Finds this device's address of the scratch array of
// sloam_project_partition_noalloc (once per device; the device is current).
int sloam_project_noalloc_init(sloam_ctx* ctx);

// Pose solves whose problem count is in device memory (*d_n_problems <=
// max_problems): LM state set-up and the fused LM kernel of
// sloam_pose_plan_exec over a work list built on the device (d_work: int4
// {problem, first residual, end residual, 0} per chunk of at most 2048
// residuals, d_first: first chunk of each problem).  d_state holds
// max_problems LM states (sloam_pose_state_bytes each); d_cyl_tab / d_pl_tab
// the landmark tables (per cylinder / plane: 32 / 16 bytes at cfg->precision
// 0, 64 / 32 bytes at precision 1).
size_t sloam_pose_state_bytes();
int sloam_pose_fused_device(sloam_ctx* ctx, int max_problems, const int32_t* d_n_problems,
                            const float* d_points, const int32_t* d_work, const int32_t* d_first,
                            const double* d_cyl, const int32_t* d_cyl_off, const double* d_planes,
                            const int32_t* d_pl_off, const double* d_pose_in,
                            const int32_t* d_free_mask, const sloam_pose_cfg* cfg, void* d_state,
                            void* d_cyl_tab, void* d_pl_tab, double* d_pose_out,
                            double* d_params_out, int32_t* d_status, double* d_cost);

// Records a HIP failure in the context and returns SLOAM_ERR_HIP.
int sloam_record_hip_error(sloam_ctx* ctx, hipError_t e, const char* what);

#define SLOAM_HIP(ctx, call)                                            \
  do {                                                                  \
    hipError_t e_ = (call);                                             \
    if (e_ != hipSuccess) return sloam_record_hip_error((ctx), e_, #call); \
  } while (0)

// Brackets one kernel launch with events when timing is enabled.
struct SloamKernelScope {
  sloam_ctx* ctx;
  const char* name;
  hipEvent_t start = nullptr;
  SloamKernelScope(sloam_ctx* c, const char* n);
  ~SloamKernelScope();
};

// Device allocation helpers that report through the context.
int sloam_dev_alloc(sloam_ctx* ctx, void** p, size_t bytes);
void sloam_dev_free(void* p);

// Fill kernel (used instead of hipMemsetAsync so that a captured graph holds
// kernel nodes only).
int sloam_fill_u32(sloam_ctx* ctx, uint32_t* d, size_t n, uint32_t value);

// Stable LSD radix sort (radix.hip) of n (key, value) pairs by key bits
// [begin_bit, end_bit), on the context stream, without memsets (graph-capturable).
// keys / vals (nullable) hold the input; keys_tmp / vals_tmp are same-sized
// scratch; hist holds sloam_radix_hist_ints(n) ints.  Sorts by the key bits
// [begin_bit, end_bit); d_n (nullable) is a device-side count that bounds the
// sort below n (the capacity the launches are sized for).  With keys_out
// (vals_out) given, they receive the buffers that hold the result, keys or
// keys_tmp by the pass count's parity, and no copy is made; without, the
// result is copied back into keys / vals.
int sloam_radix_sort(sloam_ctx* ctx, unsigned long long* keys, int32_t* vals,
                     unsigned long long* keys_tmp, int32_t* vals_tmp, int n, int begin_bit,
                     int end_bit, const int32_t* d_n, int32_t* hist,
                     unsigned long long** keys_out = nullptr, int32_t** vals_out = nullptr);
size_t sloam_radix_hist_ints(int n);

// sloam_fit_lines over n_sets sets of which the first *d_n_sets (device,
// nullable: all) are fitted; the others are left untouched.
int sloam_fit_lines_dev(sloam_ctx* ctx, const float* d_xyz, const int32_t* d_offsets,
                        int32_t n_sets, const int32_t* d_n_sets, const sloam_line_fit_cfg* cfg,
                        float* d_coeff, int32_t* d_stats, uint8_t* d_inliers);

// ---------------------------------------------------------------------------
// counter-based random numbers: Philox4x32-10
// J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11.  Round multipliers and Weyl key
// increments are the published constants of the 4x32 variant.
// ---------------------------------------------------------------------------
struct SlU4 {
  uint32_t x, y, z, w;
};

__host__ __device__ inline void sl_mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

__host__ __device__ inline SlU4 sl_philox4x32_10(SlU4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    sl_mulhilo(0xD2511F53u, c.x, &hi0, &lo0);
    sl_mulhilo(0xCD9E8D57u, c.z, &hi1, &lo1);
    SlU4 o;
    o.x = hi1 ^ c.y ^ k0;
    o.y = lo1;
    o.z = hi0 ^ c.w ^ k1;
    o.w = lo0;
    c = o;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Uniform integer in [0, m) from a 32-bit word: floor(x * m / 2^32).
__host__ __device__ inline uint32_t sl_bounded(uint32_t x, uint32_t m) {
  return (uint32_t)(((uint64_t)x * (uint64_t)m) >> 32);
}

// Three distinct indices in [0, n), n >= 3, from three random words.
__host__ __device__ inline void sl_draw3(uint32_t x0, uint32_t x1, uint32_t x2, uint32_t n,
                                         int32_t* i0, int32_t* i1, int32_t* i2) {
  uint32_t a = sl_bounded(x0, n);
  uint32_t b = sl_bounded(x1, n - 1u);
  if (b >= a) b += 1u;
  uint32_t lo = a < b ? a : b;
  uint32_t hi = a < b ? b : a;
  uint32_t c = sl_bounded(x2, n - 2u);
  if (c >= lo) c += 1u;
  if (c >= hi) c += 1u;
  *i0 = (int32_t)a;
  *i1 = (int32_t)b;
  *i2 = (int32_t)c;
}

// ---------------------------------------------------------------------------
// FP64 symmetric 3x3 eigen-solver (cyclic Jacobi).  Only +,-,*,/,sqrt, which
// are correctly rounded on gfx950 and on the host, so the result is the same
// on both for the same input.
// a = {a00, a01, a02, a11, a12, a22}; on return w = eigenvalues and the
// columns of V (row-major V[3*r + c]) the eigenvectors.
// ---------------------------------------------------------------------------
__host__ __device__ inline void sl_jacobi_rotate(double A[3][3], double V[3][3], int p, int q) {
  const double apq = A[p][q];
  if (apq == 0.0) return;
  const double theta = 0.5 * (A[q][q] - A[p][p]) / apq;
  double t;
  const double at = theta < 0.0 ? -theta : theta;
  if (at > 1.0e100) {
    t = 0.5 / theta;
  } else {
    t = 1.0 / (at + sqrt(1.0 + theta * theta));
    if (theta < 0.0) t = -t;
  }
  const double c = 1.0 / sqrt(1.0 + t * t);
  const double s = t * c;
  const double tau = s / (1.0 + c);
  const double h = t * apq;
  A[p][p] = A[p][p] - h;
  A[q][q] = A[q][q] + h;
  A[p][q] = 0.0;
  A[q][p] = 0.0;
  for (int r = 0; r < 3; ++r) {
    if (r == p || r == q) continue;
    const double g = A[r][p];
    const double hh = A[r][q];
    const double np = g - s * (hh + g * tau);
    const double nq = hh + s * (g - hh * tau);
    A[r][p] = np;
    A[p][r] = np;
    A[r][q] = nq;
    A[q][r] = nq;
  }
  for (int r = 0; r < 3; ++r) {
    const double g = V[r][p];
    const double hh = V[r][q];
    V[r][p] = g - s * (hh + g * tau);
    V[r][q] = hh + s * (g - hh * tau);
  }
}

__host__ __device__ inline double sl_fabs(double x) { return x < 0.0 ? -x : x; }

__host__ __device__ inline void sl_eig3(const double a[6], double w[3], double V[3][3]) {
  double A[3][3] = {{a[0], a[1], a[2]}, {a[1], a[3], a[4]}, {a[2], a[4], a[5]}};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) V[r][c] = (r == c) ? 1.0 : 0.0;
  const double scale = sl_fabs(a[0]) + sl_fabs(a[3]) + sl_fabs(a[5]) +
                       2.0 * (sl_fabs(a[1]) + sl_fabs(a[2]) + sl_fabs(a[4]));
  for (int sweep = 0; sweep < 32; ++sweep) {
    const double off = sl_fabs(A[0][1]) + sl_fabs(A[0][2]) + sl_fabs(A[1][2]);
    if (!(off > 1.0e-17 * scale)) break;
    sl_jacobi_rotate(A, V, 0, 1);
    sl_jacobi_rotate(A, V, 0, 2);
    sl_jacobi_rotate(A, V, 1, 2);
  }
  w[0] = A[0][0];
  w[1] = A[1][1];
  w[2] = A[2][2];
}

// Flips v so that its largest-magnitude component is positive (lowest index
// on ties).  Written with selects only: no dynamically indexed arrays.
__host__ __device__ inline void sl_sign_normalise(double v[3]) {
  const double a0 = sl_fabs(v[0]), a1 = sl_fabs(v[1]), a2 = sl_fabs(v[2]);
  const bool pick1 = a1 > a0;
  const double best01 = pick1 ? a1 : a0;
  const double val01 = pick1 ? v[1] : v[0];
  const double lead = (a2 > best01) ? v[2] : val01;
  if (lead < 0.0) {
    v[0] = -v[0];
    v[1] = -v[1];
    v[2] = -v[2];
  }
}

// Eigenvector column k of V, sign-normalised (see sl_sign_normalise).
__host__ __device__ inline void sl_eigvec(const double V[3][3], int k, double out[3]) {
  out[0] = k == 0 ? V[0][0] : (k == 1 ? V[0][1] : V[0][2]);
  out[1] = k == 0 ? V[1][0] : (k == 1 ? V[1][1] : V[1][2]);
  out[2] = k == 0 ? V[2][0] : (k == 1 ? V[2][1] : V[2][2]);
  sl_sign_normalise(out);
}

__host__ __device__ inline int sl_argmax3(const double w[3]) {
  int k = 0;
  if (w[1] > w[k]) k = 1;
  if (w[2] > w[k]) k = 2;
  return k;
}

__host__ __device__ inline int sl_argmin3(const double w[3]) {
  int k = 0;
  if (w[1] < w[k]) k = 1;
  if (w[2] < w[k]) k = 2;
  return k;
}

// Orthonormal basis (e1, e2) of the plane orthogonal to the unit vector a:
// e1 = normalise(r - (r.a) a) with r the coordinate axis least aligned with a
// (lowest index on ties), e2 = a x e1.  Selects only (no dynamic indexing).
__host__ __device__ inline void sl_basis(const double a[3], double e1[3], double e2[3]) {
  const double a0 = sl_fabs(a[0]), a1 = sl_fabs(a[1]), a2 = sl_fabs(a[2]);
  const bool k1 = a1 < a0;
  const double best01 = k1 ? a1 : a0;
  const bool k2 = a2 < best01;
  const int k = k2 ? 2 : (k1 ? 1 : 0);
  const double d = k == 0 ? a[0] : (k == 1 ? a[1] : a[2]);
  double x = (k == 0 ? 1.0 : 0.0) - d * a[0];
  double y = (k == 1 ? 1.0 : 0.0) - d * a[1];
  double z = (k == 2 ? 1.0 : 0.0) - d * a[2];
  const double nrm = sqrt(x * x + y * y + z * z);
  x = x / nrm;
  y = y / nrm;
  z = z / nrm;
  e1[0] = x;
  e1[1] = y;
  e1[2] = z;
  e2[0] = a[1] * z - a[2] * y;
  e2[1] = a[2] * x - a[0] * z;
  e2[2] = a[0] * y - a[1] * x;
}

// ---------------------------------------------------------------------------
// wave / block reductions
// ---------------------------------------------------------------------------
__device__ inline double sl_wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ inline long long sl_wave_sum_i64(long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
