// Batched point-level cylinder RANSAC for gfx950.
//
// Three kernels per batch (all on the context stream):
//   cyl_seed    one workgroup per cluster: bounding-box centre m, integer-exact
//               covariance sums, FP64 Jacobi eigen-solve -> seed axis and an
//               FP32 basis (e1, e2) of the plane orthogonal to it.
//   cyl_score   one workgroup per (cluster, hypothesis chunk), clusters largest
//               first: builds the chunk's hypotheses into LDS, streams the
//               cluster's points through LDS tiles as (u, v, s = u^2 + v^2) and
//               counts inliers with a wavefront ballot + popcount, 64 points per
//               compare, eight hypotheses per pass over a tile.
//   cyl_refit   one workgroup per cluster (one wave for small clusters): best
//               hypothesis (max count, lowest index), inlier bit mask, FP64
//               LO-RANSAC refit (Gauss-Newton on the 5-parameter geometric
//               cylinder residual with inlier re-selection).
//
// The inlier predicate is the reference's point-to-cylinder distance
// (sloam/src/objects/cylinder.cpp:196-203, Cylinder::distance(const PointT&)):
// | ||(p - c) x a|| - r | < thr for a unit axis a.  With (u, v) the coordinates
// of p in the plane orthogonal to a and (cu, cv) the circle centre,
// ||(p - c) x a||^2 = (u - cu)^2 + (v - cv)^2 = s + A u + B v + K with
// A = -2 cu, B = -2 cv, K = cu^2 + cv^2, so the test becomes
// L < fma(u, A, fma(v, B, s)) < U with L = (r - thr)^2 - K (or -inf when
// r <= thr) and U = (r + thr)^2 - K: two FMAs and two compares per pair.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <numeric>
#include <tuple>

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kHB = 8;  // hypotheses held in registers per pass over a tile
constexpr int kMaxLds = 65536;  // LDS bytes per workgroup this kernel asks for
// Refit tiers: a cluster of n points is refitted by one workgroup of
// kTierBlock[t] threads, t the first tier with n <= the tier's bound.  The
// tiers run concurrently (the context stream and its auxiliary streams).
// When the batch fills every SIMD, a cluster's time is set by its points per
// SIMD whatever the workgroup size: measured on the 2000-cluster batch, 64 /
// 256 threads split at 768 points beat three tiers with 512-thread
// workgroups and beat staging the points in LDS (which lowers occupancy).  A
// batch with fewer clusters than the device has SIMDs leaves SIMDs idle; a
// 256-thread workgroup then spreads its cluster over the four SIMDs of its CU
// (a quarter of the serial FP64 work per SIMD), so only clusters of up to
// kTierMaxSmall points keep one wave.
constexpr int kTiers = 2;
constexpr int kTierMax[kTiers] = {768, 1 << 30};
constexpr int kTierMaxSmall[kTiers] = {256, 1 << 30};
constexpr int kTierBlock[kTiers] = {64, 256};
static_assert(kTierBlock[0] == 64 && kTierBlock[kTiers - 1] == 256,
              "launch_refit instantiates the refit kernel for 64 and 256 threads");
static_assert(kTiers >= 1 && kTiers <= 3, "one stream per tier: the context stream and two auxiliary ones");

struct CylFrame {
  float m[3];
  float e1[3];
  float e2[3];
  int32_t n;
  double axis[3];
  double reserved;
};

struct HypOut {
  float A, B, L, U;  // predicate parameters
  float cu, cv, r;   // circle in the local frame
  int valid;
};

__device__ inline void local_uvs(const CylFrame& f, float x, float y, float z, float* u, float* v,
                                 float* s) {
  const float dx = x - f.m[0];
  const float dy = y - f.m[1];
  const float dz = z - f.m[2];
  const float uu = (dx * f.e1[0] + dy * f.e1[1]) + dz * f.e1[2];
  const float vv = (dx * f.e2[0] + dy * f.e2[1]) + dz * f.e2[2];
  *u = uu;
  *v = vv;
  *s = uu * uu + vv * vv;
}

// Predicate parameters of the circle (cu, cv, r) of the local plane.
__device__ inline HypOut hyp_from_circle(float cu, float cv, float r, float thr,
                                         float max_radius) {
  HypOut h;
  h.cu = cu;
  h.cv = cv;
  h.r = r;
  h.valid = (r <= max_radius) ? 1 : 0;  // false for NaN / inf (degenerate triangle)
  if (h.valid) {
    h.A = -2.0f * h.cu;
    h.B = -2.0f * h.cv;
    const float K = h.cu * h.cu + h.cv * h.cv;
    const float hi = r + thr;
    h.U = hi * hi - K;
    const float lo = r - thr;
    h.L = lo > 0.0f ? lo * lo - K : -INFINITY;
  } else {
    h.A = 0.0f;
    h.B = 0.0f;
    h.L = INFINITY;
    h.U = -INFINITY;
  }
  return h;
}

// Circumcircle of three points of the local plane and the predicate parameters.
__device__ inline HypOut make_hyp(float ua, float va, float ub, float vb, float uc, float vc,
                                  float thr, float max_radius) {
  const float bx = ub - ua;
  const float by = vb - va;
  const float cx = uc - ua;
  const float cy = vc - va;
  const float d = 2.0f * (bx * cy - by * cx);
  const float b2 = bx * bx + by * by;
  const float c2 = cx * cx + cy * cy;
  const float ox = (cy * b2 - by * c2) / d;
  const float oy = (bx * c2 - cx * b2) / d;
  const float r = sqrtf(ox * ox + oy * oy);
  return hyp_from_circle(ua + ox, va + oy, r, thr, max_radius);
}

// A proposed circle (centre in the local frame, radius) as a hypothesis.
__device__ inline HypOut proposal_hyp(const CylFrame& f, const float* prop, float thr,
                                      float max_radius) {
  float cu, cv, s;
  local_uvs(f, prop[3], prop[4], prop[5], &cu, &cv, &s);
  return hyp_from_circle(cu, cv, prop[6], thr, max_radius);
}

// Hypothesis h of cluster c: the proposal for h == 0 when the cluster has a
// valid one (prop_c != nullptr), else the circumcircle of a Philox sample.
__device__ inline HypOut hyp_for(const float* __restrict__ P, int n, int c, int h,
                                 const CylFrame& f, float thr, float max_radius, uint32_t k0,
                                 uint32_t k1, uint32_t tag, int32_t idx[3],
                                 const float* prop_c) {
  if (prop_c && h == 0) {
    idx[0] = idx[1] = idx[2] = -1;
    return proposal_hyp(f, prop_c, thr, max_radius);
  }
  SlU4 ctr = {(uint32_t)h, (uint32_t)c, tag, 0u};
  const SlU4 x = sl_philox4x32_10(ctr, k0, k1);
  sl_draw3(x.x, x.y, x.z, (uint32_t)n, &idx[0], &idx[1], &idx[2]);
  float u[3], v[3], s;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float* q = P + 3 * (size_t)idx[k];
    local_uvs(f, q[0], q[1], q[2], &u[k], &v[k], &s);
  }
  return make_hyp(u[0], v[0], u[1], v[1], u[2], v[2], thr, max_radius);
}

__device__ inline bool is_inlier(float u, float v, float s, const HypOut& h) {
  const float t = __builtin_fmaf(u, h.A, __builtin_fmaf(v, h.B, s));
  return (t > h.L) && (t < h.U);
}

// ---------------------------------------------------------------------------
// seed: bounding-box centre, integer covariance, FP64 Jacobi, basis
// ---------------------------------------------------------------------------
// prop (nullable): proposals (SLOAM_PN_PROPOSAL_STRIDE floats per cluster);
// a valid proposal's circle, taken into the cluster's seed frame, is
// hypothesis 0.  active (nullable): clusters with active[c] == 0 are skipped.
__device__ inline const float* cluster_prop(const float* prop, int c) {
  if (!prop) return nullptr;
  const float* q = prop + (size_t)c * SLOAM_PN_PROPOSAL_STRIDE;
  return q[7] > 0.0f ? q : nullptr;
}

__global__ __launch_bounds__(kBlock) void cyl_seed_kernel(const float* __restrict__ xyz,
                                                          const int32_t* __restrict__ offsets,
                                                          int n_clusters,
                                                          const int32_t* __restrict__ active,
                                                          CylFrame* __restrict__ frames) {
  __shared__ float red_f[kWaves][6];
  __shared__ long long red_i[kWaves][9];
  const int c = blockIdx.x;
  if (c >= n_clusters) return;
  if (active && !active[c]) return;
  const int off = offsets[c];
  const int n = offsets[c + 1] - off;
  const float* P = xyz + 3 * (size_t)off;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;

  float mn[3] = {INFINITY, INFINITY, INFINITY};
  float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = threadIdx.x; i < n; i += kBlock) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float q = P[3 * (size_t)i + k];
      mn[k] = fminf(mn[k], q);
      mx[k] = fmaxf(mx[k], q);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      mn[k] = fminf(mn[k], __shfl_xor(mn[k], o, 64));
      mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], o, 64));
    }
  }
  if (lane == 0) {
    for (int k = 0; k < 3; ++k) {
      red_f[wave][k] = mn[k];
      red_f[wave][3 + k] = mx[k];
    }
  }
  __syncthreads();
  float m[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float a = red_f[0][k], b = red_f[0][3 + k];
    for (int w = 1; w < kWaves; ++w) {
      a = fminf(a, red_f[w][k]);
      b = fmaxf(b, red_f[w][3 + k]);
    }
    m[k] = n > 0 ? 0.5f * (a + b) : 0.0f;
  }

  // integer-exact sums of q = rint((p - m) * 2^16), |q| < 2^23
  long long s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < n; i += kBlock) {
    long long q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double d = ((double)P[3 * (size_t)i + k] - (double)m[k]) * 65536.0;
      d = rint(d);
      if (d > 8388607.0) d = 8388607.0;
      if (d < -8388607.0) d = -8388607.0;
      q[k] = (long long)d;
    }
    s[0] += q[0];
    s[1] += q[1];
    s[2] += q[2];
    s[3] += q[0] * q[0];
    s[4] += q[0] * q[1];
    s[5] += q[0] * q[2];
    s[6] += q[1] * q[1];
    s[7] += q[1] * q[2];
    s[8] += q[2] * q[2];
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) s[k] = sl_wave_sum_i64(s[k]);
  if (lane == 0)
    for (int k = 0; k < 9; ++k) red_i[wave][k] = s[k];
  __syncthreads();
  if (threadIdx.x != 0) return;
  long long t[9];
  for (int k = 0; k < 9; ++k) {
    t[k] = red_i[0][k];
    for (int w = 1; w < kWaves; ++w) t[k] += red_i[w][k];
  }
  CylFrame f;
  f.m[0] = m[0];
  f.m[1] = m[1];
  f.m[2] = m[2];
  f.n = n;
  f.reserved = 0.0;
  double axis[3] = {0.0, 0.0, 1.0};
  if (n >= 2) {
    const double dn = (double)n;
    const double S0 = (double)t[0], S1 = (double)t[1], S2 = (double)t[2];
    double a6[6];
    a6[0] = (double)t[3] - S0 * S0 / dn;
    a6[1] = (double)t[4] - S0 * S1 / dn;
    a6[2] = (double)t[5] - S0 * S2 / dn;
    a6[3] = (double)t[6] - S1 * S1 / dn;
    a6[4] = (double)t[7] - S1 * S2 / dn;
    a6[5] = (double)t[8] - S2 * S2 / dn;
    double w[3], V[3][3];
    sl_eig3(a6, w, V);
    sl_eigvec(V, sl_argmax3(w), axis);
  }
  double e1[3], e2[3];
  sl_basis(axis, e1, e2);
  for (int k = 0; k < 3; ++k) {
    f.axis[k] = axis[k];
    f.e1[k] = (float)e1[k];
    f.e2[k] = (float)e2[k];
  }
  frames[c] = f;
}

// ---------------------------------------------------------------------------
// score: hypotheses x points, ballot/popcount inlier counts
// work item = {cluster, first hypothesis, hypothesis count, unused}
// LDS: hyp params [hcap] float4 | counts [hcap] int | u,v,s [tile] float
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void cyl_score_kernel(
    const float* __restrict__ xyz, const int32_t* __restrict__ offsets,
    const CylFrame* __restrict__ frames, const int4* __restrict__ work, int n_hyp, int hcap,
    int tile, float thr, float max_radius, uint32_t k0, uint32_t k1, uint32_t tag,
    const float* __restrict__ prop, const int32_t* __restrict__ active,
    int32_t* __restrict__ counts, int32_t* __restrict__ samples) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float4* hp = reinterpret_cast<float4*>(smem_raw);
  int* cnt = reinterpret_cast<int*>(hp + hcap);
  float* su = reinterpret_cast<float*>(cnt + hcap);
  float* sv = su + tile;
  float* ss = sv + tile;

  const int4 w = work[blockIdx.x];
  const int c = w.x, h0 = w.y, hc = w.z;
  if (c < 0) return;  // padding of the per-XCD work order
  if (active && !active[c]) return;
  const float* pc = cluster_prop(prop, c);
  const int off = offsets[c];
  const int n = offsets[c + 1] - off;
  const float* P = xyz + 3 * (size_t)off;
  const CylFrame f = frames[c];

  for (int t = threadIdx.x; t < hcap; t += kBlock) {
    float4 prm = make_float4(0.f, 0.f, INFINITY, -INFINITY);
    if (t < hc && n >= 3) {
      int32_t idx[3];
      const HypOut h = hyp_for(P, n, c, h0 + t, f, thr, max_radius, k0, k1, tag, idx, pc);
      prm = make_float4(h.A, h.B, h.L, h.U);
      if (samples) {
        int32_t* o = samples + 3 * ((size_t)c * n_hyp + h0 + t);
        o[0] = idx[0];
        o[1] = idx[1];
        o[2] = idx[2];
      }
    } else if (t < hc && samples) {
      int32_t* o = samples + 3 * ((size_t)c * n_hyp + h0 + t);
      o[0] = -1;
      o[1] = -1;
      o[2] = -1;
    }
    hp[t] = prm;
    cnt[t] = 0;
  }

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int hpw = hcap / kWaves;  // hypotheses per wave (multiple of kHB)

  for (int t0 = 0; t0 < n; t0 += tile) {
    const int tn = min(tile, n - t0);
    const int tn64 = (tn + 63) & ~63;
    __syncthreads();
    for (int i = threadIdx.x; i < tn64; i += kBlock) {
      float u = 0.f, v = 0.f, s = NAN;
      if (i < tn) {
        const float* q = P + 3 * (size_t)(t0 + i);
        local_uvs(f, q[0], q[1], q[2], &u, &v, &s);
      }
      su[i] = u;
      sv[i] = v;
      ss[i] = s;
    }
    __syncthreads();
    for (int hb = 0; hb < hpw; hb += kHB) {
      const int hbase = wave * hpw + hb;
      if (hbase >= hc) break;
      float A[kHB], B[kHB], L[kHB], U[kHB];
      int cc[kHB];
#pragma unroll
      for (int j = 0; j < kHB; ++j) {
        const float4 q = hp[hbase + j];
        A[j] = q.x;
        B[j] = q.y;
        L[j] = q.z;
        U[j] = q.w;
        cc[j] = 0;
      }
      // uniform trip count: i0 lives in a scalar register, the two compares
      // write scalar lane masks that are and-ed and popcounted on the SALU
      for (int i0 = 0; i0 < tn64; i0 += 64) {
        const int i = i0 + lane;
        const float u = su[i];
        const float v = sv[i];
        const float s = ss[i];
#pragma unroll
        for (int j = 0; j < kHB; ++j) {
          const float t = __builtin_fmaf(u, A[j], __builtin_fmaf(v, B[j], s));
          const uint64_t gt = __builtin_amdgcn_ballot_w64(t > L[j]);
          const uint64_t lt = __builtin_amdgcn_ballot_w64(t < U[j]);
          cc[j] += __builtin_popcountll(gt & lt);
        }
      }
      if (lane == 0) {
#pragma unroll
        for (int j = 0; j < kHB; ++j) cnt[hbase + j] += cc[j];
      }
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < hc; t += kBlock) counts[(size_t)c * n_hyp + h0 + t] = cnt[t];
}

struct CylModel {
  double a[3], c[3], r;
};

// As cyl_dist, and 1 / ||w|| from one reciprocal square root refined by two
// Newton steps (no FP64 divide in the Gauss-Newton loop). A point on the axis
// (||w|| = 0) keeps cyl_dist's answer: distance -r and a zero reciprocal.
__device__ inline double cyl_dist_inv(const CylModel& M, const double p[3], double w[3],
                                      double* h, double* inv) {
  const double q[3] = {p[0] - M.c[0], p[1] - M.c[1], p[2] - M.c[2]};
  const double hh = q[0] * M.a[0] + q[1] * M.a[1] + q[2] * M.a[2];
  w[0] = q[0] - hh * M.a[0];
  w[1] = q[1] - hh * M.a[1];
  w[2] = q[2] - hh * M.a[2];
  *h = hh;
  const double s2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  if (!(s2 > 0.0)) {
    *inv = 0.0;
    return -M.r;
  }
  double y = __builtin_amdgcn_rsq(s2);
  const double hs = 0.5 * s2;
  y = y * __builtin_fma(-hs * y, y, 1.5);
  y = y * __builtin_fma(-hs * y, y, 1.5);
  *inv = y;
  return s2 * y - M.r;
}

__device__ inline double cyl_dist(const CylModel& M, const double p[3], double w[3], double* h) {
  const double q[3] = {p[0] - M.c[0], p[1] - M.c[1], p[2] - M.c[2]};
  const double hh = q[0] * M.a[0] + q[1] * M.a[1] + q[2] * M.a[2];
  w[0] = q[0] - hh * M.a[0];
  w[1] = q[1] - hh * M.a[1];
  w[2] = q[2] - hh * M.a[2];
  *h = hh;
  return sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) - M.r;
}

__device__ inline void write_invalid(double* out, double k) {
  for (int i = 0; i < 6; ++i) out[i] = 0.0;
  out[5] = 1.0;
  out[6] = -1.0;
  out[7] = k;
}

// Cholesky solve of a symmetric positive definite N x N system, fully
// unrolled so that every index is static (no private memory).
template <int N>
__device__ inline bool cholesky_solve(const double (&A)[N][N], const double (&b)[N],
                                      double (&x)[N]) {
  double L[N][N];
  bool ok = true;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    double s = A[j][j];
#pragma unroll
    for (int k = 0; k < j; ++k) s = s - L[j][k] * L[j][k];
    ok = ok && (s > 0.0);
    const double d = sqrt(s);
    L[j][j] = d;
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      double t = A[i][j];
#pragma unroll
      for (int k = 0; k < j; ++k) t = t - L[i][k] * L[j][k];
      L[i][j] = t / d;
    }
  }
  double y[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double t = b[i];
#pragma unroll
    for (int k = 0; k < i; ++k) t = t - L[i][k] * y[k];
    y[i] = t / L[i][i];
  }
#pragma unroll
  for (int i = N - 1; i >= 0; --i) {
    double t = y[i];
#pragma unroll
    for (int k = i + 1; k < N; ++k) t = t - L[k][i] * x[k];
    x[i] = t / L[i][i];
  }
  return ok;
}

// Sums K doubles over a block of BLOCK threads; every thread gets the totals.
// lds holds (BLOCK / 64 + 1) * K doubles: the wave sums, then the totals,
// added in wave order by K threads.
template <int BLOCK, int K>
__device__ inline void bsum(double (&v)[K], double* lds) {
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = sl_wave_sum(v[k]);
  if (BLOCK == 64) return;
  constexpr int W = BLOCK / 64;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) lds[wave * K + k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int k = threadIdx.x;
    double t = lds[k];
#pragma unroll
    for (int w = 1; w < W; ++w) t += lds[w * K + k];
    lds[W * K + k] = t;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = lds[W * K + k];
}

// Sum over the workgroup of K doubles per thread, every thread getting the
// totals, through LDS: each thread's values are stored, P threads per value
// add BLOCK / P threads' values each in a fixed order, and K threads add the
// P partials (the wave shuffles of bsum move each double through two LDS
// permutes per step, six steps per value).
template <int BLOCK, int K>
struct BsumLds {
  static constexpr int P = (BLOCK / K) < 16 ? (BLOCK / K) : 16;  // threads per value
  static constexpr int L = (BLOCK + P - 1) / P;                   // values per thread
  double x[K][BLOCK + 1];
  double part[K][P];
  double tot[K];
};

template <int BLOCK, int K>
__device__ inline void bsum_lds(double (&v)[K], BsumLds<BLOCK, K>& r) {
  constexpr int P = BsumLds<BLOCK, K>::P;
  constexpr int L = BsumLds<BLOCK, K>::L;
  static_assert(P >= 1 && K * P <= BLOCK, "bsum_lds layout");
  const int t = threadIdx.x;
  __syncthreads();  // r may still be read from the previous call
#pragma unroll
  for (int k = 0; k < K; ++k) r.x[k][t] = v[k];
  __syncthreads();
  if (t < K * P) {
    const int k = t % K, p = t / K;
    const int b = p * L, e = min(BLOCK, b + L);
    double s = 0.0;
    for (int j = b; j < e; ++j) s += r.x[k][j];
    r.part[k][p] = s;
  }
  __syncthreads();
  if (t < K) {
    double s = r.part[t][0];
#pragma unroll
    for (int p = 1; p < P; ++p) s += r.part[t][p];
    r.tot[t] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = r.tot[k];
}

// ---------------------------------------------------------------------------
// refit: best hypothesis, mask, FP64 LO-RANSAC cylinder refit.
// One workgroup per cluster, its size by the cluster's size (refit tiers).
// Rounds of Gauss-Newton steps on the 5-parameter geometric cylinder residual
//   d_i = ||(p_i - c) - ((p_i - c).a) a|| - r
// over the round's inliers: the FP32 inliers of the best hypothesis in round
// 0, |d_i| < thr under the model at the start of the round afterwards.  The
// membership of a round is computed in its first pass and kept as a bit mask
// in LDS.  Output record: root[3], ray[3], radius, inliers.
// ---------------------------------------------------------------------------
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void cyl_refit_kernel(
    const float* __restrict__ xyz, const int32_t* __restrict__ offsets,
    const CylFrame* __restrict__ frames, const int32_t* __restrict__ counts,
    const int32_t* __restrict__ list, int n_hyp, float thr, float max_radius, uint32_t k0,
    uint32_t k1, uint32_t tag, int refit_rounds, int refit_iters,
    const float* __restrict__ prop, const int32_t* __restrict__ active,
    const int32_t* __restrict__ alt_flags, const int32_t* __restrict__ alt_counts, int alt_n_hyp,
    uint32_t alt_tag, int alt_rounds, int alt_iters, uint32_t* __restrict__ mask,
    int32_t* __restrict__ best_out, double* __restrict__ model) {
  constexpr int W = BLOCK / 64;
  __shared__ double red[(W + 1) * 21];  // bsum of up to 21 values
  __shared__ BsumLds<BLOCK, 21> red21;   // the Gauss-Newton sums
  __shared__ long long redk[W];
  __shared__ HypOut best_h;
  extern __shared__ uint32_t bits[];  // membership of the current round
  const int c = list[blockIdx.x];
  if (active && !active[c]) return;
  const int off = offsets[c];
  const int n = offsets[c + 1] - off;
  const float* P = xyz + 3 * (size_t)off;
  const CylFrame f = frames[c];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  double* out = model + (size_t)c * SLOAM_CYL_MODEL_STRIDE;
  // a flagged cluster takes its hypotheses from the alternate set (its own
  // counts, hypothesis count and sample stream, no proposal) and refits with
  // the alternate round and step counts
  const bool alt = alt_flags && alt_flags[c];
  const int32_t* cnts = alt ? alt_counts : counts;
  const int nh = alt ? alt_n_hyp : n_hyp;
  const uint32_t tg = alt ? alt_tag : tag;
  const int rounds = alt ? alt_rounds : refit_rounds;
  const int iters = alt ? alt_iters : refit_iters;
  const float* pc = alt ? nullptr : cluster_prop(prop, c);

  // 1. best hypothesis: key = count * 2^32 + (2^32 - 1 - index), max wins
  long long key = -1;
  for (int h = threadIdx.x; h < nh; h += BLOCK) {
    const long long kk = ((long long)cnts[(size_t)c * nh + h] << 32) |
                         (long long)(0xFFFFFFFFu - (uint32_t)h);
    key = kk > key ? kk : key;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const long long other = __shfl_xor(key, o, 64);
    key = other > key ? other : key;
  }
  if (BLOCK > 64) {
    if (lane == 0) redk[wave] = key;
    __syncthreads();
    key = redk[0];
#pragma unroll
    for (int w = 1; w < W; ++w) key = redk[w] > key ? redk[w] : key;
  }
  const int best = (n >= 3 && nh > 0) ? (int)(0xFFFFFFFFu - (uint32_t)(key & 0xFFFFFFFFll)) : -1;
  if (threadIdx.x == 0) {
    if (best >= 0) {
      int32_t idx[3];
      best_h = hyp_for(P, n, c, best, f, thr, max_radius, k0, k1, tg, idx, pc);
    } else {
      best_h.valid = 0;
      best_h.A = best_h.B = 0.f;
      best_h.L = INFINITY;
      best_h.U = -INFINITY;
      best_h.cu = best_h.cv = best_h.r = 0.f;
    }
    if (best_out) best_out[c] = best;
  }
  __syncthreads();
  const HypOut h = best_h;

  // 2. round-0 membership = FP32 inliers of the best hypothesis: global mask
  //    output and LDS bits (64 consecutive points per wave, 64-aligned)
  double k0cnt[1] = {0.0};
  for (int base = 0; base < n; base += BLOCK) {
    const int i = base + (int)threadIdx.x;
    bool in = false;
    if (i < n) {
      const float* q = P + 3 * (size_t)i;
      float u, v, s;
      local_uvs(f, q[0], q[1], q[2], &u, &v, &s);
      in = is_inlier(u, v, s, h);
    }
    const uint64_t b = __builtin_amdgcn_ballot_w64(in);
    if (lane == 0) {
      const int wbase = (base + wave * 64) >> 5;
      bits[wbase] = (uint32_t)b;
      bits[wbase + 1] = (uint32_t)(b >> 32);
    }
    if (in) k0cnt[0] += 1.0;
  }
  if (mask) {
    // global mask: iterate 64-aligned global indices so that each wave owns
    // two whole 32-bit words (or-ed atomically: neighbouring clusters share
    // boundary words)
    const int gend = off + n;
    for (int gb = (off & ~63); gb < gend; gb += BLOCK) {
      const int g = gb + (int)threadIdx.x;
      bool in = false;
      if (g >= off && g < gend) {
        const float* q = xyz + 3 * (size_t)g;
        float u, v, s;
        local_uvs(f, q[0], q[1], q[2], &u, &v, &s);
        in = is_inlier(u, v, s, h);
      }
      const uint64_t b = __builtin_amdgcn_ballot_w64(in);
      if (lane == 0 && b) {
        const int w0 = (gb + wave * 64) >> 5;
        if ((uint32_t)b) atomicOr(mask + w0, (uint32_t)b);
        if ((uint32_t)(b >> 32)) atomicOr(mask + w0 + 1, (uint32_t)(b >> 32));
      }
    }
  }
  bsum<BLOCK, 1>(k0cnt, red);
  if (k0cnt[0] < 3.0 || !h.valid) {
    if (threadIdx.x == 0) write_invalid(out, k0cnt[0]);
    return;
  }

  // 3. LO refit
  CylModel M;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    M.a[i] = f.axis[i];
    M.c[i] = (double)h.cu * (double)f.e1[i] + (double)h.cv * (double)f.e2[i];
  }
  M.r = (double)h.r;
  const double dthr = (double)thr;
  bool stop = false;
  for (int round = 0; round < rounds && !stop; ++round) {
    for (int it = 0; it < iters; ++it) {
      double e1[3], e2[3];
      sl_basis(M.a, e1, e2);
      double g[21];
#pragma unroll
      for (int t = 0; t < 21; ++t) g[t] = 0.0;
      const bool reselect = (round > 0 && it == 0);
      for (int base = 0; base < n; base += BLOCK) {
        const int i = base + (int)threadIdx.x;
        bool in = false;
        double dd = 0.0, hh = 0.0, w[3] = {0.0, 0.0, 0.0}, iw = 0.0;
        if (i < n) {
          const float* q = P + 3 * (size_t)i;
          const double x[3] = {(double)q[0] - (double)f.m[0], (double)q[1] - (double)f.m[1],
                               (double)q[2] - (double)f.m[2]};
          if (reselect) {
            dd = cyl_dist_inv(M, x, w, &hh, &iw);
            in = dd < dthr && dd > -dthr;
          } else {
            in = (bits[i >> 5] >> (i & 31)) & 1u;
            if (in) dd = cyl_dist_inv(M, x, w, &hh, &iw);
          }
        }
        if (reselect) {
          // bits of point i are written and read by the same wave only
          const uint64_t b = __builtin_amdgcn_ballot_w64(in);
          if (lane == 0) {
            const int wbase = (base + wave * 64) >> 5;
            bits[wbase] = (uint32_t)b;
            bits[wbase + 1] = (uint32_t)(b >> 32);
          }
        }
        if (in) {
          g[20] += 1.0;
          const double rho = dd + M.r;
          if (rho > 0.0) {
            const double inv = iw;
            const double n1 = (w[0] * e1[0] + w[1] * e1[1] + w[2] * e1[2]) * inv;
            const double n2 = (w[0] * e2[0] + w[1] * e2[1] + w[2] * e2[2]) * inv;
            const double J[5] = {-hh * n1, -hh * n2, -n1, -n2, -1.0};
            int t = 0;
#pragma unroll
            for (int a = 0; a < 5; ++a)
#pragma unroll
              for (int bb = a; bb < 5; ++bb) g[t++] += J[a] * J[bb];
#pragma unroll
            for (int a = 0; a < 5; ++a) g[15 + a] += J[a] * dd;
          }
        }
      }
      bsum_lds<BLOCK, 21>(g, red21);
      if (g[20] < 5.0) {
        stop = true;
        break;
      }
      double A[5][5], rhs[5], dl[5];
      {
        int t = 0;
#pragma unroll
        for (int a = 0; a < 5; ++a)
#pragma unroll
          for (int bb = a; bb < 5; ++bb) {
            A[a][bb] = g[t];
            A[bb][a] = g[t];
            ++t;
          }
#pragma unroll
        for (int a = 0; a < 5; ++a) rhs[a] = -g[15 + a];
      }
      if (!cholesky_solve<5>(A, rhs, dl)) break;
      double na[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) na[i] = M.a[i] + dl[0] * e1[i] + dl[1] * e2[i];
      const double nl = sqrt(na[0] * na[0] + na[1] * na[1] + na[2] * na[2]);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        M.a[i] = na[i] / nl;
        M.c[i] = M.c[i] + dl[2] * e1[i] + dl[3] * e2[i];
      }
      M.r = M.r + dl[4];
      if (!(M.r > 0.0)) {
        stop = true;
        break;
      }
    }
  }

  // 4. final inliers under the final model: count + centroid
  double fin[4] = {0.0, 0.0, 0.0, 0.0};
  if (M.r > 0.0) {
    for (int i = threadIdx.x; i < n; i += BLOCK) {
      const float* q = P + 3 * (size_t)i;
      const double x[3] = {(double)q[0] - (double)f.m[0], (double)q[1] - (double)f.m[1],
                           (double)q[2] - (double)f.m[2]};
      double w[3], hh;
      const double dd = cyl_dist(M, x, w, &hh);
      if (dd < dthr && dd > -dthr) {
        fin[0] += x[0];
        fin[1] += x[1];
        fin[2] += x[2];
        fin[3] += 1.0;
      }
    }
  }
  bsum<BLOCK, 4>(fin, red);
  const double k = fin[3];
  if (threadIdx.x != 0) return;
  if (!(M.r > 0.0) || k < 3.0) {
    write_invalid(out, k);
    return;
  }
  const double ht = (fin[0] / k - M.c[0]) * M.a[0] + (fin[1] / k - M.c[1]) * M.a[1] +
                    (fin[2] / k - M.c[2]) * M.a[2];
  double ray[3] = {M.a[0], M.a[1], M.a[2]};
  sl_sign_normalise(ray);
  for (int i = 0; i < 3; ++i) {
    out[i] = (double)f.m[i] + M.c[i] + ht * M.a[i];
    out[3 + i] = ray[i];
  }
  out[6] = M.r;
  out[7] = k;
}

}  // namespace

// ---------------------------------------------------------------------------
// plans
// ---------------------------------------------------------------------------
struct sloam_cyl_plan {
  sloam_ctx* ctx = nullptr;
  sloam_cyl_fit_cfg cfg{};
  int32_t n_clusters = 0;
  int64_t n_points = 0;
  int hcap = 0;
  int tile = 0;
  int n_work = 0;
  int4* d_work = nullptr;
  int32_t* d_lists = nullptr;   // small-cluster list then large-cluster list
  int tier_n[kTiers] = {};      // clusters per refit tier
  int tier_off[kTiers] = {};    // first list entry of each tier
  int tier_bits[kTiers] = {};   // LDS bytes of the membership bits
  CylFrame* d_frames = nullptr;
  int32_t* d_counts = nullptr;  // used when the caller passes no counts buffer
  int n_xcd = 1;
  int n_simd = 1024;  // SIMDs of the device (4 per CU)
  // d_work and d_lists share one allocation (the work list, then the refit
  // lists), uploaded in one copy; allocated sizes (bytes for the tables,
  // elements otherwise) only grow across sloam_cyl_plan_reset
  void* d_tables = nullptr;
  size_t cap_tables = 0, cap_frames = 0, cap_counts = 0;
  // pinned staging of the work and refit lists (bytes): their uploads are
  // asynchronous on the context stream, ahead of the runs that read them
  void* h_stage = nullptr;
  size_t cap_stage = 0;
};

static int round_up(int x, int m) { return (x + m - 1) / m * m; }

// active[c] = 1 when cluster c's model keeps fewer than min_ratio of its
// points as inliers (or is invalid); *n_active counts them
// Launches the refit tiers of plan p: concurrently, the larger tiers on
// auxiliary streams forked from and joined back into the context stream.
// alt_flags (nullable) marks the clusters that take their best hypothesis
// from alt_counts (alt_n_hyp hypotheses of sample stream alt_tag, no
// proposal) instead of counts.
static int launch_refit(sloam_cyl_plan* p, const float* d_xyz, const int32_t* d_offsets,
                 const int32_t* counts, const float* d_proposal, const int32_t* d_active,
                 const int32_t* alt_flags, const int32_t* alt_counts, int alt_n_hyp,
                 uint32_t alt_tag, int alt_rounds, int alt_iters, uint32_t* d_mask,
                 int32_t* d_best, double* d_model) {
  sloam_ctx* ctx = p->ctx;
  const sloam_cyl_fit_cfg& cfg = p->cfg;
  const uint32_t k0 = (uint32_t)cfg.seed;
  const uint32_t k1 = (uint32_t)(cfg.seed >> 32);
  const uint32_t tag = (uint32_t)cfg.stream_tag;
  int st = SLOAM_OK;
  for (int t = kTiers - 1; t >= 0; --t) {
    if (p->tier_n[t] == 0) continue;
    hipStream_t s = ctx->stream;
    if (t > 0) {
      st = sloam_ctx_fork(ctx, t - 1);
      if (st) return st;
      s = ctx->aux[t - 1];
    }
    const int32_t* list = p->d_lists + p->tier_off[t];
    const dim3 g(p->tier_n[t]);
    const size_t lds = p->tier_bits[t];
#define SLOAM_REFIT_ARGS                                                                      \
  d_xyz, d_offsets, p->d_frames, counts, list, cfg.n_hyp, cfg.thr, cfg.max_radius, k0, k1, tag, \
      cfg.refit_rounds, cfg.refit_iters, d_proposal, d_active, alt_flags, alt_counts, alt_n_hyp, \
      alt_tag, alt_rounds, alt_iters, d_mask, d_best, d_model
    if (kTierBlock[t] == 64)
      hipLaunchKernelGGL(cyl_refit_kernel<64>, g, dim3(64), lds, s, SLOAM_REFIT_ARGS);
    else
      hipLaunchKernelGGL(cyl_refit_kernel<256>, g, dim3(256), lds, s, SLOAM_REFIT_ARGS);
#undef SLOAM_REFIT_ARGS
    SLOAM_HIP(ctx, hipGetLastError());
  }
  for (int t = 1; t < kTiers; ++t)
    if (p->tier_n[t] > 0) {
      st = sloam_ctx_join(ctx, t - 1);
      if (st) return st;
    }
  return SLOAM_OK;
}

// Fallback flags of the learned fit: a cluster (n >= 3) is refitted by the
// plain method when its best verifying hypothesis holds fewer than
// min_ratio * n inliers.
__global__ __launch_bounds__(256) void cyl_decide_kernel(const int32_t* __restrict__ offsets,
                                                         int n_clusters,
                                                         const int32_t* __restrict__ counts,
                                                         int n_hyp, float min_ratio,
                                                         int32_t* __restrict__ flags,
                                                         int32_t* __restrict__ n_flagged) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  bool a = false;
  if (c < n_clusters) {
    const int n = offsets[c + 1] - offsets[c];
    int best = 0;
    for (int h = 0; h < n_hyp; ++h) best = max(best, counts[(size_t)c * n_hyp + h]);
    a = n >= 3 && (double)best < (double)min_ratio * (double)n;
    flags[c] = a ? 1 : 0;
  }
  const uint64_t b = __builtin_amdgcn_ballot_w64(a);
  if ((threadIdx.x & 63) == 0 && b && n_flagged) atomicAdd(n_flagged, __builtin_popcountll(b));
}

__global__ __launch_bounds__(256) void cyl_fallback_kernel(const int32_t* __restrict__ offsets,
                                                           int n_clusters,
                                                           const double* __restrict__ model,
                                                           float min_ratio,
                                                           int32_t* __restrict__ active,
                                                           int32_t* __restrict__ n_active) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  bool a = false;
  if (c < n_clusters) {
    const int n = offsets[c + 1] - offsets[c];
    const double* m = model + (size_t)c * SLOAM_CYL_MODEL_STRIDE;
    a = n >= 3 && !(m[6] > 0.0 && m[7] >= (double)min_ratio * (double)n);
    active[c] = a ? 1 : 0;
  }
  const uint64_t b = __builtin_amdgcn_ballot_w64(a);
  if ((threadIdx.x & 63) == 0 && b && n_active) atomicAdd(n_active, __builtin_popcountll(b));
}

// Grows the device buffer *ptr (element size elem) to hold n elements;
// contents are not kept.
static int grow(sloam_ctx* ctx, void** ptr, size_t* cap, size_t n, size_t elem) {
  if (n <= *cap) return SLOAM_OK;
  sloam_dev_free(*ptr);
  *ptr = nullptr;
  *cap = 0;
  const int st = sloam_dev_alloc(ctx, ptr, n * elem);
  if (st) return st;
  *cap = n;
  return SLOAM_OK;
}

// Lays the batch h_offsets (n_clusters + 1 entries) out on plan p, whose cfg,
// hcap and tile are set: the scoring work list and the refit tier lists,
// uploaded into buffers that only grow.
static int plan_layout(sloam_cyl_plan* p, const int32_t* h_offsets, int32_t n_clusters) {
  sloam_ctx* ctx = p->ctx;
  const sloam_cyl_fit_cfg& cfg = p->cfg;
  if (h_offsets[0] != 0) return SLOAM_ERR_INVALID_ARG;
  for (int i = 0; i < n_clusters; ++i) {
    const int64_t len = (int64_t)h_offsets[i + 1] - h_offsets[i];
    if (len < 0) return SLOAM_ERR_INVALID_ARG;
    if (len > (1 << 17)) return SLOAM_ERR_TOO_LARGE;  // integer-exact covariance bound
  }
  p->n_clusters = n_clusters;
  p->n_points = h_offsets[n_clusters];

  // work list: clusters by decreasing size (ties: lower index), each cluster
  // split into hypothesis chunks.  Workgroups go to the XCDs round robin
  // (workgroup b on XCD b mod n_xcd), each XCD with its own L2: clusters are
  // dealt to per-XCD bins by least load (largest first) and the launch order
  // interleaves the bins, so all chunks of a cluster run on one XCD and read
  // its points through one L2.  Bins are padded with empty items (cluster -1).
  std::vector<int> order(n_clusters);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (h_offsets[a + 1] - h_offsets[a]) > (h_offsets[b + 1] - h_offsets[b]);
  });
  const int n_xcd = p->n_xcd;
  const int chunks = (cfg.n_hyp + p->hcap - 1) / p->hcap;
  std::vector<std::vector<int4>> bins(n_xcd);
  std::vector<long long> load(n_xcd, 0);
  for (int c : order) {
    const int len = h_offsets[c + 1] - h_offsets[c];
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    for (int h0 = 0; h0 < cfg.n_hyp; h0 += p->hcap) {
      const int hc = std::min(p->hcap, cfg.n_hyp - h0);
      bins[x].push_back(make_int4(c, h0, hc, len));
    }
    load[x] += (long long)std::max(len, 1) * chunks;
  }
  size_t longest = 0;
  for (const auto& b : bins) longest = std::max(longest, b.size());
  std::vector<int4> work;
  work.reserve(longest * n_xcd);
  for (size_t k = 0; k < longest; ++k)
    for (int x = 0; x < n_xcd; ++x)
      work.push_back(k < bins[x].size() ? bins[x][k] : make_int4(-1, 0, 0, 0));
  p->n_work = (int)work.size();
  // refit lists, one per tier, each in decreasing size
  const int* tier_max = n_clusters < p->n_simd ? kTierMaxSmall : kTierMax;
  std::vector<int32_t> tier_list[kTiers];
  int tier_maxlen[kTiers] = {};
  for (int c : order) {
    const int len = h_offsets[c + 1] - h_offsets[c];
    int t = 0;
    while (t + 1 < kTiers && len > tier_max[t]) ++t;
    tier_list[t].push_back(c);
    tier_maxlen[t] = std::max(tier_maxlen[t], len);
  }
  std::vector<int32_t> lists;
  for (int t = 0; t < kTiers; ++t) {
    p->tier_n[t] = (int)tier_list[t].size();
    p->tier_off[t] = (int)lists.size();
    p->tier_bits[t] = round_up(std::max(tier_maxlen[t], 1), kTierBlock[t]) / 8;
    lists.insert(lists.end(), tier_list[t].begin(), tier_list[t].end());
  }
  const size_t wb = sizeof(int4) * work.size(), lb = sizeof(int32_t) * lists.size();
  int st = grow(ctx, &p->d_tables, &p->cap_tables, wb + lb, 1);
  if (!st) st = grow(ctx, (void**)&p->d_frames, &p->cap_frames, n_clusters, sizeof(CylFrame));
  if (!st)
    st = grow(ctx, (void**)&p->d_counts, &p->cap_counts, (size_t)n_clusters * cfg.n_hyp,
              sizeof(int32_t));
  if (st) return st;
  p->d_work = static_cast<int4*>(p->d_tables);
  p->d_lists = reinterpret_cast<int32_t*>(static_cast<char*>(p->d_tables) + wb);
  // both lists through the pinned staging buffer, in one copy (the caller has
  // made sure that no earlier upload from it is pending)
  if (wb + lb > p->cap_stage) {
    if (p->h_stage) (void)hipHostFree(p->h_stage);
    p->h_stage = nullptr;
    p->cap_stage = 0;
    const size_t sz = (wb + lb) + (wb + lb) / 2;
    hipError_t e = hipHostMalloc(&p->h_stage, sz);
    if (e != hipSuccess) return sloam_record_hip_error(ctx, e, "hipHostMalloc(plan staging)");
    p->cap_stage = sz;
  }
  char* stage = static_cast<char*>(p->h_stage);
  if (wb) std::memcpy(stage, work.data(), wb);
  if (lb) std::memcpy(stage + wb, lists.data(), lb);
  if (wb + lb)
    SLOAM_HIP(ctx, hipMemcpyAsync(p->d_tables, stage, wb + lb, hipMemcpyHostToDevice, ctx->stream));
  return SLOAM_OK;
}

extern "C" {

int sloam_cyl_plan_create(sloam_ctx* ctx, const int32_t* h_offsets, int32_t n_clusters,
                          const sloam_cyl_fit_cfg* cfg, sloam_cyl_plan** out) {
  if (!ctx || !h_offsets || !cfg || !out || n_clusters < 0) return SLOAM_ERR_INVALID_ARG;
  *out = nullptr;
  if (cfg->n_hyp < 1 || cfg->n_hyp > (1 << 20) || !(cfg->thr > 0.0f) ||
      !(cfg->max_radius > 0.0f) || cfg->refit_iters < 0 || cfg->refit_rounds < 0 || cfg->hyp_chunk < 0 ||
      cfg->tile_points < 0)
    return SLOAM_ERR_INVALID_ARG;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  sloam_cyl_plan* p = new (std::nothrow) sloam_cyl_plan();
  if (!p) return SLOAM_ERR_ALLOC;
  p->ctx = ctx;
  p->cfg = *cfg;

  // hypotheses per workgroup: a multiple of kWaves * kHB
  int chunk = cfg->hyp_chunk > 0 ? cfg->hyp_chunk : 256;
  chunk = std::min(chunk, round_up(cfg->n_hyp, kWaves * kHB));
  p->hcap = round_up(chunk, kWaves * kHB);
  p->tile = cfg->tile_points > 0 ? round_up(cfg->tile_points, 64) : 2048;
  // keep the workgroup's LDS within 64 KiB: shrink the point tile if needed
  const int hyp_bytes = p->hcap * (int)(sizeof(float4) + sizeof(int));
  if (hyp_bytes > kMaxLds - 64 * 12) {
    delete p;
    return SLOAM_ERR_TOO_LARGE;
  }
  p->tile = std::min(p->tile, ((kMaxLds - hyp_bytes) / 12) / 64 * 64);
  int n_xcd = 1;
  if (hipDeviceGetAttribute(&n_xcd, hipDeviceAttributeNumberOfXccs, ctx->device) != hipSuccess ||
      n_xcd < 1)
    n_xcd = 1;
  (void)hipGetLastError();
  p->n_xcd = n_xcd;
  int n_cu = 256;
  if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device) !=
          hipSuccess ||
      n_cu < 1)
    n_cu = 256;
  (void)hipGetLastError();
  p->n_simd = 4 * n_cu;
  const int st = plan_layout(p, h_offsets, n_clusters);
  if (st) {
    sloam_cyl_plan_destroy(p);
    return st;
  }
  *out = p;
  return SLOAM_OK;
}

int sloam_cyl_plan_reset(sloam_cyl_plan* p, const int32_t* h_offsets, int32_t n_clusters) {
  if (!p || !h_offsets || n_clusters < 0) return SLOAM_ERR_INVALID_ARG;
  sloam_ctx* ctx = p->ctx;
  if (ctx->capturing) return SLOAM_ERR_INVALID_ARG;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  // earlier runs of the plan read the buffers about to be rewritten
  SLOAM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const int st = plan_layout(p, h_offsets, n_clusters);
  if (st) {
    // no half-made layout: the plan holds an empty batch
    p->n_clusters = 0;
    p->n_points = 0;
    p->n_work = 0;
    for (int t = 0; t < kTiers; ++t) p->tier_n[t] = 0;
  }
  return st;
}

int sloam_cyl_fallback(sloam_ctx* ctx, const int32_t* d_offsets, int32_t n_clusters,
                       const double* d_model, float min_inlier_ratio, int32_t* d_active,
                       int32_t* d_n_active) {
  if (!ctx || n_clusters < 0 || (n_clusters > 0 && (!d_offsets || !d_model || !d_active)))
    return SLOAM_ERR_INVALID_ARG;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  if (d_n_active) {
    int st = sloam_fill_u32(ctx, (uint32_t*)d_n_active, 1, 0u);
    if (st) return st;
  }
  if (n_clusters == 0) return SLOAM_OK;
  {
    SloamKernelScope ks(ctx, "cyl_fallback");
    hipLaunchKernelGGL(cyl_fallback_kernel, dim3((n_clusters + 255) / 256), dim3(256), 0,
                       ctx->stream, d_offsets, n_clusters, d_model, min_inlier_ratio, d_active,
                       d_n_active);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}

int sloam_cyl_learned_exec(sloam_cyl_plan* verify, sloam_cyl_plan* plain, const float* d_xyz,
                           const int32_t* d_offsets, const float* d_proposal,
                           float min_inlier_ratio, int32_t* d_counts_verify,
                           int32_t* d_counts_plain, int32_t* d_fallback, int32_t* d_n_fallback,
                           int32_t* d_best, double* d_model) {
  if (!verify || !plain || !d_offsets || !d_model || !d_fallback) return SLOAM_ERR_INVALID_ARG;
  if (verify->ctx != plain->ctx || verify->n_clusters != plain->n_clusters ||
      verify->n_points != plain->n_points)
    return SLOAM_ERR_INVALID_ARG;
  const sloam_cyl_fit_cfg& vc = verify->cfg;
  const sloam_cyl_fit_cfg& pc = plain->cfg;
  if (vc.thr != pc.thr || vc.max_radius != pc.max_radius || vc.seed != pc.seed)
    return SLOAM_ERR_INVALID_ARG;
  if (verify->n_points > 0 && !d_xyz) return SLOAM_ERR_INVALID_ARG;
  sloam_ctx* ctx = verify->ctx;
  if (verify->n_clusters == 0) return SLOAM_OK;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  const uint32_t k0 = (uint32_t)vc.seed;
  const uint32_t k1 = (uint32_t)(vc.seed >> 32);
  int32_t* cv = d_counts_verify ? d_counts_verify : verify->d_counts;
  int32_t* cp = d_counts_plain ? d_counts_plain : plain->d_counts;
  const int S = verify->n_clusters;
  {
    // one axis seed per cluster serves both hypothesis sets
    SloamKernelScope ks(ctx, "cyl_seed");
    hipLaunchKernelGGL(cyl_seed_kernel, dim3(S), dim3(kBlock), 0, ctx->stream, d_xyz, d_offsets,
                       S, nullptr, verify->d_frames);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  if (verify->n_work > 0) {
    const size_t lds =
        (size_t)verify->hcap * (sizeof(float4) + sizeof(int)) + 3 * sizeof(float) * verify->tile;
    SloamKernelScope ks(ctx, "cyl_score");
    hipLaunchKernelGGL(cyl_score_kernel, dim3(verify->n_work), dim3(kBlock), lds, ctx->stream,
                       d_xyz, d_offsets, verify->d_frames, verify->d_work, vc.n_hyp, verify->hcap,
                       verify->tile, vc.thr, vc.max_radius, k0, k1, (uint32_t)vc.stream_tag,
                       d_proposal, nullptr, cv, nullptr);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  {
    SloamKernelScope ks(ctx, "cyl_decide");
    if (d_n_fallback) {
      const int st = sloam_fill_u32(ctx, (uint32_t*)d_n_fallback, 1, 0u);
      if (st) return st;
    }
    hipLaunchKernelGGL(cyl_decide_kernel, dim3((S + 255) / 256), dim3(256), 0, ctx->stream,
                       d_offsets, S, cv, vc.n_hyp, min_inlier_ratio, d_fallback, d_n_fallback);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  if (plain->n_work > 0) {
    // the plain hypotheses of the flagged clusters, in the same seed frames
    const size_t lds =
        (size_t)plain->hcap * (sizeof(float4) + sizeof(int)) + 3 * sizeof(float) * plain->tile;
    SloamKernelScope ks(ctx, "cyl_score");
    hipLaunchKernelGGL(cyl_score_kernel, dim3(plain->n_work), dim3(kBlock), lds, ctx->stream,
                       d_xyz, d_offsets, verify->d_frames, plain->d_work, pc.n_hyp, plain->hcap,
                       plain->tile, pc.thr, pc.max_radius, k0, k1, (uint32_t)pc.stream_tag,
                       nullptr, d_fallback, cp, nullptr);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  {
    SloamKernelScope ks(ctx, "cyl_refit");
    const int st = launch_refit(verify, d_xyz, d_offsets, cv, d_proposal, nullptr, d_fallback, cp,
                                pc.n_hyp, (uint32_t)pc.stream_tag, pc.refit_rounds,
                                pc.refit_iters, nullptr, d_best, d_model);
    if (st) return st;
  }
  return SLOAM_OK;
}

int sloam_cyl_autotune(sloam_ctx* ctx, const int32_t* h_offsets, int32_t n_clusters,
                       const sloam_cyl_fit_cfg* cfg, const float* d_xyz, const int32_t* d_offsets,
                       sloam_cyl_fit_cfg* out) {
  if (!ctx || !h_offsets || !cfg || !out || n_clusters < 0) return SLOAM_ERR_INVALID_ARG;
  *out = *cfg;
  if (n_clusters == 0) return SLOAM_OK;
  if (!d_xyz || !d_offsets) return SLOAM_ERR_INVALID_ARG;
  // choices are cached per batch shape: every shape gives the same results
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, int>, std::pair<int, int>> cache;
  const auto key = std::make_tuple(ctx->device, n_clusters, (int)h_offsets[n_clusters], cfg->n_hyp);
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      out->hyp_chunk = it->second.first;
      out->tile_points = it->second.second;
      return SLOAM_OK;
    }
  }
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  double* d_model = nullptr;
  int st = sloam_dev_alloc(ctx, (void**)&d_model,
                           sizeof(double) * SLOAM_CYL_MODEL_STRIDE * (size_t)n_clusters);
  if (st) return st;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  float best_ms = 1e30f;
  int best_chunk = 0, best_tile = 0;
  const int chunks[] = {64, 128, 256, 512};
  const int tiles[] = {1024, 2048, 4096};
  for (int ch : chunks) {
    if (ch > round_up(cfg->n_hyp, kWaves * kHB) && ch != 64) continue;
    for (int tl : tiles) {
      sloam_cyl_fit_cfg c = *cfg;
      c.hyp_chunk = ch;
      c.tile_points = tl;
      sloam_cyl_plan* p = nullptr;
      if (sloam_cyl_plan_create(ctx, h_offsets, n_clusters, &c, &p) != SLOAM_OK) continue;
      float ms = 0.f;
      bool ok = sloam_cyl_plan_exec(p, d_xyz, d_offsets, nullptr, nullptr, nullptr, nullptr,
                                    d_model) == SLOAM_OK;
      ok = ok && hipEventRecord(e0, ctx->stream) == hipSuccess;
      for (int r = 0; r < 3 && ok; ++r)
        ok = sloam_cyl_plan_exec(p, d_xyz, d_offsets, nullptr, nullptr, nullptr, nullptr,
                                 d_model) == SLOAM_OK;
      ok = ok && hipEventRecord(e1, ctx->stream) == hipSuccess &&
           hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
      sloam_cyl_plan_destroy(p);
      if (ok && ms < best_ms) {
        best_ms = ms;
        best_chunk = ch;
        best_tile = tl;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  sloam_dev_free(d_model);
  if (best_chunk == 0) return SLOAM_ERR_HIP;
  {
    std::lock_guard<std::mutex> lk(mu);
    cache[key] = std::make_pair(best_chunk, best_tile);
  }
  out->hyp_chunk = best_chunk;
  out->tile_points = best_tile;
  return SLOAM_OK;
}

void sloam_cyl_plan_destroy(sloam_cyl_plan* p) {
  if (!p) return;
  // pending uploads read the staging buffer
  if (p->h_stage) {
    (void)hipStreamSynchronize(p->ctx->stream);
    (void)hipHostFree(p->h_stage);
  }
  sloam_dev_free(p->d_tables);
  sloam_dev_free(p->d_frames);
  sloam_dev_free(p->d_counts);
  delete p;
}

int sloam_cyl_plan_exec(sloam_cyl_plan* p, const float* d_xyz, const int32_t* d_offsets,
                        int32_t* d_counts, int32_t* d_samples, uint32_t* d_mask, int32_t* d_best,
                        double* d_model) {
  return sloam_cyl_plan_exec_ex(p, d_xyz, d_offsets, nullptr, nullptr, d_counts, d_samples, d_mask,
                                d_best, d_model);
}

int sloam_cyl_plan_exec_ex(sloam_cyl_plan* p, const float* d_xyz, const int32_t* d_offsets,
                           const float* d_proposal, const int32_t* d_active, int32_t* d_counts,
                           int32_t* d_samples, uint32_t* d_mask, int32_t* d_best,
                           double* d_model) {
  if (!p || !d_offsets || !d_model || (p->n_points > 0 && !d_xyz)) return SLOAM_ERR_INVALID_ARG;
  if (d_active && d_mask) return SLOAM_ERR_INVALID_ARG;  // the mask is cleared for all clusters
  sloam_ctx* ctx = p->ctx;
  if (p->n_clusters == 0) return SLOAM_OK;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  const sloam_cyl_fit_cfg& cfg = p->cfg;
  const uint32_t k0 = (uint32_t)cfg.seed;
  const uint32_t k1 = (uint32_t)(cfg.seed >> 32);
  const uint32_t tag = (uint32_t)cfg.stream_tag;
  int32_t* counts = d_counts ? d_counts : p->d_counts;
  if (d_mask) {
    int st = sloam_fill_u32(ctx, d_mask, (size_t)((p->n_points + 31) / 32), 0u);
    if (st) return st;
  }
  {
    SloamKernelScope ks(ctx, "cyl_seed");
    hipLaunchKernelGGL(cyl_seed_kernel, dim3(p->n_clusters), dim3(kBlock), 0, ctx->stream, d_xyz,
                       d_offsets, p->n_clusters, d_active, p->d_frames);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  if (p->n_work > 0) {
    const size_t lds = (size_t)p->hcap * (sizeof(float4) + sizeof(int)) + 3 * sizeof(float) * p->tile;
    SloamKernelScope ks(ctx, "cyl_score");
    hipLaunchKernelGGL(cyl_score_kernel, dim3(p->n_work), dim3(kBlock), lds, ctx->stream, d_xyz,
                       d_offsets, p->d_frames, p->d_work, cfg.n_hyp, p->hcap, p->tile, cfg.thr,
                       cfg.max_radius, k0, k1, tag, d_proposal, d_active, counts, d_samples);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  {
    SloamKernelScope ks(ctx, "cyl_refit");
    const int st = launch_refit(p, d_xyz, d_offsets, counts, d_proposal, d_active, nullptr,
                                nullptr, 0, 0u, cfg.refit_rounds, cfg.refit_iters, d_mask, d_best,
                                d_model);
    if (st) return st;
  }
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}

}  // extern "C"
