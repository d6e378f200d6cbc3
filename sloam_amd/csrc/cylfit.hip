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

// order (nullable): the cluster of each workgroup, -1 for padding (the plans'
// largest-first order dealt to the XCDs as the scoring work list deals them,
// so a cluster's points are read through the L2 its scoring will read them
// through); without it workgroup b seeds cluster b
__global__ __launch_bounds__(kBlock) void cyl_seed_kernel(const float* __restrict__ xyz,
                                                          const int32_t* __restrict__ offsets,
                                                          int n_clusters,
                                                          const int32_t* __restrict__ active,
                                                          const int32_t* __restrict__ bcnt,
                                                          CylFrame* __restrict__ frames,
                                                          const int32_t* __restrict__ order) {
  __shared__ float red_f[kWaves][6];
  __shared__ long long red_i[kWaves][9];
  const int c = order ? order[blockIdx.x] : (int)blockIdx.x;
  if (c < 0 || c >= n_clusters || (bcnt && c >= bcnt[0])) return;
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
    const int32_t* __restrict__ bcnt, int32_t* __restrict__ counts,
    int32_t* __restrict__ samples) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  if (bcnt && (int)blockIdx.x >= bcnt[1]) return;  // past the batch's work items
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

// ---------------------------------------------------------------------------
// refit: best hypothesis, mask, LO-RANSAC cylinder refit with FP32 residuals
// and Jacobians, compensated FP32 sums and an FP64 5x5 solve.
//
// Rounds of Gauss-Newton steps on the 5-parameter geometric cylinder residual
//   d_i = ||(p_i - c) - ((p_i - c).a) a|| - r
// over the round's inliers: the FP32 inliers of the best hypothesis in round
// 0, |d_i| < thr under the model at the start of the round afterwards.
// Output record: root[3], ray[3], radius, inliers.
//
// Every cluster is cut into chunks of kRefitChunk points; the sums of a step
// are made per chunk by kRefitBlock threads: per point, in FP32, w, h,
// rho = ||w||, d = rho - r, n = w / rho and J = (-h n.e1, -h n.e2, -n.e1,
// -n.e2, -1); each thread adds its points' J J^T and J d terms with FMAs, the
// chunk's threads' values are added with TwoSum in groups of kRefitGroup
// threads (thread order, then group order) into one (sum, error) pair per
// term; the solve adds the chunks' pairs in chunk order and solves the
// normal equations in FP64.  The order of every addition is fixed (no
// atomics on values), so results do not depend on scheduling; the oracle
// restates the same arithmetic.  Two launch forms give the same bits:
//   per step  (large batches)  cyl_refit_init (best hypothesis, inlier mask,
//             round-0 membership bits, initial model), then per Gauss-Newton
//             step cyl_refit_step (one workgroup per chunk) and
//             cyl_refit_solve (one wave per cluster: lane q sums term q over
//             the chunks, lane 0 solves), then cyl_refit_final (the final
//             inliers' centroid sums per chunk) and cyl_refit_finish (the
//             output record, one wave per cluster);
//   fused     (small batches, e.g. a keyframe's trees)  cyl_refit_init, then
//             cyl_refit_fused: one workgroup walks every step of one cluster
//             over its chunks, its first wave solving between steps.
// The kernels of the per-step form read the previous launch's results (no
// fences between workgroups inside a launch).
// ---------------------------------------------------------------------------
constexpr int kRefitBlock = 256;
constexpr int kRefitPerThread = 8;
constexpr int kRefitChunk = kRefitBlock * kRefitPerThread;
constexpr int kRefitGroup = 32;
constexpr int kRefitGroups = kRefitBlock / kRefitGroup;
constexpr int kRefitTerms = 20;  // 15 of J J^T (upper triangle), 5 of J d
// largest cluster the plans take (points), and its chunks
constexpr int kRefitMaxPoints = 1 << 17;
constexpr int kRefitFusedChunks = kRefitMaxPoints / kRefitChunk;
// the fused refit (a workgroup walks all steps of one cluster) takes a
// host-planned batch of at most this many clusters of at most this many
// chunks: beyond, one launch per step (every chunk of every cluster in
// parallel) is faster; counted plans (a keyframe's trees) always take it
constexpr int kRefitFusedPlanClusters = 512;
constexpr int kRefitFusedPlanChunks = 4;
// clusters (one wave each) per workgroup of the per-cluster solve and finish
constexpr int kRefitWaveClusters = 4;

enum RefitStatus : int32_t {
  kRefitActive = 0,    // Gauss-Newton steps continue
  kRefitStopped = 1,   // no more steps (too few inliers, radius <= 0)
  kRefitInvalid = 2,   // record already written as invalid
  kRefitInactive = 3,  // cluster not part of this run
};

// Per-cluster refit state.  The FP32 copy of the model and the basis of the
// plane orthogonal to its axis are made whenever the FP64 model changes.
struct RefitState {
  double a[3], c[3], r;  // axis, axis point (relative to the frame centre), radius
  float af[3], cf[3], rf, e1f[3], e2f[3];
  int32_t status;
  int32_t skip_round;  // round whose remaining steps are skipped (failed solve), -1
  int32_t rounds, iters;
};

struct Comp2 {
  float s, e;  // value s + e
};

// s + e = a + b exactly (TwoSum)
__device__ inline void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  const float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

__device__ inline void comp_add(Comp2& acc, float x) {
  float s, e;
  two_sum(acc.s, x, s, e);
  acc.s = s;
  acc.e = acc.e + e;
}

__device__ inline void comp_add(Comp2& acc, const Comp2& x) {
  float s, e;
  two_sum(acc.s, x.s, s, e);
  acc.s = s;
  acc.e = acc.e + (e + x.e);
}

template <int K>
struct RefitRed {
  float v[K][kRefitGroups][kRefitGroup + 1];  // +1: group chains read distinct banks
  Comp2 grp[K][kRefitGroups];
};

// The chunk's (sum, error) pair of each of K per-thread values, into out[K]
// (global): TwoSum chains over kRefitGroup threads, then over the groups.
template <int K>
__device__ inline void chunk_partial(const float (&g)[K], RefitRed<K>& r, Comp2* out) {
  const int t = threadIdx.x;
  const int tg = t / kRefitGroup, tl = t % kRefitGroup;
#pragma unroll
  for (int k = 0; k < K; ++k) r.v[k][tg][tl] = g[k];
  __syncthreads();
  if (t < K * kRefitGroups) {
    const int k = t / kRefitGroups, q = t % kRefitGroups;
    Comp2 s = {0.0f, 0.0f};
    for (int j = 0; j < kRefitGroup; ++j) comp_add(s, r.v[k][q][j]);
    r.grp[k][q] = s;
  }
  __syncthreads();
  if (t < K) {
    Comp2 s = {0.0f, 0.0f};
#pragma unroll
    for (int q = 0; q < kRefitGroups; ++q) comp_add(s, r.grp[t][q]);
    out[t] = s;
  }
}

// Sum of one int per thread over the workgroup (every thread gets it).
__device__ inline int block_sum_int(int v, int* lds) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  int s = 0;
#pragma unroll
  for (int w = 0; w < kRefitBlock / 64; ++w) s += lds[w];
  return s;
}

// FP32 copy of the FP64 model and its basis.
__device__ inline void refit_model_f32(RefitState& S) {
  double e1[3], e2[3];
  sl_basis(S.a, e1, e2);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    S.af[i] = (float)S.a[i];
    S.cf[i] = (float)S.c[i];
    S.e1f[i] = (float)e1[i];
    S.e2f[i] = (float)e2[i];
  }
  S.rf = (float)S.r;
}

// FP32 distance of x (relative to the frame centre) to the model; w, h, rho out.
__device__ inline float refit_dist(const RefitState& S, const float x[3], float w[3], float& h,
                                   float& rho) {
  const float q0 = x[0] - S.cf[0], q1 = x[1] - S.cf[1], q2 = x[2] - S.cf[2];
  h = __builtin_fmaf(q2, S.af[2], __builtin_fmaf(q1, S.af[1], q0 * S.af[0]));
  w[0] = __builtin_fmaf(-h, S.af[0], q0);
  w[1] = __builtin_fmaf(-h, S.af[1], q1);
  w[2] = __builtin_fmaf(-h, S.af[2], q2);
  rho = sqrtf(__builtin_fmaf(w[2], w[2], __builtin_fmaf(w[1], w[1], w[0] * w[0])));
  return rho - S.rf;
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

// Refit chunk ck = {cluster, first point (cluster-relative), points, chunk
// index within the cluster}; cluster record cl = {first chunk id, chunks,
// first membership word, 0}.  Membership words: two per 64 points of the
// cluster, written by the wave that owns those 64 points.

// Round 0: best hypothesis, global mask bits, membership bits, initial model.
// A flagged cluster (alt_flags) takes its hypotheses from the alternate set
// (its own counts, hypothesis count and sample stream, no proposal) and
// refits with the alternate round and step counts.
__global__ __launch_bounds__(kRefitBlock) void cyl_refit_init_kernel(
    const float* __restrict__ xyz, const int32_t* __restrict__ offsets,
    const CylFrame* __restrict__ frames, const int32_t* __restrict__ counts,
    const int4* __restrict__ rchunks, const int4* __restrict__ rclus, int n_hyp, float thr,
    float max_radius, uint32_t k0, uint32_t k1, uint32_t tag, int refit_rounds, int refit_iters,
    const float* __restrict__ prop, const int32_t* __restrict__ active,
    const int32_t* __restrict__ alt_flags, const int32_t* __restrict__ alt_counts, int alt_n_hyp,
    uint32_t alt_tag, int alt_rounds, int alt_iters, const int32_t* __restrict__ bcnt,
    RefitState* __restrict__ state, uint32_t* __restrict__ member, uint32_t* __restrict__ mask,
    int32_t* __restrict__ best_out, double* __restrict__ model) {
  constexpr int W = kRefitBlock / 64;
  __shared__ long long redk[W];
  __shared__ HypOut best_h;
  if (bcnt && (int)blockIdx.x >= bcnt[2]) return;  // past the batch's chunks
  const int4 ck = rchunks[blockIdx.x];
  const int c = ck.x;
  if (c < 0) return;  // padding of the per-XCD interleave
  if (active && !active[c]) {
    if (ck.w == 0 && threadIdx.x == 0) state[c].status = kRefitInactive;
    return;
  }
  const int off = offsets[c];
  const int n = offsets[c + 1] - off;
  const float* P = xyz + 3 * (size_t)off;
  const CylFrame f = frames[c];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const bool alt = alt_flags && alt_flags[c];
  const int32_t* cnts = alt ? alt_counts : counts;
  const int nh = alt ? alt_n_hyp : n_hyp;
  const uint32_t tg = alt ? alt_tag : tag;
  const float* pc = alt ? nullptr : cluster_prop(prop, c);

  // best hypothesis: key = count * 2^32 + (2^32 - 1 - index), max wins
  long long key = -1;
  for (int h = threadIdx.x; h < nh; h += kRefitBlock) {
    const long long kk = ((long long)cnts[(size_t)c * nh + h] << 32) |
                         (long long)(0xFFFFFFFFu - (uint32_t)h);
    key = kk > key ? kk : key;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const long long other = __shfl_xor(key, o, 64);
    key = other > key ? other : key;
  }
  if (lane == 0) redk[wave] = key;
  __syncthreads();
  key = redk[0];
#pragma unroll
  for (int w = 1; w < W; ++w) key = redk[w] > key ? redk[w] : key;
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
  }
  __syncthreads();
  const HypOut h = best_h;

  // membership bits of the chunk (cluster-relative, 64-aligned per wave)
  const int4 cl = rclus[c];
  uint32_t* mem = member + cl.z;
  const int end = ck.y + ck.z;
#pragma unroll
  for (int j = 0; j < kRefitPerThread; ++j) {
    const int gbase = ck.y + kRefitBlock * j + 64 * wave;  // first point of this wave's group
    const int i = gbase + lane;
    bool in = false;
    if (i < end) {
      const float* q = P + 3 * (size_t)i;
      float u, v, s;
      local_uvs(f, q[0], q[1], q[2], &u, &v, &s);
      in = is_inlier(u, v, s, h);
    }
    const uint64_t b = __builtin_amdgcn_ballot_w64(in);
    if (lane == 0 && gbase < end) {
      mem[gbase >> 5] = (uint32_t)b;
      mem[(gbase >> 5) + 1] = (uint32_t)(b >> 32);
    }
  }
  if (mask) {
    // global mask: 64-aligned global indices, so that each wave owns two whole
    // 32-bit words (or-ed atomically: neighbouring chunks share boundary words)
    const int gs = off + ck.y, ge = off + end;
    for (int gb = (gs & ~63); gb < ge; gb += kRefitBlock) {
      const int g = gb + (int)threadIdx.x;
      bool in = false;
      if (g >= gs && g < ge) {
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
  if (ck.w != 0 || threadIdx.x != 0) return;
  // the cluster's first chunk: initial model (the inlier count of the best
  // hypothesis is its score)
  if (best_out) best_out[c] = best;
  RefitState S;
  S.status = kRefitActive;
  S.skip_round = -1;
  S.rounds = alt ? alt_rounds : refit_rounds;
  S.iters = alt ? alt_iters : refit_iters;
  const int kin = best >= 0 ? cnts[(size_t)c * nh + best] : 0;
  if (kin < 3 || !h.valid) {
    S.status = kRefitInvalid;
    write_invalid(model + (size_t)c * SLOAM_CYL_MODEL_STRIDE, (double)kin);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    S.a[i] = f.axis[i];
    S.c[i] = (double)h.cu * (double)f.e1[i] + (double)h.cv * (double)f.e2[i];
  }
  S.r = (double)h.r;
  refit_model_f32(S);
  state[c] = S;
}

// One Gauss-Newton step's partials of the chunk [first, end) of a cluster
// (points P, frame f, model S, membership bits mem), by all threads of the
// workgroup: the chunk's (sum, error) pairs into out[kRefitTerms] and its
// inlier count into *out_count.  In a re-selecting step the membership bits
// of the chunk are rewritten from the model.
__device__ inline void refit_step_chunk(const float* __restrict__ P, const CylFrame& f,
                                        const RefitState& S, uint32_t* __restrict__ mem,
                                        int first, int end, bool reselect, float thr,
                                        RefitRed<kRefitTerms>& red, int* cnt_w, Comp2* out,
                                        int32_t* out_count) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  float g[kRefitTerms];
#pragma unroll
  for (int q = 0; q < kRefitTerms; ++q) g[q] = 0.0f;
  int cnt = 0;
  // every point of the thread loaded before any is used (the loads overlap)
  float px[kRefitPerThread][3];
#pragma unroll
  for (int j = 0; j < kRefitPerThread; ++j) {
    const int i = first + kRefitBlock * j + (int)threadIdx.x;
    if (i < end) {
      const float* q = P + 3 * (size_t)i;
      px[j][0] = q[0];
      px[j][1] = q[1];
      px[j][2] = q[2];
    }
  }
#pragma unroll 2
  for (int j = 0; j < kRefitPerThread; ++j) {
    const int gbase = first + kRefitBlock * j + 64 * wave;
    const int i = gbase + lane;
    bool in = false;
    float d = 0.f, hh = 0.f, rho = 0.f, w[3] = {0.f, 0.f, 0.f};
    if (i < end) {
      const float x[3] = {px[j][0] - f.m[0], px[j][1] - f.m[1], px[j][2] - f.m[2]};
      d = refit_dist(S, x, w, hh, rho);
      in = reselect ? (d < thr && d > -thr) : ((mem[i >> 5] >> (i & 31)) & 1u) != 0u;
    }
    if (reselect) {
      const uint64_t b = __builtin_amdgcn_ballot_w64(in);
      if (lane == 0 && gbase < end) {
        mem[gbase >> 5] = (uint32_t)b;
        mem[(gbase >> 5) + 1] = (uint32_t)(b >> 32);
      }
    }
    if (in) {
      ++cnt;
      if (rho > 0.0f) {
        const float inv = 1.0f / rho;
        const float n1 =
            __builtin_fmaf(w[2], S.e1f[2], __builtin_fmaf(w[1], S.e1f[1], w[0] * S.e1f[0])) * inv;
        const float n2 =
            __builtin_fmaf(w[2], S.e2f[2], __builtin_fmaf(w[1], S.e2f[1], w[0] * S.e2f[0])) * inv;
        const float J[5] = {-hh * n1, -hh * n2, -n1, -n2, -1.0f};
        int t = 0;
#pragma unroll
        for (int a = 0; a < 5; ++a)
#pragma unroll
          for (int b = a; b < 5; ++b) {
            g[t] = __builtin_fmaf(J[a], J[b], g[t]);
            ++t;
          }
#pragma unroll
        for (int a = 0; a < 5; ++a) g[15 + a] = __builtin_fmaf(J[a], d, g[15 + a]);
      }
    }
  }
  chunk_partial<kRefitTerms>(g, red, out);
  const int bc = block_sum_int(cnt, cnt_w);
  if (threadIdx.x == 0) *out_count = bc;
}

// Term q of a step's sums over a cluster's nck chunks: the chunks' (sum,
// error) pairs of term q added in chunk order, then sum + error in FP64.
__device__ inline double refit_term(const Comp2* __restrict__ partial, int nck, int q) {
  Comp2 t = {0.0f, 0.0f};
  for (int k = 0; k < nck; ++k) comp_add(t, partial[(size_t)k * kRefitTerms + q]);
  return (double)t.s + (double)t.e;
}

// The Gauss-Newton step's solve of one cluster (one thread) from its summed
// terms gs and inlier count: the FP64 5x5 normal equations, the next model in
// S.  Too few inliers stop the cluster; a failed factorisation skips the
// rest of the round.
__device__ inline void refit_solve_terms(RefitState& S, const double (&gs)[kRefitTerms],
                                         int tot_cnt, int round) {
  if (tot_cnt < 5) {
    S.status = kRefitStopped;
    return;
  }
  double A[5][5], rhs[5], dl[5];
  {
    int t = 0;
#pragma unroll
    for (int a = 0; a < 5; ++a)
#pragma unroll
      for (int bb = a; bb < 5; ++bb) {
        A[a][bb] = gs[t];
        A[bb][a] = gs[t];
        ++t;
      }
#pragma unroll
    for (int a = 0; a < 5; ++a) rhs[a] = -gs[15 + a];
  }
  if (!cholesky_solve<5>(A, rhs, dl)) {
    S.skip_round = round;
    return;
  }
  double e1[3], e2[3];
  sl_basis(S.a, e1, e2);
  double na[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) na[i] = S.a[i] + dl[0] * e1[i] + dl[1] * e2[i];
  const double nl = sqrt(na[0] * na[0] + na[1] * na[1] + na[2] * na[2]);
  RefitState N = S;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    N.a[i] = na[i] / nl;
    N.c[i] = S.c[i] + dl[2] * e1[i] + dl[3] * e2[i];
  }
  N.r = S.r + dl[4];
  if (!(N.r > 0.0)) N.status = kRefitStopped;
  refit_model_f32(N);
  S = N;
}

// Sum of one int per lane over a wave (every lane gets it).
__device__ inline int wave_sum_int(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// The solve by one wave: lane q < kRefitTerms sums term q over the chunks
// (the loads of all terms in flight at once), lane 0 solves; S is lane 0's.
__device__ inline void refit_solve_wave(RefitState& S, const Comp2* __restrict__ partial,
                                        const int32_t* __restrict__ pcount, int nck, int round) {
  const int lane = threadIdx.x & 63;
  const double mine = lane < kRefitTerms ? refit_term(partial, nck, lane) : 0.0;
  int cnt = 0;
  for (int k = lane; k < nck; k += 64) cnt += pcount[k];
  cnt = wave_sum_int(cnt);
  double gs[kRefitTerms];
#pragma unroll
  for (int q = 0; q < kRefitTerms; ++q) gs[q] = __shfl(mine, q, 64);
  if (lane == 0) refit_solve_terms(S, gs, cnt, round);
}

// The final pass over the chunk [first, end): the chunk sums of the final
// inliers' coordinates (frame-relative) into out[3] and their count.
__device__ inline void refit_final_chunk(const float* __restrict__ P, const CylFrame& f,
                                         const RefitState& S, int first, int end, float thr,
                                         RefitRed<3>& red, int* cnt_w, Comp2* out,
                                         int32_t* out_count) {
  const bool pos = S.r > 0.0;
  float g[3] = {0.0f, 0.0f, 0.0f};
  int cnt = 0;
  if (pos) {
    float px[kRefitPerThread][3];
#pragma unroll
    for (int j = 0; j < kRefitPerThread; ++j) {
      const int i = first + kRefitBlock * j + (int)threadIdx.x;
      if (i < end) {
        const float* q = P + 3 * (size_t)i;
        px[j][0] = q[0];
        px[j][1] = q[1];
        px[j][2] = q[2];
      }
    }
#pragma unroll 2
    for (int j = 0; j < kRefitPerThread; ++j) {
      const int i = first + kRefitBlock * j + (int)threadIdx.x;
      if (i < end) {
        const float x[3] = {px[j][0] - f.m[0], px[j][1] - f.m[1], px[j][2] - f.m[2]};
        float w[3], hh, rho;
        const float d = refit_dist(S, x, w, hh, rho);
        if (d < thr && d > -thr) {
          g[0] = g[0] + x[0];
          g[1] = g[1] + x[1];
          g[2] = g[2] + x[2];
          ++cnt;
        }
      }
    }
  }
  chunk_partial<3>(g, red, out);
  const int bc = block_sum_int(cnt, cnt_w);
  if (threadIdx.x == 0) *out_count = bc;
}

// The output record of one cluster by one wave from its final pass's chunk
// partials: lanes 0..2 sum the coordinates over the chunks, lane 0 writes
// root (the final inliers' centroid projected on the axis), ray, radius,
// inliers; invalid without a positive radius or three inliers.
__device__ inline void refit_finish_wave(const RefitState& S, const CylFrame& f,
                                         const Comp2* __restrict__ partial,
                                         const int32_t* __restrict__ pcount, int nck,
                                         double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const double mine = lane < 3 ? refit_term(partial, nck, lane) : 0.0;
  int tot_cnt = 0;
  for (int k = lane; k < nck; k += 64) tot_cnt += pcount[k];
  tot_cnt = wave_sum_int(tot_cnt);
  const double sum[3] = {__shfl(mine, 0, 64), __shfl(mine, 1, 64), __shfl(mine, 2, 64)};
  if (lane != 0) return;
  const double k = (double)tot_cnt;
  if (!(S.r > 0.0) || tot_cnt < 3) {
    write_invalid(out, k);
    return;
  }
  double mean[3];
  for (int i = 0; i < 3; ++i) mean[i] = sum[i] / k;
  const double ht =
      (mean[0] - S.c[0]) * S.a[0] + (mean[1] - S.c[1]) * S.a[1] + (mean[2] - S.c[2]) * S.a[2];
  double ray[3] = {S.a[0], S.a[1], S.a[2]};
  sl_sign_normalise(ray);
  for (int i = 0; i < 3; ++i) {
    out[i] = (double)f.m[i] + S.c[i] + ht * S.a[i];
    out[3 + i] = ray[i];
  }
  out[6] = S.r;
  out[7] = k;
}

// One Gauss-Newton step (launch `step` of the refit: round step / iters,
// iteration step % iters of each cluster's own schedule), one workgroup per
// chunk.
__global__ __launch_bounds__(kRefitBlock) void cyl_refit_step_kernel(
    const float* __restrict__ xyz, const int32_t* __restrict__ offsets,
    const CylFrame* __restrict__ frames, const int4* __restrict__ rchunks,
    const int4* __restrict__ rclus, const RefitState* __restrict__ state,
    uint32_t* __restrict__ member, Comp2* __restrict__ partial, int32_t* __restrict__ pcount,
    int step, float thr, const int32_t* __restrict__ bcnt) {
  __shared__ RefitRed<kRefitTerms> red;
  __shared__ int cnt_w[kRefitBlock / 64];
  if (bcnt && (int)blockIdx.x >= bcnt[2]) return;
  const int4 ck = rchunks[blockIdx.x];
  const int c = ck.x;
  if (c < 0) return;
  const RefitState S = state[c];
  if (S.status != kRefitActive) return;
  if (step >= S.rounds * S.iters) return;
  const int round = step / S.iters, it = step % S.iters;
  if (S.skip_round == round) return;
  const int4 cl = rclus[c];
  const int chunk_id = cl.x + ck.w;
  refit_step_chunk(xyz + 3 * (size_t)offsets[c], frames[c], S, member + cl.z, ck.y, ck.y + ck.z,
                   round > 0 && it == 0, thr, red, cnt_w,
                   partial + (size_t)chunk_id * kRefitTerms, pcount + chunk_id);
}

// The Gauss-Newton step's solve after the step's chunk partials, one wave
// per cluster (kRefitWaveClusters per workgroup).
__global__ __launch_bounds__(64 * kRefitWaveClusters) void cyl_refit_solve_kernel(
    int n_clusters, const int4* __restrict__ rclus, RefitState* __restrict__ state,
    const Comp2* __restrict__ partial, const int32_t* __restrict__ pcount, int step,
    const int32_t* __restrict__ bcnt) {
  const int c = blockIdx.x * kRefitWaveClusters + (int)(threadIdx.x >> 6);
  if (c >= n_clusters || (bcnt && c >= bcnt[0])) return;
  RefitState S = state[c];
  if (S.status != kRefitActive) return;
  if (step >= S.rounds * S.iters) return;
  const int round = step / S.iters;
  if (S.skip_round == round) return;
  const int4 cl = rclus[c];
  refit_solve_wave(S, partial + (size_t)cl.x * kRefitTerms, pcount + cl.x, cl.y, round);
  if ((threadIdx.x & 63) == 0) state[c] = S;
}

// Final inliers under the final model and their centroid's chunk sums.
__global__ __launch_bounds__(kRefitBlock) void cyl_refit_final_kernel(
    const float* __restrict__ xyz, const int32_t* __restrict__ offsets,
    const CylFrame* __restrict__ frames, const int4* __restrict__ rchunks,
    const int4* __restrict__ rclus, const RefitState* __restrict__ state,
    Comp2* __restrict__ partial, int32_t* __restrict__ pcount, float thr,
    const int32_t* __restrict__ bcnt) {
  __shared__ RefitRed<3> red;
  __shared__ int cnt_w[kRefitBlock / 64];
  if (bcnt && (int)blockIdx.x >= bcnt[2]) return;
  const int4 ck = rchunks[blockIdx.x];
  const int c = ck.x;
  if (c < 0) return;
  const RefitState S = state[c];
  if (S.status == kRefitInvalid || S.status == kRefitInactive) return;
  const int4 cl = rclus[c];
  const int chunk_id = cl.x + ck.w;
  refit_final_chunk(xyz + 3 * (size_t)offsets[c], frames[c], S, ck.y, ck.y + ck.z, thr, red,
                    cnt_w, partial + (size_t)chunk_id * kRefitTerms, pcount + chunk_id);
}

// The output record of each cluster from the final pass's chunk partials,
// one wave per cluster (kRefitWaveClusters per workgroup).
__global__ __launch_bounds__(64 * kRefitWaveClusters) void cyl_refit_finish_kernel(
    int n_clusters, const CylFrame* __restrict__ frames, const int4* __restrict__ rclus,
    const RefitState* __restrict__ state, const Comp2* __restrict__ partial,
    const int32_t* __restrict__ pcount, const int32_t* __restrict__ bcnt,
    double* __restrict__ model) {
  const int c = blockIdx.x * kRefitWaveClusters + (int)(threadIdx.x >> 6);
  if (c >= n_clusters || (bcnt && c >= bcnt[0])) return;
  const RefitState S = state[c];
  if (S.status == kRefitInvalid || S.status == kRefitInactive) return;
  const int4 cl = rclus[c];
  refit_finish_wave(S, frames[c], partial + (size_t)cl.x * kRefitTerms, pcount + cl.x, cl.y,
                    model + (size_t)c * SLOAM_CYL_MODEL_STRIDE);
}

// Every Gauss-Newton step, the final pass and the output record of one
// cluster in one workgroup: the workgroup of the cluster's first refit chunk
// (the others exit) walks the cluster's chunks in order at every step, keeps
// the chunk partials and counts in LDS and solves on its first thread
// between steps.  Every sum is the same function over the same chunk as in
// the launch-per-step sequence, so the records carry the same bits, without
// its launches and per-cluster solve kernels.  Clusters of up to
// kRefitFusedChunks chunks (every plan's bound).
__global__ __launch_bounds__(kRefitBlock) void cyl_refit_fused_kernel(
    const float* __restrict__ xyz, const int32_t* __restrict__ offsets,
    const CylFrame* __restrict__ frames, const int4* __restrict__ rchunks,
    const int4* __restrict__ rclus, RefitState* __restrict__ state, uint32_t* __restrict__ member,
    float thr, const int32_t* __restrict__ bcnt, double* __restrict__ model) {
  __shared__ RefitRed<kRefitTerms> red;
  __shared__ RefitRed<3> red3;
  __shared__ int cnt_w[kRefitBlock / 64];
  __shared__ Comp2 part[kRefitFusedChunks * kRefitTerms];
  __shared__ int32_t pc[kRefitFusedChunks];
  __shared__ RefitState s_next;
  if (bcnt && (int)blockIdx.x >= bcnt[2]) return;
  const int4 ck = rchunks[blockIdx.x];
  const int c = ck.x;
  if (c < 0 || ck.w != 0) return;  // one workgroup per cluster
  RefitState S = state[c];
  if (S.status == kRefitInvalid || S.status == kRefitInactive) return;
  const int4 cl = rclus[c];
  if (cl.y > kRefitFusedChunks) return;  // excluded by the plans' cluster bound
  const int off = offsets[c];
  const int n = offsets[c + 1] - off;
  const float* P = xyz + 3 * (size_t)off;
  const CylFrame f = frames[c];
  uint32_t* mem = member + cl.z;
  const int steps = S.rounds * S.iters;
  for (int step = 0; step < steps && S.status == kRefitActive; ++step) {
    const int round = step / S.iters, it = step % S.iters;
    if (S.skip_round == round) continue;
    for (int k = 0; k < cl.y; ++k) {
      const int first = k * kRefitChunk;
      refit_step_chunk(P, f, S, mem, first, first + min(kRefitChunk, n - first),
                       round > 0 && it == 0, thr, red, cnt_w, part + k * kRefitTerms, pc + k);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
      RefitState N = S;
      refit_solve_wave(N, part, pc, cl.y, round);
      if (threadIdx.x == 0) s_next = N;
    }
    __syncthreads();
    S = s_next;
  }
  for (int k = 0; k < cl.y; ++k) {
    const int first = k * kRefitChunk;
    refit_final_chunk(P, f, S, first, first + min(kRefitChunk, n - first), thr, red3, cnt_w,
                      part + k * kRefitTerms, pc + k);
  }
  __syncthreads();
  if (threadIdx.x == 0) state[c] = S;
  if (threadIdx.x < 64)
    refit_finish_wave(S, f, part, pc, cl.y, model + (size_t)c * SLOAM_CYL_MODEL_STRIDE);
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
  // refit: chunks {cluster, first point, points, chunk index} in launch
  // order, and per cluster {first chunk id, chunks, first membership word, 0}
  int n_rchunks = 0;
  int4* d_rchunks = nullptr;
  int4* d_rclus = nullptr;
  int n_seed = 0;                    // seed workgroups (clusters + padding)
  int32_t* d_seed_order = nullptr;   // their clusters (in d_tables)
  CylFrame* d_frames = nullptr;
  int32_t* d_counts = nullptr;  // used when the caller passes no counts buffer
  int32_t* d_rescue = nullptr;  // per cluster: refitted again by the plain method
  RefitState* d_state = nullptr;
  uint32_t* d_member = nullptr;  // membership bits of the current round
  Comp2* d_partial = nullptr;    // per chunk: kRefitTerms (sum, error) pairs
  int32_t* d_pcount = nullptr;   // per chunk: inliers counted
  int n_xcd = 1;
  // counted plans: the layout is made on the device at every exec from the
  // offsets and the cluster count at d_n (at most cap clusters and
  // cap_points points); n_clusters, n_work and n_rchunks are then the grid
  // sizes, d_cnt the batch's own {clusters, work items, refit chunks,
  // overflow}
  bool counted = false;
  // refit in one workgroup per cluster (cyl_refit_fused) instead of one
  // launch per Gauss-Newton step
  bool fused = true;
  const int32_t* d_n = nullptr;
  int64_t cap_points = 0;
  int cpc = 0;  // work items per cluster
  int32_t* d_cnt = nullptr;
  // d_work, d_rchunks and d_rclus share one allocation, uploaded in one copy;
  // allocated sizes (bytes for the tables, elements otherwise) only grow
  // across sloam_cyl_plan_reset
  void* d_tables = nullptr;
  size_t cap_tables = 0, cap_frames = 0, cap_counts = 0, cap_state = 0, cap_member = 0,
         cap_partial = 0, cap_rescue = 0;
  // pinned staging of the tables (bytes): their uploads are asynchronous on
  // the context stream, ahead of the runs that read them
  void* h_stage = nullptr;
  size_t cap_stage = 0;
};

static int round_up(int x, int m) { return (x + m - 1) / m * m; }

// Launches the refit of plan p: the init pass, then either the fused refit
// (one workgroup per cluster) or one launch per Gauss-Newton step (as many as
// the longest schedule among the clusters) and the final pass.
// alt_flags (nullable) marks the clusters that take their best hypothesis
// from alt_counts (alt_n_hyp hypotheses of sample stream alt_tag, no
// proposal) and refit with alt_rounds x alt_iters steps, instead of counts.
static int launch_refit(sloam_cyl_plan* p, const float* d_xyz, const int32_t* d_offsets,
                        const int32_t* counts, const float* d_proposal, const int32_t* d_active,
                        const int32_t* alt_flags, const int32_t* alt_counts, int alt_n_hyp,
                        uint32_t alt_tag, int alt_rounds, int alt_iters, uint32_t* d_mask,
                        int32_t* d_best, double* d_model, bool force_fused = false) {
  sloam_ctx* ctx = p->ctx;
  const sloam_cyl_fit_cfg& cfg = p->cfg;
  if (p->n_rchunks == 0) return SLOAM_OK;
  const uint32_t k0 = (uint32_t)cfg.seed;
  const uint32_t k1 = (uint32_t)(cfg.seed >> 32);
  const uint32_t tag = (uint32_t)cfg.stream_tag;
  const dim3 grid(p->n_rchunks), block(kRefitBlock);
  const int32_t* cnt = p->counted ? p->d_cnt : nullptr;
  hipLaunchKernelGGL(cyl_refit_init_kernel, grid, block, 0, ctx->stream, d_xyz, d_offsets,
                     p->d_frames, counts, p->d_rchunks, p->d_rclus, cfg.n_hyp, cfg.thr,
                     cfg.max_radius, k0, k1, tag, cfg.refit_rounds, cfg.refit_iters, d_proposal,
                     d_active, alt_flags, alt_counts, alt_n_hyp, alt_tag, alt_rounds, alt_iters,
                     cnt, p->d_state, p->d_member, d_mask, d_best, d_model);
  SLOAM_HIP(ctx, hipGetLastError());
  if (p->fused || force_fused) {
    hipLaunchKernelGGL(cyl_refit_fused_kernel, grid, block, 0, ctx->stream, d_xyz, d_offsets,
                       p->d_frames, p->d_rchunks, p->d_rclus, p->d_state, p->d_member, cfg.thr,
                       cnt, d_model);
    SLOAM_HIP(ctx, hipGetLastError());
    return SLOAM_OK;
  }
  // one wave per cluster
  const dim3 cgrid((std::max(p->n_clusters, 1) + kRefitWaveClusters - 1) / kRefitWaveClusters),
      cblock(64 * kRefitWaveClusters);
  int steps = cfg.refit_rounds * cfg.refit_iters;
  if (alt_flags) steps = std::max(steps, alt_rounds * alt_iters);
  for (int s = 0; s < steps; ++s) {
    hipLaunchKernelGGL(cyl_refit_step_kernel, grid, block, 0, ctx->stream, d_xyz, d_offsets,
                       p->d_frames, p->d_rchunks, p->d_rclus, p->d_state, p->d_member,
                       p->d_partial, p->d_pcount, s, cfg.thr, cnt);
    SLOAM_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(cyl_refit_solve_kernel, cgrid, cblock, 0, ctx->stream, p->n_clusters,
                       p->d_rclus, p->d_state, p->d_partial, p->d_pcount, s, cnt);
    SLOAM_HIP(ctx, hipGetLastError());
  }
  hipLaunchKernelGGL(cyl_refit_final_kernel, grid, block, 0, ctx->stream, d_xyz, d_offsets,
                     p->d_frames, p->d_rchunks, p->d_rclus, p->d_state, p->d_partial,
                     p->d_pcount, cfg.thr, cnt);
  SLOAM_HIP(ctx, hipGetLastError());
  hipLaunchKernelGGL(cyl_refit_finish_kernel, cgrid, cblock, 0, ctx->stream, p->n_clusters,
                     p->d_frames, p->d_rclus, p->d_state, p->d_partial, p->d_pcount, cnt,
                     d_model);
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}

// Layout of a counted plan, made on the device: the work list (cluster c's
// hypothesis chunks at c * cpc .. c * cpc + cpc - 1, clusters in index order),
// the refit chunks and per-cluster records (prefix sums over the clusters),
// and cnt = {clusters, work items, refit chunks, overflow}.  A batch beyond
// the capacities sets overflow and zero counts (nothing runs).  One
// workgroup of 256 threads.
__global__ __launch_bounds__(256) void cyl_layout_kernel(
    const int32_t* __restrict__ offsets, const int32_t* __restrict__ d_n, int cap,
    long long cap_points, int n_hyp, int hcap, int cpc, int cap_rchunks, long long cap_words,
    int4* __restrict__ work, int4* __restrict__ rchunks, int4* __restrict__ rclus,
    int32_t* __restrict__ cnt) {
  __shared__ int wck[4], wwd[4];
  __shared__ long long s_ck, s_wd;
  __shared__ int s_over;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nin = *d_n;
  const int n = min(max(nin, 0), cap);
  if (tid == 0) {
    s_ck = 0;
    s_wd = 0;
    s_over = (nin > cap || nin < 0) ? 1 : 0;
  }
  __syncthreads();
  for (int c0 = 0; c0 < n; c0 += 256) {
    const int c = c0 + tid;
    int len = 0, nck = 0, nw = 0;
    if (c < n) {
      len = offsets[c + 1] - offsets[c];
      nck = max(1, (len + kRefitChunk - 1) / kRefitChunk);
      nw = 2 * ((len + 63) / 64);
      if (len > kRefitMaxPoints || len < 0) s_over = 1;  // the host plans' bound
    }
    // inclusive wave scans, then the waves' totals
    int sc = nck, sw = nw;
    for (int o = 1; o < 64; o <<= 1) {
      const int a = __shfl_up(sc, o, 64), b = __shfl_up(sw, o, 64);
      if (lane >= o) {
        sc += a;
        sw += b;
      }
    }
    if (lane == 63) {
      wck[wave] = sc;
      wwd[wave] = sw;
    }
    __syncthreads();
    long long cb = s_ck, wb = s_wd;
    for (int w = 0; w < wave; ++w) {
      cb += wck[w];
      wb += wwd[w];
    }
    cb += sc - nck;
    wb += sw - nw;
    if (c < n) {
      rclus[c] = make_int4((int)cb, nck, (int)wb, 0);
      for (int k = 0; k < nck; ++k)
        if (cb + k < cap_rchunks)
          rchunks[cb + k] = make_int4(c, k * kRefitChunk, min(kRefitChunk, len - k * kRefitChunk), k);
      for (int j = 0; j < cpc; ++j)
        work[(size_t)c * cpc + j] = make_int4(c, j * hcap, min(hcap, n_hyp - j * hcap), len);
    }
    __syncthreads();
    if (tid == 0) {
      for (int w = 0; w < 4; ++w) {
        s_ck += wck[w];
        s_wd += wwd[w];
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    const bool over = s_over || s_ck > cap_rchunks || s_wd > cap_words ||
                      (long long)offsets[n] - (long long)offsets[0] > cap_points;
    cnt[0] = over ? 0 : n;
    cnt[1] = over ? 0 : n * cpc;
    cnt[2] = over ? 0 : (int)s_ck;
    cnt[3] = over ? 1 : 0;
  }
}

// Enqueues the device layout of a counted plan (nothing for other plans).
static int layout_counted(sloam_cyl_plan* p, const int32_t* d_offsets) {
  if (!p->counted) return SLOAM_OK;
  sloam_ctx* ctx = p->ctx;
  const long long cap_words = (long long)p->cap_member;
  hipLaunchKernelGGL(cyl_layout_kernel, dim3(1), dim3(256), 0, ctx->stream, d_offsets, p->d_n,
                     p->n_clusters, (long long)p->cap_points, p->cfg.n_hyp, p->hcap, p->cpc,
                     p->n_rchunks, cap_words, p->d_work, p->d_rchunks, p->d_rclus, p->d_cnt);
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}

// Fallback flags of the learned fit: a cluster (n >= 3) is refitted by the
// plain method when its best verifying hypothesis holds fewer than
// min_ratio * n inliers.
__global__ __launch_bounds__(256) void cyl_decide_kernel(const int32_t* __restrict__ offsets,
                                                         int n_clusters,
                                                         const int32_t* __restrict__ counts,
                                                         int n_hyp, float min_ratio,
                                                         const int32_t* __restrict__ bcnt,
                                                         int32_t* __restrict__ flags,
                                                         int32_t* __restrict__ n_flagged) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  bool a = false;
  if (c < n_clusters && (!bcnt || c < bcnt[0])) {
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

// Rescue flags of the learned fit, after its refit: a cluster (n >= 3) that
// the verify stage kept is fitted again by the plain method when its refitted
// model is invalid or keeps fewer than min_ratio * n inliers; it joins the
// fallback flags (and their count).
__global__ __launch_bounds__(256) void cyl_rescue_kernel(const int32_t* __restrict__ offsets,
                                                         int n_clusters,
                                                         const double* __restrict__ model,
                                                         float min_ratio,
                                                         const int32_t* __restrict__ bcnt,
                                                         int32_t* __restrict__ fallback,
                                                         int32_t* __restrict__ rescue,
                                                         int32_t* __restrict__ n_flagged) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  bool a = false;
  if (c < n_clusters) {
    if (!bcnt || c < bcnt[0]) {
      const int n = offsets[c + 1] - offsets[c];
      const double* m = model + (size_t)c * SLOAM_CYL_MODEL_STRIDE;
      a = fallback[c] == 0 && n >= 3 && !(m[6] > 0.0 && m[7] >= (double)min_ratio * (double)n);
      if (a) fallback[c] = 1;
    }
    rescue[c] = a ? 1 : 0;
  }
  const uint64_t b = __builtin_amdgcn_ballot_w64(a);
  if ((threadIdx.x & 63) == 0 && b && n_flagged) atomicAdd(n_flagged, __builtin_popcountll(b));
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
// hcap and tile are set: the scoring work list and the refit chunks,
// uploaded into buffers that only grow.
static int plan_layout(sloam_cyl_plan* p, const int32_t* h_offsets, int32_t n_clusters) {
  sloam_ctx* ctx = p->ctx;
  const sloam_cyl_fit_cfg& cfg = p->cfg;
  if (h_offsets[0] != 0) return SLOAM_ERR_INVALID_ARG;
  for (int i = 0; i < n_clusters; ++i) {
    const int64_t len = (int64_t)h_offsets[i + 1] - h_offsets[i];
    if (len < 0) return SLOAM_ERR_INVALID_ARG;
    if (len > kRefitMaxPoints) return SLOAM_ERR_TOO_LARGE;  // integer-exact covariance bound
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
  // the seed: one workgroup per cluster, in each bin's order, bins interleaved
  std::vector<std::vector<int32_t>> sbins(n_xcd);
  for (int x = 0; x < n_xcd; ++x)
    for (const int4& w : bins[x])
      if (sbins[x].empty() || sbins[x].back() != w.x) sbins[x].push_back(w.x);
  size_t slongest = 0;
  for (const auto& b : sbins) slongest = std::max(slongest, b.size());
  std::vector<int32_t> seed;
  seed.reserve(slongest * n_xcd);
  for (size_t k = 0; k < slongest; ++k)
    for (int x = 0; x < n_xcd; ++x) seed.push_back(k < sbins[x].size() ? sbins[x][k] : -1);
  p->n_seed = (int)seed.size();
  size_t longest = 0;
  for (const auto& b : bins) longest = std::max(longest, b.size());
  std::vector<int4> work;
  work.reserve(longest * n_xcd);
  for (size_t k = 0; k < longest; ++k)
    for (int x = 0; x < n_xcd; ++x)
      work.push_back(k < bins[x].size() ? bins[x][k] : make_int4(-1, 0, 0, 0));
  p->n_work = (int)work.size();
  // refit chunks: every cluster (even an empty one) has at least one; the
  // clusters are dealt to per-XCD bins as above and the bins interleaved, so
  // all chunks of a cluster run on one XCD
  std::vector<std::vector<int4>> rbins(n_xcd);
  std::vector<long long> rload(n_xcd, 0);
  std::vector<int4> rclus(n_clusters);
  int n_chunk_ids = 0;
  size_t n_words = 0;
  for (int c = 0; c < n_clusters; ++c) {
    const int len = h_offsets[c + 1] - h_offsets[c];
    const int nck = std::max(1, (len + kRefitChunk - 1) / kRefitChunk);
    rclus[c] = make_int4(n_chunk_ids, nck, (int)n_words, 0);
    n_chunk_ids += nck;
    n_words += 2 * (size_t)((len + 63) / 64);
  }
  for (int c : order) {
    const int len = h_offsets[c + 1] - h_offsets[c];
    const int x = (int)(std::min_element(rload.begin(), rload.end()) - rload.begin());
    for (int k = 0; k < rclus[c].y; ++k) {
      const int first = k * kRefitChunk;
      rbins[x].push_back(make_int4(c, first, std::min(kRefitChunk, len - first), k));
    }
    rload[x] += std::max(len, 1);
  }
  size_t rlongest = 0;
  for (const auto& b : rbins) rlongest = std::max(rlongest, b.size());
  std::vector<int4> rchunks;
  rchunks.reserve(rlongest * n_xcd);
  for (size_t k = 0; k < rlongest; ++k)
    for (int x = 0; x < n_xcd; ++x)
      rchunks.push_back(k < rbins[x].size() ? rbins[x][k] : make_int4(-1, 0, 0, 0));
  p->n_rchunks = (int)rchunks.size();
  int max_chunks = 0;
  for (int c = 0; c < n_clusters; ++c) max_chunks = std::max(max_chunks, rclus[c].y);
  p->fused = n_clusters <= kRefitFusedPlanClusters && max_chunks <= kRefitFusedPlanChunks;
  const size_t wb = sizeof(int4) * work.size(), cb = sizeof(int4) * rchunks.size(),
               lb = sizeof(int4) * rclus.size(), sb = sizeof(int32_t) * seed.size();
  int st = grow(ctx, &p->d_tables, &p->cap_tables, wb + cb + lb + sb, 1);
  if (!st) st = grow(ctx, (void**)&p->d_frames, &p->cap_frames, n_clusters, sizeof(CylFrame));
  if (!st)
    st = grow(ctx, (void**)&p->d_counts, &p->cap_counts, (size_t)n_clusters * cfg.n_hyp,
              sizeof(int32_t));
  if (!st)
    st = grow(ctx, (void**)&p->d_state, &p->cap_state, n_clusters, sizeof(RefitState));
  if (!st) st = grow(ctx, (void**)&p->d_member, &p->cap_member, std::max<size_t>(n_words, 1),
                     sizeof(uint32_t));
  if (!st) st = grow(ctx, (void**)&p->d_rescue, &p->cap_rescue, n_clusters, sizeof(int32_t));
  if (!st && (size_t)n_chunk_ids > p->cap_partial) {
    sloam_dev_free(p->d_pcount);
    p->d_pcount = nullptr;
    st = grow(ctx, (void**)&p->d_partial, &p->cap_partial, n_chunk_ids,
              sizeof(Comp2) * kRefitTerms);
    if (!st) st = sloam_dev_alloc(ctx, (void**)&p->d_pcount, sizeof(int32_t) * p->cap_partial);
  }
  if (st) return st;
  p->d_work = static_cast<int4*>(p->d_tables);
  p->d_rchunks = reinterpret_cast<int4*>(static_cast<char*>(p->d_tables) + wb);
  p->d_rclus = reinterpret_cast<int4*>(static_cast<char*>(p->d_tables) + wb + cb);
  p->d_seed_order = reinterpret_cast<int32_t*>(static_cast<char*>(p->d_tables) + wb + cb + lb);
  // the tables through the pinned staging buffer, in one copy (the caller
  // has made sure that no earlier upload from it is pending)
  const size_t tb = wb + cb + lb + sb;
  if (tb > p->cap_stage) {
    if (p->h_stage) (void)hipHostFree(p->h_stage);
    p->h_stage = nullptr;
    p->cap_stage = 0;
    const size_t sz = tb + tb / 2;
    hipError_t e = hipHostMalloc(&p->h_stage, sz);
    if (e != hipSuccess) return sloam_record_hip_error(ctx, e, "hipHostMalloc(plan staging)");
    p->cap_stage = sz;
  }
  char* stage = static_cast<char*>(p->h_stage);
  if (wb) std::memcpy(stage, work.data(), wb);
  if (cb) std::memcpy(stage + wb, rchunks.data(), cb);
  if (lb) std::memcpy(stage + wb + cb, rclus.data(), lb);
  if (sb) std::memcpy(stage + wb + cb + lb, seed.data(), sb);
  if (tb) SLOAM_HIP(ctx, hipMemcpyAsync(p->d_tables, stage, tb, hipMemcpyHostToDevice, ctx->stream));
  return SLOAM_OK;
}

extern "C" {

}  // extern "C"

// A plan with cfg's launch shapes and no batch yet (nullptr when cfg is
// invalid; *st says why).
static sloam_cyl_plan* plan_new(sloam_ctx* ctx, const sloam_cyl_fit_cfg* cfg, int* st) {
  *st = SLOAM_ERR_INVALID_ARG;
  if (cfg->n_hyp < 1 || cfg->n_hyp > (1 << 20) || !(cfg->thr > 0.0f) ||
      !(cfg->max_radius > 0.0f) || cfg->refit_iters < 0 || cfg->refit_rounds < 0 ||
      cfg->hyp_chunk < 0 || cfg->tile_points < 0)
    return nullptr;
  sloam_cyl_plan* p = new (std::nothrow) sloam_cyl_plan();
  if (!p) {
    *st = SLOAM_ERR_ALLOC;
    return nullptr;
  }
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
    *st = SLOAM_ERR_TOO_LARGE;
    return nullptr;
  }
  p->tile = std::min(p->tile, ((kMaxLds - hyp_bytes) / 12) / 64 * 64);
  int n_xcd = 1;
  if (hipDeviceGetAttribute(&n_xcd, hipDeviceAttributeNumberOfXccs, ctx->device) != hipSuccess ||
      n_xcd < 1)
    n_xcd = 1;
  (void)hipGetLastError();
  p->n_xcd = n_xcd;
  *st = SLOAM_OK;
  return p;
}

extern "C" {

int sloam_cyl_plan_create(sloam_ctx* ctx, const int32_t* h_offsets, int32_t n_clusters,
                          const sloam_cyl_fit_cfg* cfg, sloam_cyl_plan** out) {
  if (!ctx || !h_offsets || !cfg || !out || n_clusters < 0) return SLOAM_ERR_INVALID_ARG;
  *out = nullptr;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  int st = SLOAM_OK;
  sloam_cyl_plan* p = plan_new(ctx, cfg, &st);
  if (!p) return st;
  st = plan_layout(p, h_offsets, n_clusters);
  if (st) {
    sloam_cyl_plan_destroy(p);
    return st;
  }
  *out = p;
  return SLOAM_OK;
}

int sloam_cyl_plan_create_counted(sloam_ctx* ctx, int32_t cap_clusters, int64_t cap_points,
                                  const int32_t* d_n_clusters, const sloam_cyl_fit_cfg* cfg,
                                  sloam_cyl_plan** out) {
  if (!ctx || !cfg || !out || !d_n_clusters || cap_clusters < 1 || cap_points < 0 ||
      cap_points > ((int64_t)1 << 31) - 1 || cap_clusters > (1 << 20))
    return SLOAM_ERR_INVALID_ARG;
  *out = nullptr;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  int st = SLOAM_OK;
  sloam_cyl_plan* p = plan_new(ctx, cfg, &st);
  if (!p) return st;
  p->counted = true;
  p->fused = true;
  p->d_n = d_n_clusters;
  p->cap_points = cap_points;
  p->cpc = (cfg->n_hyp + p->hcap - 1) / p->hcap;
  p->n_clusters = cap_clusters;
  p->n_points = cap_points;
  p->n_work = cap_clusters * p->cpc;
  p->n_rchunks = (int)(cap_clusters + cap_points / kRefitChunk + 1);
  const size_t words = (size_t)(cap_points / 32) + 2 * (size_t)cap_clusters + 2;
  const size_t tb = sizeof(int4) * ((size_t)p->n_work + p->n_rchunks + cap_clusters);
  st = grow(ctx, &p->d_tables, &p->cap_tables, tb, 1);
  if (!st) st = grow(ctx, (void**)&p->d_frames, &p->cap_frames, cap_clusters, sizeof(CylFrame));
  if (!st)
    st = grow(ctx, (void**)&p->d_counts, &p->cap_counts, (size_t)cap_clusters * cfg->n_hyp,
              sizeof(int32_t));
  if (!st)
    st = grow(ctx, (void**)&p->d_state, &p->cap_state, cap_clusters, sizeof(RefitState));
  if (!st) st = grow(ctx, (void**)&p->d_member, &p->cap_member, words, sizeof(uint32_t));
  if (!st) st = grow(ctx, (void**)&p->d_rescue, &p->cap_rescue, cap_clusters, sizeof(int32_t));
  if (!st)
    st = grow(ctx, (void**)&p->d_partial, &p->cap_partial, p->n_rchunks,
              sizeof(Comp2) * kRefitTerms);
  if (!st) st = sloam_dev_alloc(ctx, (void**)&p->d_pcount, sizeof(int32_t) * p->cap_partial);
  if (!st) st = sloam_dev_alloc(ctx, (void**)&p->d_cnt, sizeof(int32_t) * 4);
  if (st) {
    sloam_cyl_plan_destroy(p);
    return st;
  }
  p->d_work = static_cast<int4*>(p->d_tables);
  p->d_rchunks = p->d_work + p->n_work;
  p->d_rclus = p->d_rchunks + p->n_rchunks;
  *out = p;
  return SLOAM_OK;
}

int sloam_cyl_plan_counts(const sloam_cyl_plan* p, const int32_t** d_counts) {
  if (!p || !d_counts || !p->counted) return SLOAM_ERR_INVALID_ARG;
  *d_counts = p->d_cnt;
  return SLOAM_OK;
}

int sloam_cyl_plan_reset(sloam_cyl_plan* p, const int32_t* h_offsets, int32_t n_clusters) {
  if (!p || !h_offsets || n_clusters < 0 || p->counted) return SLOAM_ERR_INVALID_ARG;
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
    p->n_rchunks = 0;
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
  if (verify->counted != plain->counted ||
      (verify->counted && (verify->d_n != plain->d_n || verify->cap_points != plain->cap_points)))
    return SLOAM_ERR_INVALID_ARG;
  int st0 = layout_counted(verify, d_offsets);
  if (!st0) st0 = layout_counted(plain, d_offsets);
  if (st0) return st0;
  const int32_t* vcnt = verify->counted ? verify->d_cnt : nullptr;
  {
    // one axis seed per cluster serves both hypothesis sets
    SloamKernelScope ks(ctx, "cyl_seed");
    const bool ord = !verify->counted && verify->n_seed > 0;
    hipLaunchKernelGGL(cyl_seed_kernel, dim3(ord ? verify->n_seed : S), dim3(kBlock), 0,
                       ctx->stream, d_xyz, d_offsets, S, nullptr, vcnt, verify->d_frames,
                       ord ? verify->d_seed_order : nullptr);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  if (verify->n_work > 0) {
    const size_t lds =
        (size_t)verify->hcap * (sizeof(float4) + sizeof(int)) + 3 * sizeof(float) * verify->tile;
    SloamKernelScope ks(ctx, "cyl_score");
    hipLaunchKernelGGL(cyl_score_kernel, dim3(verify->n_work), dim3(kBlock), lds, ctx->stream,
                       d_xyz, d_offsets, verify->d_frames, verify->d_work, vc.n_hyp, verify->hcap,
                       verify->tile, vc.thr, vc.max_radius, k0, k1, (uint32_t)vc.stream_tag,
                       d_proposal, nullptr, vcnt, cv, nullptr);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  {
    SloamKernelScope ks(ctx, "cyl_decide");
    if (d_n_fallback) {
      const int st = sloam_fill_u32(ctx, (uint32_t*)d_n_fallback, 1, 0u);
      if (st) return st;
    }
    hipLaunchKernelGGL(cyl_decide_kernel, dim3((S + 255) / 256), dim3(256), 0, ctx->stream,
                       d_offsets, S, cv, vc.n_hyp, min_inlier_ratio, vcnt, d_fallback,
                       d_n_fallback);
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
                       nullptr, d_fallback, plain->counted ? plain->d_cnt : nullptr, cp, nullptr);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  {
    SloamKernelScope ks(ctx, "cyl_refit");
    const int st = launch_refit(verify, d_xyz, d_offsets, cv, d_proposal, nullptr, d_fallback, cp,
                                pc.n_hyp, (uint32_t)pc.stream_tag, pc.refit_rounds,
                                pc.refit_iters, nullptr, d_best, d_model);
    if (st) return st;
  }
  {
    // rescue: the verified clusters whose refit failed or kept fewer than
    // min_inlier_ratio * n inliers take the plain method's fit after all
    // (scored with the plain hypotheses in the same seed frames, refitted
    // alone in the fused form)
    SloamKernelScope ks(ctx, "cyl_rescue");
    hipLaunchKernelGGL(cyl_rescue_kernel, dim3((S + 255) / 256), dim3(256), 0, ctx->stream,
                       d_offsets, S, d_model, min_inlier_ratio, vcnt, d_fallback,
                       verify->d_rescue, d_n_fallback);
    SLOAM_HIP(ctx, hipGetLastError());
    if (plain->n_work > 0) {
      const size_t lds =
          (size_t)plain->hcap * (sizeof(float4) + sizeof(int)) + 3 * sizeof(float) * plain->tile;
      hipLaunchKernelGGL(cyl_score_kernel, dim3(plain->n_work), dim3(kBlock), lds, ctx->stream,
                         d_xyz, d_offsets, verify->d_frames, plain->d_work, pc.n_hyp,
                         plain->hcap, plain->tile, pc.thr, pc.max_radius, k0, k1,
                         (uint32_t)pc.stream_tag, nullptr, verify->d_rescue,
                         plain->counted ? plain->d_cnt : nullptr, cp, nullptr);
      SLOAM_HIP(ctx, hipGetLastError());
    }
    const int st = launch_refit(verify, d_xyz, d_offsets, cv, nullptr, verify->d_rescue,
                                verify->d_rescue, cp, pc.n_hyp, (uint32_t)pc.stream_tag,
                                pc.refit_rounds, pc.refit_iters, nullptr, d_best, d_model,
                                /*force_fused=*/true);
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
  sloam_dev_free(p->d_state);
  sloam_dev_free(p->d_member);
  sloam_dev_free(p->d_partial);
  sloam_dev_free(p->d_pcount);
  sloam_dev_free(p->d_rescue);
  sloam_dev_free(p->d_cnt);
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
  if (p->counted && d_mask) return SLOAM_ERR_INVALID_ARG;  // its size is the batch's
  int st0 = layout_counted(p, d_offsets);
  if (st0) return st0;
  const int32_t* cnt = p->counted ? p->d_cnt : nullptr;
  if (d_mask) {
    int st = sloam_fill_u32(ctx, d_mask, (size_t)((p->n_points + 31) / 32), 0u);
    if (st) return st;
  }
  {
    SloamKernelScope ks(ctx, "cyl_seed");
    const bool ord = !p->counted && p->n_seed > 0;
    hipLaunchKernelGGL(cyl_seed_kernel, dim3(ord ? p->n_seed : p->n_clusters), dim3(kBlock), 0,
                       ctx->stream, d_xyz, d_offsets, p->n_clusters, d_active, cnt, p->d_frames,
                       ord ? p->d_seed_order : nullptr);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  if (p->n_work > 0) {
    const size_t lds = (size_t)p->hcap * (sizeof(float4) + sizeof(int)) + 3 * sizeof(float) * p->tile;
    SloamKernelScope ks(ctx, "cyl_score");
    hipLaunchKernelGGL(cyl_score_kernel, dim3(p->n_work), dim3(kBlock), lds, ctx->stream, d_xyz,
                       d_offsets, p->d_frames, p->d_work, cfg.n_hyp, p->hcap, p->tile, cfg.thr,
                       cfg.max_radius, k0, k1, tag, d_proposal, d_active, cnt, counts, d_samples);
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
