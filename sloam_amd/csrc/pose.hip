// Batched pose estimation (Levenberg-Marquardt over point-to-cylinder and
// point-to-plane residuals) for gfx950.
//
// Reference: sloam::sloam::OptimizePose (sloam/src/core/sloam.cpp:174-255) with
// CylinderCost (sloam/include/objects/cylinder.h:90-132) and PlaneCost
// (sloam/include/objects/plane.h:76-116), HuberLoss(0.1), Ceres LM.
//
// Per LM iteration two kernels run on the context stream:
//   pose_eval   one workgroup per chunk of 256 * kRpt residuals of one problem:
//               FP32 residual r, 1x6 Jacobian [d/dt, d/dtheta] (left
//               perturbation), Huber IRLS weight w = min(1, delta/|r|); sums
//               J'WJ (21), J'Wr (6) and the cost (1) per thread in FP32, then an
//               error-free (TwoSum) tree over the workgroup; the workgroup's
//               28 (sum, compensation) pairs go to HBM as FP64.
//   pose_solve  one wave per problem: FP64 sum of the chunk partials in chunk
//               order, then the LM trust-region logic and the FP64 6x6
//               Cholesky solve of (J'WJ + D/radius) dx = -J'Wr; writes the next
//               pose to evaluate.
// pose_fused    one workgroup per problem runs every iteration in one launch:
//               the same chunk evaluations and the same FP64 sum of chunk
//               partials in chunk order, then the same LM step in thread 0,
//               so its results are bit-identical to the two-kernel loop.  It
//               serves small problems (a keyframe's solve), where the loop's
//               2 * (max_iters + 1) launches cost more than the work.
// Coordinates are evaluated relative to a per-problem anchor (the initial
// translation) so that FP32 residuals keep sub-micrometre resolution.
//
// Precision FP64 (sloam_pose_cfg.precision 1) evaluates residuals, Jacobians
// and sums in FP64 with the landmarks in FP64 and absolute coordinates, the
// arithmetic Ceres does: the solves of one keyframe use it, so that their
// termination tests (relative cost change against 1e-6) see the same costs as
// Ceres'.  FP32 residuals carry rounding of ~1e-6 of the cost of a keyframe's
// problem, the size of that tolerance, which moves the iteration at which an
// LM run stops and, in a flat valley of the cost, its result by ~1e-4.
#include "common.hpp"
#include "detmath.h"

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

namespace {

constexpr int kPBlock = 256;
constexpr int kRpt = 8;  // residuals per thread per workgroup
constexpr int kSums = 28;
// the chunk reduction: kParts threads per sum add kPBlock / kParts lanes'
// values each, then one thread per sum adds the kParts partials
constexpr int kParts = 8;
static_assert(kSums * kParts <= kPBlock && kPBlock % kParts == 0, "chunk reduction layout");

// LDS of the chunk reduction: every thread's sums, then the partials
struct ChunkRed {
  float x[kSums][kPBlock + 1];
  float ps[kSums][kParts];
  float pc[kSums][kParts];
};

// the same in FP64 (plain sums in the same order)
struct ChunkRed64 {
  double x[kSums][kPBlock + 1];
  double ps[kSums][kParts];
};

struct LMState {
  double x[7];   // accepted pose: t, q (x, y, z, w)
  double xe[7];  // pose to evaluate next
  double anchor[3];
  double F0, F;
  double A[21];
  double g[6];
  double radius, factor;
  double model, dn, xn;  // of the pending candidate
  double aa[6], aae[6];   // angle-axis parametrization: accepted / to evaluate [t, w]
  int32_t iter, term, done, phase;  // phase 0: next evaluation is the initial one
  int32_t free_mask;                // bit i: parameter i is optimised
  int32_t pad;
};

struct PoseCfgDev {
  int32_t max_iters, stop_conv, param;  // param 0: quaternion / left perturbation, 1: angle-axis
  double delta, ftol, gtol, ptol, radius0, min_rel;
};

// --------------------------------------------------------------------------
// small FP64 helpers (one thread)
// --------------------------------------------------------------------------
__device__ inline void quat_to_R(const double q[4], double R[3][3]) {
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  R[0][0] = 1 - 2 * (y * y + z * z);
  R[0][1] = 2 * (x * y - z * w);
  R[0][2] = 2 * (x * z + y * w);
  R[1][0] = 2 * (x * y + z * w);
  R[1][1] = 1 - 2 * (x * x + z * z);
  R[1][2] = 2 * (y * z - x * w);
  R[2][0] = 2 * (x * z - y * w);
  R[2][1] = 2 * (y * z + x * w);
  R[2][2] = 1 - 2 * (x * x + y * y);
}

__device__ inline void quat_left_update(const double q[4], const double dth[3], double out[4]) {
  const double ang = sqrt(dth[0] * dth[0] + dth[1] * dth[1] + dth[2] * dth[2]);
  double d[4];
  if (ang > 1e-300) {
    double sh, ch;
    sloam_dm::sincos(0.5 * ang, &sh, &ch);
    const double s = sh / ang;
    d[0] = dth[0] * s;
    d[1] = dth[1] * s;
    d[2] = dth[2] * s;
    d[3] = ch;
  } else {
    d[0] = 0.5 * dth[0];
    d[1] = 0.5 * dth[1];
    d[2] = 0.5 * dth[2];
    d[3] = 1.0;
  }
  const double x = d[3] * q[0] + d[0] * q[3] + d[1] * q[2] - d[2] * q[1];
  const double y = d[3] * q[1] - d[0] * q[2] + d[1] * q[3] + d[2] * q[0];
  const double z = d[3] * q[2] + d[0] * q[1] - d[1] * q[0] + d[2] * q[3];
  const double w = d[3] * q[3] - d[0] * q[0] - d[1] * q[1] - d[2] * q[2];
  const double n = sqrt(x * x + y * y + z * z + w * w);
  out[0] = x / n;
  out[1] = y / n;
  out[2] = z / n;
  out[3] = w / n;
}

// angle-axis <-> quaternion (w, x, y, z order internally)
__device__ inline void quat_to_aa(const double qxyzw[4], double aa[3]) {
  const double q1 = qxyzw[0], q2 = qxyzw[1], q3 = qxyzw[2], c = qxyzw[3];
  const double s2 = q1 * q1 + q2 * q2 + q3 * q3;
  if (s2 > 0.0) {
    const double st = sqrt(s2);
    const double two_theta =
        2.0 * ((c < 0.0) ? sloam_dm::atan2(-st, -c) : sloam_dm::atan2(st, c));
    const double k = two_theta / st;
    aa[0] = q1 * k;
    aa[1] = q2 * k;
    aa[2] = q3 * k;
  } else {
    aa[0] = 2.0 * q1;
    aa[1] = 2.0 * q2;
    aa[2] = 2.0 * q3;
  }
}

__device__ inline void aa_to_quat(const double aa[3], double qxyzw[4]) {
  const double t2 = aa[0] * aa[0] + aa[1] * aa[1] + aa[2] * aa[2];
  if (t2 > 0.0) {
    const double th = sqrt(t2);
    const double h = 0.5 * th;
    double sh, ch;
    sloam_dm::sincos(h, &sh, &ch);
    const double k = sh / th;
    qxyzw[0] = aa[0] * k;
    qxyzw[1] = aa[1] * k;
    qxyzw[2] = aa[2] * k;
    qxyzw[3] = ch;
  } else {
    qxyzw[0] = 0.5 * aa[0];
    qxyzw[1] = 0.5 * aa[1];
    qxyzw[2] = 0.5 * aa[2];
    qxyzw[3] = 1.0;
  }
}

// left Jacobian of SO(3): exp(w + dw) ~ exp(Jl(w) dw) exp(w)
__device__ inline void left_jacobian(const double w[3], double J[3][3]) {
  const double t2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  double a, b;
  if (t2 < 1e-12) {
    a = 0.5 - t2 / 24.0;
    b = 1.0 / 6.0 - t2 / 120.0;
  } else {
    const double t = sqrt(t2);
    double st, ct;
    sloam_dm::sincos(t, &st, &ct);
    a = (1.0 - ct) / t2;
    b = (t - st) / (t2 * t);
  }
  const double K[3][3] = {{0.0, -w[2], w[1]}, {w[2], 0.0, -w[0]}, {-w[1], w[0], 0.0}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double k2 = 0.0;
      for (int m = 0; m < 3; ++m) k2 += K[i][m] * K[m][j];
      J[i][j] = (i == j ? 1.0 : 0.0) + a * K[i][j] + b * k2;
    }
}

__device__ inline bool chol6(const double A[6][6], const double b[6], double x[6]) {
  double L[6][6];
  bool ok = true;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    double s = A[j][j];
#pragma unroll
    for (int k = 0; k < j; ++k) s = s - L[j][k] * L[j][k];
    ok = ok && (s > 0.0);
    const double d = sqrt(s);
    L[j][j] = d;
#pragma unroll
    for (int i = j + 1; i < 6; ++i) {
      double t = A[i][j];
#pragma unroll
      for (int k = 0; k < j; ++k) t = t - L[i][k] * L[j][k];
      L[i][j] = t / d;
    }
  }
  double y[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double t = b[i];
#pragma unroll
    for (int k = 0; k < i; ++k) t = t - L[i][k] * y[k];
    y[i] = t / L[i][i];
  }
#pragma unroll
  for (int i = 5; i >= 0; --i) {
    double t = y[i];
#pragma unroll
    for (int k = i + 1; k < 6; ++k) t = t - L[k][i] * x[k];
    x[i] = t / L[i][i];
  }
  return ok;
}

// error-free transformation: a + b = s + e exactly
__device__ inline void two_sum(float a, float b, float* s, float* e) {
  const float ss = a + b;
  const float bp = ss - a;
  *e = (a - (ss - bp)) + (b - bp);
  *s = ss;
}

// --------------------------------------------------------------------------
// prep: LM state init + landmark tables: FP32 relative to the anchor, or FP64
// absolute (F64)
// --------------------------------------------------------------------------
// d_n_problems (nullable): the problem count in device memory (a grid of
// n_problems workgroups then serves any count up to n_problems)
template <bool F64>
__global__ __launch_bounds__(kPBlock) void pose_prep_kernel(
    int n_problems, const int32_t* __restrict__ d_n_problems, const double* __restrict__ pose_in,
    const int32_t* __restrict__ free_mask, const double* __restrict__ cyl,
    const int32_t* __restrict__ cyl_off, const double* __restrict__ pl,
    const int32_t* __restrict__ pl_off, PoseCfgDev cfg, LMState* __restrict__ st,
    void* __restrict__ cyl_tab, void* __restrict__ pl_tab) {
  const int k = blockIdx.x;
  if (k >= n_problems || (d_n_problems && k >= *d_n_problems)) return;
  const double* x0 = pose_in + 7 * (size_t)k;
  const double ax = x0[0], ay = x0[1], az = x0[2];
  for (int i = cyl_off[k] + (int)threadIdx.x; i < cyl_off[k + 1]; i += kPBlock) {
    const double* c = cyl + 8 * (size_t)i;
    const double rn = sqrt(c[3] * c[3] + c[4] * c[4] + c[5] * c[5]);
    if constexpr (F64) {
      double4* t = static_cast<double4*>(cyl_tab) + 2 * (size_t)i;
      t[0] = make_double4(c[0], c[1], c[2], c[6]);
      t[1] = make_double4(c[3] / rn, c[4] / rn, c[5] / rn, 0.0);
    } else {
      float4* t = static_cast<float4*>(cyl_tab) + 2 * (size_t)i;
      t[0] = make_float4((float)(c[0] - ax), (float)(c[1] - ay), (float)(c[2] - az), (float)c[6]);
      t[1] = make_float4((float)(c[3] / rn), (float)(c[4] / rn), (float)(c[5] / rn), 0.f);
    }
  }
  for (int i = pl_off[k] + (int)threadIdx.x; i < pl_off[k + 1]; i += kPBlock) {
    const double* q = pl + 4 * (size_t)i;
    const double nn = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    if constexpr (F64) {
      static_cast<double4*>(pl_tab)[i] = make_double4(q[0] / nn, q[1] / nn, q[2] / nn, q[3] / nn);
    } else {
      const double dr = (q[3] + q[0] * ax + q[1] * ay + q[2] * az) / nn;
      static_cast<float4*>(pl_tab)[i] =
          make_float4((float)(q[0] / nn), (float)(q[1] / nn), (float)(q[2] / nn), (float)dr);
    }
  }
  if (threadIdx.x == 0) {
    LMState s;
    for (int i = 0; i < 7; ++i) {
      s.x[i] = x0[i];
      s.xe[i] = x0[i];
    }
    s.anchor[0] = ax;
    s.anchor[1] = ay;
    s.anchor[2] = az;
    s.F0 = s.F = 0.0;
    for (int i = 0; i < 21; ++i) s.A[i] = 0.0;
    for (int i = 0; i < 6; ++i) s.g[i] = 0.0;
    s.radius = cfg.radius0;
    s.factor = 2.0;
    s.model = s.dn = s.xn = 0.0;
    s.iter = 0;
    s.term = 1;
    s.done = 0;
    s.phase = 0;
    s.free_mask = free_mask ? free_mask[k] : 0x3f;
    s.pad = 0;
    s.aa[0] = s.aae[0] = x0[0];
    s.aa[1] = s.aae[1] = x0[1];
    s.aa[2] = s.aae[2] = x0[2];
    {
      double w[3];
      quat_to_aa(x0 + 3, w);
      for (int i = 0; i < 3; ++i) s.aa[3 + i] = s.aae[3 + i] = w[i];
    }
    st[k] = s;
  }
}

// --------------------------------------------------------------------------
// eval: residuals, Jacobians, compensated partial sums
// --------------------------------------------------------------------------
// pose to evaluate, relative to the anchor, in FP32; for the angle-axis
// parametrization the rotation columns of the Jacobian are mapped by the left
// Jacobian Jl(w) and fixed parameters are masked (sub: either applies)
struct EvalFrame {
  float R[3][3], t[3], Jl[3][3], msk[6];
  int sub;
};

__device__ inline void eval_frame(const LMState& s, int param, EvalFrame& F) {
  double q[4] = {s.xe[3], s.xe[4], s.xe[5], s.xe[6]};
  double Jd[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  if (param == 1) {
    const double w[3] = {s.aae[3], s.aae[4], s.aae[5]};
    aa_to_quat(w, q);
    left_jacobian(w, Jd);
  }
  double Rd[3][3];
  quat_to_R(q, Rd);
  const double* tt = param == 1 ? s.aae : s.xe;
  const int fm = s.free_mask;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      F.R[i][j] = (float)Rd[i][j];
      F.Jl[i][j] = (float)Jd[i][j];
    }
    F.t[i] = (float)(tt[i] - s.anchor[i]);
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) F.msk[i] = ((fm >> i) & 1) ? 1.0f : 0.0f;
  F.sub = param == 1 || fm != 0x3f;
}

// One chunk [b, e) of at most kPBlock * kRpt residuals over a workgroup of
// kPBlock threads: thread t takes residuals b + j * kPBlock + t.  Returns the
// chunk's 28 sums as FP64 (sum + compensation) in threads 0..27.
//
// The kernel is bound by VALU issue (~160 instructions per residual against
// 16 bytes of HBM).  kSub false (the quaternion parametrization, all
// parameters free) drops the left-Jacobian map and the masks, whose identity
// and unit factors change no bit.  The divides and the square root stay
// correctly rounded: the hardware reciprocals would save ~35 instructions per
// residual but move the LM path of small ill-conditioned problems by more
// than 1e-4 within a fixed iteration count.
template <bool kSub>
__device__ inline double eval_chunk(const float4* __restrict__ pts, int b, int e,
                                    const EvalFrame& F, const float4* __restrict__ cylf,
                                    const float4* __restrict__ plf, float delta, ChunkRed& red) {
  const float d2 = delta * delta;
  float acc[kSums];
#pragma unroll
  for (int i = 0; i < kSums; ++i) acc[i] = 0.f;
  // the thread's residual loads are all issued before any arithmetic, and
  // each residual's landmark load one residual ahead of its use, so that a
  // wave keeps several loads in flight instead of waiting on one dependent
  // pair per residual (loading every landmark up front costs more in
  // occupancy than it gains)
  float4 pv[kRpt];
#pragma unroll
  for (int j = 0; j < kRpt; ++j) {
    const int i = b + j * kPBlock + (int)threadIdx.x;
    pv[j] = i < e ? pts[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  const int n_mine = e > b + (int)threadIdx.x ? (e - b - (int)threadIdx.x + kPBlock - 1) / kPBlock
                                               : 0;
  auto gather = [&](int j, float4& A, float4& B) {
    const int id = __float_as_int(pv[j].w);
    if (id >= 0) {
      A = cylf[2 * (size_t)id];
      B = cylf[2 * (size_t)id + 1];
    } else {
      A = plf[-id - 1];
      B = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  };
  float4 na = make_float4(0.f, 0.f, 0.f, 0.f), nb = na;
  if (n_mine > 0) gather(0, na, nb);
#pragma unroll
  for (int j = 0; j < kRpt; ++j) {
    if (j >= n_mine) break;
    const float4 p = pv[j];
    const int id = __float_as_int(p.w);
    const float4 L0 = na, L1 = nb;
    if (j + 1 < n_mine) gather(j + 1, na, nb);
    float rp[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
      rp[a] = __builtin_fmaf(F.R[a][0], p.x, __builtin_fmaf(F.R[a][1], p.y, F.R[a][2] * p.z));
    const float lp[3] = {rp[0] + F.t[0], rp[1] + F.t[1], rp[2] + F.t[2]};
    float dir[3], r;
    if (id >= 0) {
      const float v0 = lp[0] - L0.x, v1 = lp[1] - L0.y, v2 = lp[2] - L0.z;
      const float h = __builtin_fmaf(v0, L1.x, __builtin_fmaf(v1, L1.y, v2 * L1.z));
      const float w0 = __builtin_fmaf(-h, L1.x, v0);
      const float w1 = __builtin_fmaf(-h, L1.y, v1);
      const float w2 = __builtin_fmaf(-h, L1.z, v2);
      const float rho = sqrtf(__builtin_fmaf(w0, w0, __builtin_fmaf(w1, w1, w2 * w2)));
      r = rho - L0.w;
      const float inv = rho > 0.f ? 1.0f / rho : 0.f;
      dir[0] = w0 * inv;
      dir[1] = w1 * inv;
      dir[2] = w2 * inv;
    } else {
      const float4 L = L0;
      const float s = __builtin_fmaf(L.x, lp[0], __builtin_fmaf(L.y, lp[1],
                                                               __builtin_fmaf(L.z, lp[2], L.w)));
      r = fabsf(s);
      const float sg = s < 0.f ? -1.f : 1.f;
      dir[0] = sg * L.x;
      dir[1] = sg * L.y;
      dir[2] = sg * L.z;
    }
    float J[6];
    {
      const float jt0 = __builtin_fmaf(rp[1], dir[2], -rp[2] * dir[1]);
      const float jt1 = __builtin_fmaf(rp[2], dir[0], -rp[0] * dir[2]);
      const float jt2 = __builtin_fmaf(rp[0], dir[1], -rp[1] * dir[0]);
      if constexpr (kSub) {
        J[0] = dir[0] * F.msk[0];
        J[1] = dir[1] * F.msk[1];
        J[2] = dir[2] * F.msk[2];
#pragma unroll
        for (int c = 0; c < 3; ++c)
          J[3 + c] =
              __builtin_fmaf(jt0, F.Jl[0][c], __builtin_fmaf(jt1, F.Jl[1][c], jt2 * F.Jl[2][c])) *
              F.msk[3 + c];
      } else {
        J[0] = dir[0];
        J[1] = dir[1];
        J[2] = dir[2];
        J[3] = jt0;
        J[4] = jt1;
        J[5] = jt2;
      }
    }
    const float s2 = r * r;
    float wgt, rho;
    if (s2 <= d2) {
      wgt = 1.f;
      rho = s2;
    } else {
      const float ar = fabsf(r);
      wgt = delta / ar;
      rho = __builtin_fmaf(2.f * delta, ar, -d2);
    }
    float wJ[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) wJ[a] = wgt * J[a];
    int tt = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int bb = a; bb < 6; ++bb) {
        acc[tt] = __builtin_fmaf(wJ[a], J[bb], acc[tt]);
        ++tt;
      }
#pragma unroll
    for (int a = 0; a < 6; ++a) acc[21 + a] = __builtin_fmaf(wJ[a], r, acc[21 + a]);
    acc[27] = __builtin_fmaf(0.5f, rho, acc[27]);
  }
  // error-free reduction in a fixed order: every thread's sums to LDS; each of
  // kParts threads per sum runs a TwoSum chain over kPBlock / kParts lanes;
  // one thread per sum chains the kParts (sum, compensation) pairs
#pragma unroll
  for (int i = 0; i < kSums; ++i) red.x[i][threadIdx.x] = acc[i];
  __syncthreads();
  if (threadIdx.x < kSums * kParts) {
    const int i = threadIdx.x % kSums, part = threadIdx.x / kSums;
    const float* row = &red.x[i][part * (kPBlock / kParts)];
    float sm = row[0], c = 0.f;
    for (int j = 1; j < kPBlock / kParts; ++j) {
      float ns, err;
      two_sum(sm, row[j], &ns, &err);
      sm = ns;
      c += err;
    }
    red.ps[i][part] = sm;
    red.pc[i][part] = c;
  }
  __syncthreads();
  double out = 0.0;
  if (threadIdx.x < kSums) {
    const int i = threadIdx.x;
    float sm = red.ps[i][0], c = red.pc[i][0];
    for (int v = 1; v < kParts; ++v) {
      float ns, err;
      two_sum(sm, red.ps[i][v], &ns, &err);
      sm = ns;
      c = c + red.pc[i][v] + err;
    }
    out = (double)sm + (double)c;
  }
  __syncthreads();  // red is reused by the next chunk
  return out;
}

// FP64 evaluation: the pose in FP64, absolute coordinates; the residual,
// Jacobian and Huber forms and the order of their products follow the oracle's
// restatement of CylinderCost / PlaneCost and Ceres' corrector
struct EvalFrame64 {
  double R[3][3], t[3], Jl[3][3];
  int fm, param;
};

__device__ inline void eval_frame64(const LMState& s, int param, EvalFrame64& F) {
  double q[4] = {s.xe[3], s.xe[4], s.xe[5], s.xe[6]};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) F.Jl[i][j] = i == j ? 1.0 : 0.0;
  if (param == 1) {
    const double w[3] = {s.aae[3], s.aae[4], s.aae[5]};
    aa_to_quat(w, q);
    left_jacobian(w, F.Jl);
  }
  quat_to_R(q, F.R);
  const double* tt = param == 1 ? s.aae : s.xe;
  for (int i = 0; i < 3; ++i) F.t[i] = tt[i];
  F.fm = s.free_mask;
  F.param = param;
}

// One chunk as eval_chunk, in FP64; returns the chunk's 28 sums in threads
// 0..27
__device__ inline double eval_chunk64(const float4* __restrict__ pts, int b, int e,
                                      const EvalFrame64& F, const double4* __restrict__ cyld,
                                      const double4* __restrict__ pld, double delta,
                                      ChunkRed64& red) {
  const double d2 = delta * delta;
  double acc[kSums];
#pragma unroll
  for (int i = 0; i < kSums; ++i) acc[i] = 0.0;
  float4 pv[kRpt];
#pragma unroll
  for (int j = 0; j < kRpt; ++j) {
    const int i = b + j * kPBlock + (int)threadIdx.x;
    pv[j] = i < e ? pts[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  const int n_mine = e > b + (int)threadIdx.x ? (e - b - (int)threadIdx.x + kPBlock - 1) / kPBlock
                                               : 0;
  auto gather = [&](int j, double4& A, double4& B) {
    const int id = __float_as_int(pv[j].w);
    if (id >= 0) {
      A = cyld[2 * (size_t)id];
      B = cyld[2 * (size_t)id + 1];
    } else {
      A = pld[-id - 1];
      B = make_double4(0.0, 0.0, 0.0, 0.0);
    }
  };
  double4 na = make_double4(0.0, 0.0, 0.0, 0.0), nb = na;
  if (n_mine > 0) gather(0, na, nb);
#pragma unroll
  for (int j = 0; j < kRpt; ++j) {
    if (j >= n_mine) break;
    const float4 p = pv[j];
    const int id = __float_as_int(p.w);
    const double4 L0 = na, L1 = nb;
    if (j + 1 < n_mine) gather(j + 1, na, nb);
    const double P[3] = {(double)p.x, (double)p.y, (double)p.z};
    double rp[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) rp[a] = F.R[a][0] * P[0] + F.R[a][1] * P[1] + F.R[a][2] * P[2];
    const double lp[3] = {rp[0] + F.t[0], rp[1] + F.t[1], rp[2] + F.t[2]};
    double dir[3], r;
    if (id >= 0) {
      const double v0 = lp[0] - L0.x, v1 = lp[1] - L0.y, v2 = lp[2] - L0.z;
      const double h = v0 * L1.x + v1 * L1.y + v2 * L1.z;
      const double w0 = v0 - h * L1.x, w1 = v1 - h * L1.y, w2 = v2 - h * L1.z;
      const double rho = sqrt(w0 * w0 + w1 * w1 + w2 * w2);
      r = rho - L0.w;
      dir[0] = rho > 0.0 ? w0 / rho : 0.0;
      dir[1] = rho > 0.0 ? w1 / rho : 0.0;
      dir[2] = rho > 0.0 ? w2 / rho : 0.0;
    } else {
      const double sv = L0.x * lp[0] + L0.y * lp[1] + L0.z * lp[2] + L0.w;
      r = fabs(sv);
      const double sg = sv < 0.0 ? -1.0 : 1.0;
      dir[0] = sg * L0.x;
      dir[1] = sg * L0.y;
      dir[2] = sg * L0.z;
    }
    double J[6];
    J[0] = dir[0];
    J[1] = dir[1];
    J[2] = dir[2];
    J[3] = rp[1] * dir[2] - rp[2] * dir[1];
    J[4] = rp[2] * dir[0] - rp[0] * dir[2];
    J[5] = rp[0] * dir[1] - rp[1] * dir[0];
    if (F.param == 1) {
      const double jt[3] = {J[3], J[4], J[5]};
#pragma unroll
      for (int c = 0; c < 3; ++c)
        J[3 + c] = jt[0] * F.Jl[0][c] + jt[1] * F.Jl[1][c] + jt[2] * F.Jl[2][c];
    }
#pragma unroll
    for (int a = 0; a < 6; ++a)
      if (!((F.fm >> a) & 1)) J[a] = 0.0;
    const double s2 = r * r;
    double wgt, rho;
    if (s2 <= d2) {
      wgt = 1.0;
      rho = s2;
    } else {
      const double ar = sqrt(s2);
      wgt = delta / ar;
      rho = 2.0 * delta * ar - d2;
    }
    acc[27] += 0.5 * rho;
    int tt = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      const double wa = wgt * J[a];
#pragma unroll
      for (int bb = a; bb < 6; ++bb) {
        acc[tt] += wa * J[bb];
        ++tt;
      }
      acc[21 + a] += wa * r;
    }
  }
  // the reduction of eval_chunk with plain FP64 sums
#pragma unroll
  for (int i = 0; i < kSums; ++i) red.x[i][threadIdx.x] = acc[i];
  __syncthreads();
  if (threadIdx.x < kSums * kParts) {
    const int i = threadIdx.x % kSums, part = threadIdx.x / kSums;
    const double* row = &red.x[i][part * (kPBlock / kParts)];
    double sm = row[0];
    for (int j = 1; j < kPBlock / kParts; ++j) sm += row[j];
    red.ps[i][part] = sm;
  }
  __syncthreads();
  double out = 0.0;
  if (threadIdx.x < kSums) {
    const int i = threadIdx.x;
    out = red.ps[i][0];
    for (int v = 1; v < kParts; ++v) out += red.ps[i][v];
  }
  __syncthreads();  // red is reused by the next chunk
  return out;
}

// the evaluation of one precision: frame, chunk, LDS and table types
template <bool F64>
struct PoseEval {
  using Frame = EvalFrame;
  using Red = ChunkRed;
  __device__ static void frame(const LMState& s, int param, Frame& F) { eval_frame(s, param, F); }
  __device__ static double chunk(const float4* pts, int b, int e, const Frame& F,
                                 const void* cyl_tab, const void* pl_tab, double delta, Red& red) {
    const float4* ct = static_cast<const float4*>(cyl_tab);
    const float4* pt = static_cast<const float4*>(pl_tab);
    return F.sub ? eval_chunk<true>(pts, b, e, F, ct, pt, (float)delta, red)
                 : eval_chunk<false>(pts, b, e, F, ct, pt, (float)delta, red);
  }
};

template <>
struct PoseEval<true> {
  using Frame = EvalFrame64;
  using Red = ChunkRed64;
  __device__ static void frame(const LMState& s, int param, Frame& F) {
    eval_frame64(s, param, F);
  }
  __device__ static double chunk(const float4* pts, int b, int e, const Frame& F,
                                 const void* cyl_tab, const void* pl_tab, double delta, Red& red) {
    return eval_chunk64(pts, b, e, F, static_cast<const double4*>(cyl_tab),
                        static_cast<const double4*>(pl_tab), delta, red);
  }
};

// work item = {problem, first residual, end residual, unused}
template <bool F64>
__global__ __launch_bounds__(kPBlock) void pose_eval_kernel(
    const float4* __restrict__ pts, const int4* __restrict__ work, const LMState* __restrict__ st,
    const void* __restrict__ cyl_tab, const void* __restrict__ pl_tab, double delta, int param,
    double* __restrict__ partials) {
  using E = PoseEval<F64>;
  __shared__ typename E::Red red;
  const int4 w = work[blockIdx.x];
  const int k = w.x;
  if (st[k].done) return;
  typename E::Frame F;
  E::frame(st[k], param, F);
  const double p = E::chunk(pts, w.y, w.z, F, cyl_tab, pl_tab, delta, red);
  if (threadIdx.x < kSums) partials[(size_t)blockIdx.x * kSums + threadIdx.x] = p;
}

// --------------------------------------------------------------------------
// solve: LM trust-region logic and the FP64 6x6 step on the summed evaluation
// tot[28] of the pose S.xe (one thread)
// --------------------------------------------------------------------------
__device__ inline void lm_step(LMState& S, const double* tot, const PoseCfgDev& cfg) {
  // evaluation result of S.xe
  double A[6][6], g[6];
  {
    int tt = 0;
    for (int a = 0; a < 6; ++a)
      for (int b = a; b < 6; ++b) {
        A[a][b] = tot[tt];
        A[b][a] = tot[tt];
        ++tt;
      }
    for (int a = 0; a < 6; ++a) g[a] = tot[21 + a];
  }
  const double Fe = tot[27];
  auto store_eval = [&]() {
    int tt = 0;
    for (int a = 0; a < 6; ++a)
      for (int b = a; b < 6; ++b) S.A[tt++] = A[a][b];
    for (int a = 0; a < 6; ++a) S.g[a] = g[a];
    S.F = Fe;
  };
  bool done = false;
  if (S.phase == 0) {
    store_eval();
    S.F0 = Fe;
    double gmax = 0.0;
    for (int a = 0; a < 6; ++a) gmax = fmax(gmax, fabs(g[a]));
    if (gmax <= cfg.gtol) {
      S.term = 0;
      if (cfg.stop_conv) done = true;
    }
  } else {
    const double change = S.F - Fe;
    const double q = S.model > 0.0 ? change / S.model : -1.0;
    if (isfinite(Fe) && q > cfg.min_rel) {
      for (int i = 0; i < 7; ++i) S.x[i] = S.xe[i];
      for (int i = 0; i < 6; ++i) S.aa[i] = S.aae[i];
      const double Fprev = S.F;
      store_eval();
      const double c3 = 2.0 * q - 1.0;
      S.radius = S.radius / fmax(1.0 / 3.0, 1.0 - c3 * c3 * c3);
      S.radius = fmin(S.radius, 1e16);
      S.factor = 2.0;
      double gmax = 0.0;
      for (int a = 0; a < 6; ++a) gmax = fmax(gmax, fabs(S.g[a]));
      if (fabs(change) <= cfg.ftol * Fprev || gmax <= cfg.gtol ||
          S.dn <= cfg.ptol * (S.xn + cfg.ptol)) {
        S.term = 0;
        if (cfg.stop_conv) done = true;
      }
    } else {
      S.radius = S.radius / S.factor;
      S.factor = S.factor * 2.0;
      if (S.dn <= cfg.ptol * (S.xn + cfg.ptol)) {
        S.term = 0;
        if (cfg.stop_conv) done = true;
      }
    }
  }
  // next candidate
  bool have_candidate = false;
  while (!done && S.iter < cfg.max_iters) {
    S.iter += 1;
    double Acur[6][6], Aug[6][6], rhs[6], dx[6];
    {
      int tt = 0;
      for (int a = 0; a < 6; ++a)
        for (int b = a; b < 6; ++b) {
          Acur[a][b] = S.A[tt];
          Acur[b][a] = S.A[tt];
          ++tt;
        }
    }
    for (int a = 0; a < 6; ++a) {
      for (int b = 0; b < 6; ++b) Aug[a][b] = Acur[a][b];
      double dg = Acur[a][a];
      dg = dg < 1e-6 ? 1e-6 : (dg > 1e32 ? 1e32 : dg);
      Aug[a][a] += dg / S.radius;
      rhs[a] = -S.g[a];
    }
    if (!chol6(Aug, rhs, dx)) {
      S.radius = S.radius / S.factor;
      S.factor = S.factor * 2.0;
      continue;
    }
    double gdx = 0.0, dAd = 0.0;
    for (int a = 0; a < 6; ++a) {
      gdx += S.g[a] * dx[a];
      double tt = 0.0;
      for (int b = 0; b < 6; ++b) tt += Acur[a][b] * dx[b];
      dAd += dx[a] * tt;
    }
    S.model = -(gdx + 0.5 * dAd);
    double xn = 0.0, dn = 0.0;
    if (cfg.param == 1) {
      for (int a = 0; a < 6; ++a) xn += S.aa[a] * S.aa[a];
    } else {
      for (int a = 0; a < 7; ++a) xn += S.x[a] * S.x[a];
    }
    for (int a = 0; a < 6; ++a) dn += dx[a] * dx[a];
    S.xn = sqrt(xn);
    S.dn = sqrt(dn);
    if (cfg.param == 1) {
      for (int a = 0; a < 6; ++a) S.aae[a] = S.aa[a] + dx[a];
      double qn[4];
      aa_to_quat(S.aae + 3, qn);
      for (int a = 0; a < 3; ++a) S.xe[a] = S.aae[a];
      for (int a = 0; a < 4; ++a) S.xe[3 + a] = qn[a];
    } else {
      for (int a = 0; a < 3; ++a) S.xe[a] = S.x[a] + dx[a];
      double qn[4];
      quat_left_update(S.x + 3, dx + 3, qn);
      for (int a = 0; a < 4; ++a) S.xe[3 + a] = qn[a];
    }
    have_candidate = true;
    break;
  }
  S.phase = 1;
  if (!have_candidate) S.done = 1;
}

__device__ inline void write_outputs(const LMState& S, int k, const PoseCfgDev& cfg,
                                     double* __restrict__ pose_out, int32_t* __restrict__ status,
                                     double* __restrict__ cost, double* __restrict__ params_out) {
  for (int i = 0; i < 7; ++i) pose_out[7 * (size_t)k + i] = S.x[i];
  if (params_out) {
    if (cfg.param == 1) {
      for (int i = 0; i < 6; ++i) params_out[6 * (size_t)k + i] = S.aa[i];
    } else {
      double w[3];
      quat_to_aa(S.x + 3, w);
      for (int i = 0; i < 3; ++i) {
        params_out[6 * (size_t)k + i] = S.x[i];
        params_out[6 * (size_t)k + 3 + i] = w[i];
      }
    }
  }
  if (status) {
    status[2 * k] = S.term;
    status[2 * k + 1] = S.iter;
  }
  if (cost) {
    cost[2 * k] = S.F0;
    cost[2 * k + 1] = S.F;
  }
}

// FP64 chunk sums in chunk order, one wave per problem
__global__ __launch_bounds__(64) void pose_solve_kernel(
    int n_problems, const int32_t* __restrict__ first_work, const double* __restrict__ partials,
    PoseCfgDev cfg, LMState* __restrict__ st, double* __restrict__ pose_out,
    int32_t* __restrict__ status, double* __restrict__ cost, double* __restrict__ params_out) {
  __shared__ double tot[kSums];
  const int k = blockIdx.x;
  if (k >= n_problems) return;
  if (st[k].done) return;
  if (threadIdx.x < kSums) {
    // the chunks' sums added in chunk order; loaded eight at a time so that
    // the loads overlap
    double s = 0.0;
    const int w0 = first_work[k], w1 = first_work[k + 1];
    for (int wk = w0; wk < w1; wk += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        v[u] = wk + u < w1 ? partials[(size_t)(wk + u) * kSums + threadIdx.x] : 0.0;
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (wk + u < w1) s += v[u];
    }
    tot[threadIdx.x] = s;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  LMState S = st[k];
  lm_step(S, tot, cfg);
  st[k] = S;
  write_outputs(S, k, cfg, pose_out, status, cost, params_out);
}

// --------------------------------------------------------------------------
// fused: every iteration of one problem in one workgroup
// --------------------------------------------------------------------------
template <bool F64>
__global__ __launch_bounds__(kPBlock) void pose_fused_kernel(
    int n_problems, const int32_t* __restrict__ d_n_problems, const float4* __restrict__ pts,
    const int4* __restrict__ work, const int32_t* __restrict__ first_work,
    const void* __restrict__ cyl_tab, const void* __restrict__ pl_tab, PoseCfgDev cfg,
    LMState* __restrict__ st, double* __restrict__ pose_out, int32_t* __restrict__ status,
    double* __restrict__ cost, double* __restrict__ params_out) {
  using E = PoseEval<F64>;
  __shared__ typename E::Red red;
  __shared__ double tot[kSums];
  __shared__ LMState sS;
  const int k = blockIdx.x;
  if (k >= n_problems || (d_n_problems && k >= *d_n_problems)) return;
  if (threadIdx.x == 0) sS = st[k];
  __syncthreads();
  const int w0 = first_work[k], w1 = first_work[k + 1];
  bool ran = false;
  for (int it = 0; it <= cfg.max_iters; ++it) {
    if (sS.done) break;
    typename E::Frame F;
    E::frame(sS, cfg.param, F);
    double acc = 0.0;
    for (int wk = w0; wk < w1; ++wk) {
      const int4 w = work[wk];
      const double p = E::chunk(pts, w.y, w.z, F, cyl_tab, pl_tab, cfg.delta, red);
      if (threadIdx.x < kSums) acc += p;
    }
    if (threadIdx.x < kSums) tot[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      LMState S = sS;
      lm_step(S, tot, cfg);
      sS = S;
    }
    ran = true;
    __syncthreads();
  }
  if (threadIdx.x == 0 && ran) {
    const LMState S = sS;
    st[k] = S;
    write_outputs(S, k, cfg, pose_out, status, cost, params_out);
  }
}

}  // namespace

struct sloam_pose_plan {
  sloam_ctx* ctx = nullptr;
  sloam_pose_cfg cfg{};
  int32_t mode = SLOAM_POSE_MODE_AUTO;
  int32_t n_problems = 0;
  int64_t n_res = 0;
  int32_t n_cyl = 0, n_pl = 0;
  int n_work = 0;
  int max_chunks = 0;  // most chunks of one problem
  std::vector<int4> h_work;
  std::vector<int32_t> h_first;
  // device buffers, grow-only (capacities in elements)
  int4* d_work = nullptr;
  int32_t* d_first = nullptr;
  LMState* d_state = nullptr;
  double* d_partials = nullptr;
  void* d_cyl_tab = nullptr;  // landmark tables (pose_prep_kernel), bytes
  void* d_pl_tab = nullptr;
  size_t cap_work = 0, cap_first = 0, cap_state = 0, cap_partials = 0, cap_cyl = 0, cap_pl = 0;
};

namespace {

// a problem set whose largest problem has at most this many chunks is solved
// by pose_fused in AUTO mode
constexpr int kFusedAutoChunks = 8;

int grow(sloam_ctx* ctx, void** p, size_t* cap, size_t want, size_t elem) {
  if (want <= *cap) return SLOAM_OK;
  sloam_dev_free(*p);
  *p = nullptr;
  *cap = 0;
  const int st = sloam_dev_alloc(ctx, p, want * elem);
  if (st) return st;
  *cap = want;
  return SLOAM_OK;
}

// (Re)targets the plan at a problem set: work list, chunk table, buffers.
int plan_build(sloam_pose_plan* p, const int32_t* h_res_offsets, int32_t n_problems,
               int32_t n_cylinders, int32_t n_planes) {
  if (!h_res_offsets || n_problems < 0 || n_cylinders < 0 || n_planes < 0)
    return SLOAM_ERR_INVALID_ARG;
  if (h_res_offsets[0] != 0) return SLOAM_ERR_INVALID_ARG;
  for (int k = 0; k < n_problems; ++k)
    if (h_res_offsets[k + 1] < h_res_offsets[k]) return SLOAM_ERR_INVALID_ARG;
  sloam_ctx* ctx = p->ctx;
  // the stream may still read the previous work list and buffers
  SLOAM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  p->n_problems = n_problems;
  p->n_res = h_res_offsets[n_problems];
  p->n_cyl = n_cylinders;
  p->n_pl = n_planes;
  const int chunk = kPBlock * kRpt;
  p->h_work.clear();
  p->h_first.assign(n_problems + 1, 0);
  p->max_chunks = 0;
  for (int k = 0; k < n_problems; ++k) {
    p->h_first[k] = (int32_t)p->h_work.size();
    const int b = h_res_offsets[k], e = h_res_offsets[k + 1];
    if (e == b) p->h_work.push_back(make_int4(k, b, e, 0));  // keeps one (empty) chunk
    for (int s = b; s < e; s += chunk)
      p->h_work.push_back(make_int4(k, s, std::min(e, s + chunk), 0));
    p->max_chunks = std::max(p->max_chunks, (int)p->h_work.size() - p->h_first[k]);
  }
  p->h_first[n_problems] = (int32_t)p->h_work.size();
  p->n_work = (int)p->h_work.size();
  int st = grow(ctx, (void**)&p->d_work, &p->cap_work, p->h_work.size(), sizeof(int4));
  if (!st) st = grow(ctx, (void**)&p->d_first, &p->cap_first, p->h_first.size(), sizeof(int32_t));
  if (!st) st = grow(ctx, (void**)&p->d_state, &p->cap_state, (size_t)n_problems, sizeof(LMState));
  if (!st)
    st = grow(ctx, (void**)&p->d_partials, &p->cap_partials, (size_t)kSums * p->h_work.size(),
              sizeof(double));
  // landmark table entries: 2 float4 / 2 double4 per cylinder, 1 per plane
  const size_t tab = p->cfg.precision == 1 ? sizeof(double4) : sizeof(float4);
  if (!st) st = grow(ctx, &p->d_cyl_tab, &p->cap_cyl, 2 * tab * (size_t)n_cylinders, 1);
  if (!st) st = grow(ctx, &p->d_pl_tab, &p->cap_pl, tab * (size_t)n_planes, 1);
  if (st) return st;
  // the host vectors stay alive in the plan until the next build, which
  // synchronises the stream first
  if (!p->h_work.empty())
    SLOAM_HIP(ctx, hipMemcpyAsync(p->d_work, p->h_work.data(), sizeof(int4) * p->h_work.size(),
                                  hipMemcpyHostToDevice, ctx->stream));
  SLOAM_HIP(ctx, hipMemcpyAsync(p->d_first, p->h_first.data(),
                                sizeof(int32_t) * p->h_first.size(), hipMemcpyHostToDevice,
                                ctx->stream));
  return SLOAM_OK;
}

bool valid_cfg(const sloam_pose_cfg* cfg) {
  return cfg->max_iters >= 0 && cfg->huber_delta > 0.0 && cfg->initial_radius > 0.0 &&
         cfg->parametrization >= 0 && cfg->parametrization <= 1 && cfg->precision >= 0 &&
         cfg->precision <= 1;
}

}  // namespace

static PoseCfgDev pose_cfg_dev(const sloam_pose_cfg& c) {
  PoseCfgDev dc;
  dc.max_iters = c.max_iters;
  dc.stop_conv = c.stop_on_convergence;
  dc.param = c.parametrization;
  dc.delta = c.huber_delta;
  dc.ftol = c.function_tolerance;
  dc.gtol = c.gradient_tolerance;
  dc.ptol = c.parameter_tolerance;
  dc.radius0 = c.initial_radius;
  dc.min_rel = c.min_relative_decrease;
  return dc;
}

size_t sloam_pose_state_bytes() { return sizeof(LMState); }

namespace {

// pose_prep + pose_fused of one precision
template <bool F64>
void launch_prep(sloam_ctx* ctx, int grid, int n_problems, const int32_t* d_n_problems,
                 const double* d_pose_in, const int32_t* d_free_mask, const double* d_cyl,
                 const int32_t* d_cyl_off, const double* d_planes, const int32_t* d_pl_off,
                 const PoseCfgDev& dc, LMState* st, void* cyl_tab, void* pl_tab) {
  hipLaunchKernelGGL(pose_prep_kernel<F64>, dim3(grid), dim3(kPBlock), 0, ctx->stream,
                     n_problems, d_n_problems, d_pose_in, d_free_mask, d_cyl, d_cyl_off, d_planes,
                     d_pl_off, dc, st, cyl_tab, pl_tab);
}

template <bool F64>
void launch_fused(sloam_ctx* ctx, int grid, int n_problems, const int32_t* d_n_problems,
                  const float* d_points, const int4* d_work, const int32_t* d_first,
                  const void* cyl_tab, const void* pl_tab, const PoseCfgDev& dc, LMState* st,
                  double* d_pose_out, int32_t* d_status, double* d_cost, double* d_params_out) {
  hipLaunchKernelGGL(pose_fused_kernel<F64>, dim3(grid), dim3(kPBlock), 0, ctx->stream,
                     n_problems, d_n_problems, reinterpret_cast<const float4*>(d_points), d_work,
                     d_first, cyl_tab, pl_tab, dc, st, d_pose_out, d_status, d_cost,
                     d_params_out);
}

}  // namespace

int sloam_pose_fused_device(sloam_ctx* ctx, int max_problems, const int32_t* d_n_problems,
                            const float* d_points, const int32_t* d_work, const int32_t* d_first,
                            const double* d_cyl, const int32_t* d_cyl_off, const double* d_planes,
                            const int32_t* d_pl_off, const double* d_pose_in,
                            const int32_t* d_free_mask, const sloam_pose_cfg* cfg, void* d_state,
                            void* d_cyl_tab, void* d_pl_tab, double* d_pose_out,
                            double* d_params_out, int32_t* d_status, double* d_cost) {
  if (max_problems <= 0) return SLOAM_OK;
  if (!valid_cfg(cfg)) return SLOAM_ERR_INVALID_ARG;
  const PoseCfgDev dc = pose_cfg_dev(*cfg);
  LMState* st = static_cast<LMState*>(d_state);
  const int4* work = reinterpret_cast<const int4*>(d_work);
  if (cfg->precision == 1) {
    launch_prep<true>(ctx, max_problems, max_problems, d_n_problems, d_pose_in, d_free_mask, d_cyl,
                      d_cyl_off, d_planes, d_pl_off, dc, st, d_cyl_tab, d_pl_tab);
    SLOAM_HIP(ctx, hipGetLastError());
    launch_fused<true>(ctx, max_problems, max_problems, d_n_problems, d_points, work, d_first,
                       d_cyl_tab, d_pl_tab, dc, st, d_pose_out, d_status, d_cost, d_params_out);
  } else {
    launch_prep<false>(ctx, max_problems, max_problems, d_n_problems, d_pose_in, d_free_mask,
                       d_cyl, d_cyl_off, d_planes, d_pl_off, dc, st, d_cyl_tab, d_pl_tab);
    SLOAM_HIP(ctx, hipGetLastError());
    launch_fused<false>(ctx, max_problems, max_problems, d_n_problems, d_points, work, d_first,
                        d_cyl_tab, d_pl_tab, dc, st, d_pose_out, d_status, d_cost, d_params_out);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}

extern "C" {

int sloam_pose_plan_create(sloam_ctx* ctx, const int32_t* h_res_offsets, int32_t n_problems,
                           int32_t n_cylinders, int32_t n_planes, const sloam_pose_cfg* cfg,
                           sloam_pose_plan** out) {
  if (!ctx || !h_res_offsets || !cfg || !out || n_problems < 0 || n_cylinders < 0 ||
      n_planes < 0)
    return SLOAM_ERR_INVALID_ARG;
  *out = nullptr;
  if (!valid_cfg(cfg)) return SLOAM_ERR_INVALID_ARG;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  sloam_pose_plan* p = new (std::nothrow) sloam_pose_plan();
  if (!p) return SLOAM_ERR_ALLOC;
  p->ctx = ctx;
  p->cfg = *cfg;
  const int st = plan_build(p, h_res_offsets, n_problems, n_cylinders, n_planes);
  if (st) {
    sloam_pose_plan_destroy(p);
    return st;
  }
  *out = p;
  return SLOAM_OK;
}

int sloam_pose_plan_reset(sloam_pose_plan* p, const int32_t* h_res_offsets, int32_t n_problems,
                          int32_t n_cylinders, int32_t n_planes, const sloam_pose_cfg* cfg) {
  if (!p) return SLOAM_ERR_INVALID_ARG;
  if (cfg) {
    if (!valid_cfg(cfg)) return SLOAM_ERR_INVALID_ARG;
    p->cfg = *cfg;
  }
  SLOAM_HIP(p->ctx, hipSetDevice(p->ctx->device));
  return plan_build(p, h_res_offsets, n_problems, n_cylinders, n_planes);
}

int sloam_pose_plan_set_mode(sloam_pose_plan* p, int32_t mode) {
  if (!p || mode < SLOAM_POSE_MODE_AUTO || mode > SLOAM_POSE_MODE_FUSED)
    return SLOAM_ERR_INVALID_ARG;
  p->mode = mode;
  return SLOAM_OK;
}

int sloam_pose_plan_uses_fused(const sloam_pose_plan* p) {
  if (!p) return SLOAM_ERR_INVALID_ARG;
  return p->mode == SLOAM_POSE_MODE_FUSED ||
         (p->mode == SLOAM_POSE_MODE_AUTO && p->max_chunks <= kFusedAutoChunks);
}

void sloam_pose_plan_destroy(sloam_pose_plan* p) {
  if (!p) return;
  if (p->ctx) (void)hipStreamSynchronize(p->ctx->stream);
  sloam_dev_free(p->d_work);
  sloam_dev_free(p->d_first);
  sloam_dev_free(p->d_state);
  sloam_dev_free(p->d_partials);
  sloam_dev_free(p->d_cyl_tab);
  sloam_dev_free(p->d_pl_tab);
  delete p;
}

int sloam_pose_plan_exec(sloam_pose_plan* p, const float* d_points, const double* d_cyl,
                         const int32_t* d_cyl_offsets, const double* d_planes,
                         const int32_t* d_plane_offsets, const double* d_pose_in,
                         const int32_t* d_free_mask, double* d_pose_out, double* d_params_out,
                         int32_t* d_status, double* d_cost) {
  if (!p || !d_cyl_offsets || !d_plane_offsets || !d_pose_in || !d_pose_out)
    return SLOAM_ERR_INVALID_ARG;
  if ((p->n_res > 0 && !d_points) || (p->n_cyl > 0 && !d_cyl) || (p->n_pl > 0 && !d_planes))
    return SLOAM_ERR_INVALID_ARG;
  sloam_ctx* ctx = p->ctx;
  if (p->n_problems == 0) return SLOAM_OK;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  const sloam_pose_cfg& c = p->cfg;
  const PoseCfgDev dc = pose_cfg_dev(c);
  const bool f64 = c.precision == 1;
  {
    SloamKernelScope ks(ctx, "pose_prep");
    if (f64)
      launch_prep<true>(ctx, p->n_problems, p->n_problems, nullptr, d_pose_in, d_free_mask, d_cyl,
                        d_cyl_offsets, d_planes, d_plane_offsets, dc, p->d_state, p->d_cyl_tab,
                        p->d_pl_tab);
    else
      launch_prep<false>(ctx, p->n_problems, p->n_problems, nullptr, d_pose_in, d_free_mask,
                         d_cyl, d_cyl_offsets, d_planes, d_plane_offsets, dc, p->d_state,
                         p->d_cyl_tab, p->d_pl_tab);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  if (sloam_pose_plan_uses_fused(p)) {
    SloamKernelScope ks(ctx, "pose_fused");
    if (f64)
      launch_fused<true>(ctx, p->n_problems, p->n_problems, nullptr, d_points, p->d_work,
                         p->d_first, p->d_cyl_tab, p->d_pl_tab, dc, p->d_state, d_pose_out,
                         d_status, d_cost, d_params_out);
    else
      launch_fused<false>(ctx, p->n_problems, p->n_problems, nullptr, d_points, p->d_work,
                          p->d_first, p->d_cyl_tab, p->d_pl_tab, dc, p->d_state, d_pose_out,
                          d_status, d_cost, d_params_out);
    SLOAM_HIP(ctx, hipGetLastError());
    return SLOAM_OK;
  }
  for (int it = 0; it <= c.max_iters; ++it) {
    {
      SloamKernelScope ks(ctx, "pose_eval");
      auto kern = f64 ? pose_eval_kernel<true> : pose_eval_kernel<false>;
      hipLaunchKernelGGL(kern, dim3(p->n_work), dim3(kPBlock), 0, ctx->stream,
                         reinterpret_cast<const float4*>(d_points), p->d_work, p->d_state,
                         static_cast<const void*>(p->d_cyl_tab),
                         static_cast<const void*>(p->d_pl_tab), c.huber_delta, c.parametrization,
                         p->d_partials);
    }
    SLOAM_HIP(ctx, hipGetLastError());
    {
      SloamKernelScope ks(ctx, "pose_solve");
      hipLaunchKernelGGL(pose_solve_kernel, dim3(p->n_problems), dim3(64), 0, ctx->stream,
                         p->n_problems, p->d_first, p->d_partials, dc, p->d_state, d_pose_out,
                         d_status, d_cost, d_params_out);
    }
    SLOAM_HIP(ctx, hipGetLastError());
  }
  return SLOAM_OK;
}

}  // extern "C"
