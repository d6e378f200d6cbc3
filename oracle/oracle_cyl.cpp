// TEST INFRASTRUCTURE ONLY -- CPU oracle of the batched point-level cylinder
// RANSAC ("plain method", fixed hypothesis count).
//
// Inlier predicate: the reference's point-to-cylinder distance,
//   sloam/src/objects/cylinder.cpp:196-203  Cylinder::distance(const PointT&)
//   d = ||p - (root + ((p-root).ray / ray.ray) ray)|| - radius,  inlier <=> |d| < thr
// evaluated for a unit ray in the squared, axis-orthogonal form
//   (r - thr)^2 < (u - cu)^2 + (v - cv)^2 < (r + thr)^2.
// The reference itself fits cylinders by a PCL line RANSAC over trellis vertex
// medians (cylinder.cpp:116-124); point-level hypothesis scoring has no
// counterpart there, so the hypothesis generator below is this project's
// definition: Philox4x32-10 counter-based samples of 3 distinct points, their
// circumcircle in the plane orthogonal to the PCA seed axis.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle_common.hpp"

namespace oracle {
namespace {

struct Frame {
  float m[3];
  float e1[3];
  float e2[3];
  double axis[3];
};

struct Hyp {
  float A, B, L, U;
  float cu, cv, r;
  bool valid;
};

// Local coordinates of p in the plane orthogonal to the seed axis (FP32).
inline void to_local(const Frame& f, const float* p, float& u, float& v, float& s) {
  const float dx = p[0] - f.m[0];
  const float dy = p[1] - f.m[1];
  const float dz = p[2] - f.m[2];
  u = (dx * f.e1[0] + dy * f.e1[1]) + dz * f.e1[2];
  v = (dx * f.e2[0] + dy * f.e2[1]) + dz * f.e2[2];
  s = u * u + v * v;
}

// Seed: bounding-box centre, covariance from integer-exact sums of
// q = rint((p - m) * 2^16) (|q| clamped below 2^23), FP64 Jacobi, major axis.
Frame seed_frame(const float* P, int n) {
  Frame f;
  float lo[3] = {INFINITY, INFINITY, INFINITY};
  float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], P[3 * i + k]);
      hi[k] = std::max(hi[k], P[3 * i + k]);
    }
  for (int k = 0; k < 3; ++k) f.m[k] = n > 0 ? 0.5f * (lo[k] + hi[k]) : 0.0f;
  int64_t S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < n; ++i) {
    int64_t q[3];
    for (int k = 0; k < 3; ++k) {
      double d = std::rint(((double)P[3 * i + k] - (double)f.m[k]) * 65536.0);
      d = std::min(8388607.0, std::max(-8388607.0, d));
      q[k] = (int64_t)d;
    }
    S[0] += q[0];
    S[1] += q[1];
    S[2] += q[2];
    S[3] += q[0] * q[0];
    S[4] += q[0] * q[1];
    S[5] += q[0] * q[2];
    S[6] += q[1] * q[1];
    S[7] += q[1] * q[2];
    S[8] += q[2] * q[2];
  }
  double axis[3] = {0.0, 0.0, 1.0};
  if (n >= 2) {
    const double dn = (double)n;
    const double m0 = (double)S[0], m1 = (double)S[1], m2 = (double)S[2];
    const double cov[6] = {(double)S[3] - m0 * m0 / dn, (double)S[4] - m0 * m1 / dn,
                           (double)S[5] - m0 * m2 / dn, (double)S[6] - m1 * m1 / dn,
                           (double)S[7] - m1 * m2 / dn, (double)S[8] - m2 * m2 / dn};
    double w[3], V[3][3];
    eig3(cov, w, V);
    eigvec(V, argmax3(w), axis);
  }
  double e1[3], e2[3];
  basis(axis, e1, e2);
  for (int k = 0; k < 3; ++k) {
    f.axis[k] = axis[k];
    f.e1[k] = (float)e1[k];
    f.e2[k] = (float)e2[k];
  }
  return f;
}

// Predicate bounds of the circle (cu, cv, r) of the local plane, FP32.
Hyp circle_params(float cu, float cv, float r, float thr, float max_radius) {
  Hyp h;
  h.cu = cu;
  h.cv = cv;
  h.r = r;
  h.valid = r <= max_radius;
  if (h.valid) {
    h.A = -2.0f * h.cu;
    h.B = -2.0f * h.cv;
    const float K = h.cu * h.cu + h.cv * h.cv;
    const float up = r + thr;
    h.U = up * up - K;
    const float dn = r - thr;
    h.L = dn > 0.0f ? dn * dn - K : -INFINITY;
  } else {
    h.A = h.B = 0.0f;
    h.L = INFINITY;
    h.U = -INFINITY;
  }
  return h;
}

// Circumcircle of (a, b, c) in the local plane, FP32, and predicate bounds.
Hyp circle_hyp(const float a[2], const float b[2], const float c[2], float thr, float max_radius) {
  const float bx = b[0] - a[0], by = b[1] - a[1];
  const float cx = c[0] - a[0], cy = c[1] - a[1];
  const float den = 2.0f * (bx * cy - by * cx);
  const float bb = bx * bx + by * by;
  const float cc = cx * cx + cy * cy;
  const float ox = (cy * bb - by * cc) / den;
  const float oy = (bx * cc - cx * bb) / den;
  const float r = std::sqrt(ox * ox + oy * oy);
  return circle_params(a[0] + ox, a[1] + oy, r, thr, max_radius);
}

// Hypothesis h: the proposed circle for h == 0 when the cluster has a valid
// proposal (prop_c: axis[3], centre[3], radius, valid; the centre is taken
// into the seed frame), else the circumcircle of a Philox sample.
Hyp hypothesis(const float* P, int n, int cluster, int h, const Frame& f, float thr,
               float max_radius, uint32_t k0, uint32_t k1, uint32_t tag, int32_t idx[3],
               const float* prop_c = nullptr) {
  if (prop_c && h == 0) {
    idx[0] = idx[1] = idx[2] = -1;
    float cu, cv, s;
    to_local(f, prop_c + 3, cu, cv, s);
    return circle_params(cu, cv, prop_c[6], thr, max_radius);
  }
  const U4 ctr = {{(uint32_t)h, (uint32_t)cluster, tag, 0u}};
  const U4 x = philox4x32_10(ctr, k0, k1);
  draw3(x.v[0], x.v[1], x.v[2], (uint32_t)n, idx);
  float pts[3][2];
  for (int k = 0; k < 3; ++k) {
    float s;
    to_local(f, P + 3 * idx[k], pts[k][0], pts[k][1], s);
  }
  return circle_hyp(pts[0], pts[1], pts[2], thr, max_radius);
}

inline bool inlier(float u, float v, float s, const Hyp& h) {
  const float t = std::fmaf(u, h.A, std::fmaf(v, h.B, s));
  return t > h.L && t < h.U;
}

// LO-RANSAC refit of the best hypothesis h, all in FP64 (refit_fp64; the
// error figure of the FP32 refit below is taken against it).
// `rounds` rounds of `iters` Gauss-Newton steps of the 5-parameter geometric
// cylinder fit
//   d_i = ||(p_i - c) - ((p_i - c).a) a|| - r
// (axis a perturbed in its orthogonal plane, axis point c moved in that plane,
// radius r).  The inliers of round 0 are the FP32 inliers of h; the inliers of
// round j > 0 are |d_i| < thr under the model at the start of round j.  A round
// stops early when fewer than 5 points are inliers or the normal matrix is not
// positive definite.  Output: root (axis point at the height of the final
// inlier centroid), ray (unit, largest |component| positive), radius and the
// final inlier count (|d_i| < thr under the final model).
struct CylModel {
  double a[3], c[3], r;
};

inline double cyl_dist(const CylModel& M, const double p[3], double w[3], double* h) {
  const double q[3] = {p[0] - M.c[0], p[1] - M.c[1], p[2] - M.c[2]};
  const double hh = q[0] * M.a[0] + q[1] * M.a[1] + q[2] * M.a[2];
  w[0] = q[0] - hh * M.a[0];
  w[1] = q[1] - hh * M.a[1];
  w[2] = q[2] - hh * M.a[2];
  *h = hh;
  return std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) - M.r;
}

void refit_fp64(const float* P, int n, const Frame& f, const Hyp& h, int rounds, int iters,
                float thr, double out[8]) {
  auto invalid = [&](double cnt) {
    for (int i = 0; i < 6; ++i) out[i] = 0.0;
    out[5] = 1.0;
    out[6] = -1.0;
    out[7] = cnt;
  };
  std::vector<double> X(3 * (size_t)n);
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) X[3 * i + c] = (double)P[3 * i + c] - (double)f.m[c];
  std::vector<char> in((size_t)n, 0);
  int k0 = 0;
  for (int i = 0; i < n; ++i) {
    float u, v, s;
    to_local(f, P + 3 * i, u, v, s);
    in[i] = inlier(u, v, s, h) ? 1 : 0;
    k0 += in[i];
  }
  if (k0 < 3 || !h.valid) {
    invalid((double)k0);
    return;
  }
  CylModel M;
  for (int c = 0; c < 3; ++c) {
    M.a[c] = f.axis[c];
    M.c[c] = (double)h.cu * (double)f.e1[c] + (double)h.cv * (double)f.e2[c];
  }
  M.r = (double)h.r;
  const double dthr = (double)thr;
  bool stop = false;
  for (int round = 0; round < rounds && !stop; ++round) {
    if (round > 0) {
      for (int i = 0; i < n; ++i) {
        double w[3], hh;
        const double d = cyl_dist(M, &X[3 * i], w, &hh);
        in[i] = (d < dthr && d > -dthr) ? 1 : 0;
      }
    }
    for (int it = 0; it < iters; ++it) {
      double e1[3], e2[3];
      basis(M.a, e1, e2);
      double A[5][5] = {};
      double g[5] = {};
      int cnt = 0;
      for (int i = 0; i < n; ++i) {
        if (!in[i]) continue;
        ++cnt;
        double w[3], hh;
        const double d = cyl_dist(M, &X[3 * i], w, &hh);
        const double rho = d + M.r;
        if (!(rho > 0.0)) continue;
        const double inv = 1.0 / rho;
        const double n1 = (w[0] * e1[0] + w[1] * e1[1] + w[2] * e1[2]) * inv;
        const double n2 = (w[0] * e2[0] + w[1] * e2[1] + w[2] * e2[2]) * inv;
        const double J[5] = {-hh * n1, -hh * n2, -n1, -n2, -1.0};
        for (int a = 0; a < 5; ++a) {
          for (int b = a; b < 5; ++b) A[a][b] += J[a] * J[b];
          g[a] += J[a] * d;
        }
      }
      if (cnt < 5) {
        stop = true;
        break;
      }
      for (int a = 0; a < 5; ++a)
        for (int b = 0; b < a; ++b) A[a][b] = A[b][a];
      const double rhs[5] = {-g[0], -g[1], -g[2], -g[3], -g[4]};
      double dl[5];
      if (!cholesky_solve5(A, rhs, dl)) break;
      double na[3];
      for (int c = 0; c < 3; ++c) na[c] = M.a[c] + dl[0] * e1[c] + dl[1] * e2[c];
      const double nl = std::sqrt(na[0] * na[0] + na[1] * na[1] + na[2] * na[2]);
      for (int c = 0; c < 3; ++c) {
        M.a[c] = na[c] / nl;
        M.c[c] = M.c[c] + dl[2] * e1[c] + dl[3] * e2[c];
      }
      M.r = M.r + dl[4];
      if (!(M.r > 0.0)) {
        stop = true;
        break;
      }
    }
  }
  // final inliers under the final model
  double mu[3] = {0, 0, 0};
  int k = 0;
  if (M.r > 0.0) {
    for (int i = 0; i < n; ++i) {
      double w[3], hh;
      const double d = cyl_dist(M, &X[3 * i], w, &hh);
      if (d < dthr && d > -dthr) {
        for (int c = 0; c < 3; ++c) mu[c] += X[3 * i + c];
        ++k;
      }
    }
  }
  if (!(M.r > 0.0) || k < 3) {
    invalid((double)k);
    return;
  }
  const double ht = ((mu[0] / k - M.c[0]) * M.a[0] + (mu[1] / k - M.c[1]) * M.a[1] +
                     (mu[2] / k - M.c[2]) * M.a[2]);
  double ray[3] = {M.a[0], M.a[1], M.a[2]};
  int big = 0;
  for (int c = 1; c < 3; ++c)
    if (std::fabs(ray[c]) > std::fabs(ray[big])) big = c;
  if (ray[big] < 0)
    for (int c = 0; c < 3; ++c) ray[c] = -ray[c];
  for (int c = 0; c < 3; ++c) {
    out[c] = (double)f.m[c] + M.c[c] + ht * M.a[c];
    out[3 + c] = ray[c];
  }
  out[6] = M.r;
  out[7] = (double)k;
}

// ---------------------------------------------------------------------------
// LO-RANSAC refit with FP32 residuals and Jacobians (refit_fp32): the same
// rounds, Gauss-Newton steps, re-selection and output as refit_fp64, with
//   - the point coordinates x = p - m (m the seed frame's centre) and the
//     model taken to FP32 for each step (axis a, axis point c relative to m,
//     radius r, and the FP64 basis e1, e2 of the plane orthogonal to a);
//   - per point, in FP32 with explicit FMAs:
//       q = x - c, h = q.a, w = q - h a, rho = sqrt(w.w), d = rho - r,
//       n1 = (w.e1) / rho, n2 = (w.e2) / rho, J = (-h n1, -h n2, -n1, -n2, -1);
//   - the 20 sums of J J^T (upper triangle) and J d over the round's inliers
//     accumulated in a fixed order: the cluster is cut into chunks of
//     kRefitChunk points, thread t of a chunk's kRefitBlock threads adds the
//     products of points t, t + kRefitBlock, ... with FMAs into an FP32
//     value, the chunk's kRefitBlock values are added by TwoSum (an error-free
//     sum and its rounding error) in groups of kRefitGroup threads in thread
//     order and the groups in order, and the chunks' (sum, error) pairs are
//     added in chunk order the same way; the total is sum + error in FP64;
//   - the 5 x 5 normal equations solved in FP64 (Cholesky), the model updated
//     in FP64.
// The final inliers (|d| < thr) and the centroid of their x use the same
// per-point FP32 distance and the same summation order.  This is the
// arithmetic of the GPU's chunked refit (sloam_amd/csrc/cylfit.hip), which it
// matches bit for bit.
// ---------------------------------------------------------------------------
constexpr int kRefitBlock = 256;
constexpr int kRefitPerThread = 8;
constexpr int kRefitChunk = kRefitBlock * kRefitPerThread;
constexpr int kRefitGroup = 32;

struct Comp {
  float s = 0.0f, e = 0.0f;
};

// s + e = a + b exactly (Knuth's TwoSum)
inline void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  const float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

inline void comp_add(Comp& acc, float x) {
  float s, e;
  two_sum(acc.s, x, s, e);
  acc.s = s;
  acc.e = acc.e + e;
}

inline void comp_add(Comp& acc, const Comp& x) {
  float s, e;
  two_sum(acc.s, x.s, s, e);
  acc.s = s;
  acc.e = acc.e + (e + x.e);
}

// Chunk partial of K per-thread values v[t][k] (t < kRefitBlock).
template <int K>
void chunk_reduce(const float (*v)[K], Comp* out) {
  constexpr int G = kRefitBlock / kRefitGroup;
  for (int k = 0; k < K; ++k) {
    Comp tot;
    for (int g = 0; g < G; ++g) {
      Comp gs;
      for (int t = g * kRefitGroup; t < (g + 1) * kRefitGroup; ++t) comp_add(gs, v[t][k]);
      comp_add(tot, gs);
    }
    out[k] = tot;
  }
}

struct ModelF {
  float a[3], c[3], r, e1[3], e2[3];
};

inline ModelF model_f32(const CylModel& M) {
  ModelF F;
  double e1[3], e2[3];
  basis(M.a, e1, e2);
  for (int i = 0; i < 3; ++i) {
    F.a[i] = (float)M.a[i];
    F.c[i] = (float)M.c[i];
    F.e1[i] = (float)e1[i];
    F.e2[i] = (float)e2[i];
  }
  F.r = (float)M.r;
  return F;
}

// FP32 distance of x (relative to the frame centre) to the model; w, h, rho out.
inline float dist_f32(const ModelF& F, const float x[3], float w[3], float& h, float& rho) {
  const float q0 = x[0] - F.c[0], q1 = x[1] - F.c[1], q2 = x[2] - F.c[2];
  h = std::fmaf(q2, F.a[2], std::fmaf(q1, F.a[1], q0 * F.a[0]));
  w[0] = std::fmaf(-h, F.a[0], q0);
  w[1] = std::fmaf(-h, F.a[1], q1);
  w[2] = std::fmaf(-h, F.a[2], q2);
  rho = std::sqrt(std::fmaf(w[2], w[2], std::fmaf(w[1], w[1], w[0] * w[0])));
  return rho - F.r;
}

void refit_fp32(const float* P, int n, const Frame& f, const Hyp& h, int k0, int rounds,
                int iters, float thr, double out[8]) {
  auto invalid = [&](double cnt) {
    for (int i = 0; i < 6; ++i) out[i] = 0.0;
    out[5] = 1.0;
    out[6] = -1.0;
    out[7] = cnt;
  };
  if (k0 < 3 || !h.valid) {
    invalid((double)k0);
    return;
  }
  std::vector<float> X(3 * (size_t)n);
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) X[3 * i + c] = P[3 * i + c] - f.m[c];
  std::vector<char> in((size_t)n, 0);
  for (int i = 0; i < n; ++i) {
    float u, v, s;
    to_local(f, P + 3 * i, u, v, s);
    in[i] = inlier(u, v, s, h) ? 1 : 0;
  }
  CylModel M;
  for (int c = 0; c < 3; ++c) {
    M.a[c] = f.axis[c];
    M.c[c] = (double)h.cu * (double)f.e1[c] + (double)h.cv * (double)f.e2[c];
  }
  M.r = (double)h.r;
  const int n_chunks = std::max(1, (n + kRefitChunk - 1) / kRefitChunk);
  static thread_local std::vector<float> vbuf;
  vbuf.assign((size_t)kRefitBlock * 20, 0.0f);
  float(*v)[20] = reinterpret_cast<float(*)[20]>(vbuf.data());
  bool stop = false;
  for (int round = 0; round < rounds && !stop; ++round) {
    for (int it = 0; it < iters; ++it) {
      const ModelF F = model_f32(M);
      const bool reselect = (round > 0 && it == 0);
      Comp tot[20];
      long long cnt = 0;
      for (int k = 0; k < n_chunks; ++k) {
        for (int t = 0; t < kRefitBlock; ++t) {
          float* g = v[t];
          for (int j = 0; j < 20; ++j) g[j] = 0.0f;
          for (int j = 0; j < kRefitPerThread; ++j) {
            const int i = k * kRefitChunk + t + kRefitBlock * j;
            if (i >= n) break;
            float w[3], hh, rho;
            const float d = dist_f32(F, &X[3 * (size_t)i], w, hh, rho);
            if (reselect) in[i] = (d < thr && d > -thr) ? 1 : 0;
            if (!in[i]) continue;
            ++cnt;
            if (!(rho > 0.0f)) continue;
            const float inv = 1.0f / rho;
            const float n1 = std::fmaf(w[2], F.e1[2], std::fmaf(w[1], F.e1[1], w[0] * F.e1[0])) * inv;
            const float n2 = std::fmaf(w[2], F.e2[2], std::fmaf(w[1], F.e2[1], w[0] * F.e2[0])) * inv;
            const float J[5] = {-hh * n1, -hh * n2, -n1, -n2, -1.0f};
            int q = 0;
            for (int a = 0; a < 5; ++a)
              for (int b = a; b < 5; ++b, ++q) g[q] = std::fmaf(J[a], J[b], g[q]);
            for (int a = 0; a < 5; ++a) g[15 + a] = std::fmaf(J[a], d, g[15 + a]);
          }
        }
        Comp part[20];
        chunk_reduce<20>(v, part);
        for (int q = 0; q < 20; ++q) comp_add(tot[q], part[q]);
      }
      if (cnt < 5) {
        stop = true;
        break;
      }
      double gs[20];
      for (int q = 0; q < 20; ++q) gs[q] = (double)tot[q].s + (double)tot[q].e;
      double A[5][5], rhs[5], dl[5];
      int q = 0;
      for (int a = 0; a < 5; ++a)
        for (int b = a; b < 5; ++b, ++q) {
          A[a][b] = gs[q];
          A[b][a] = gs[q];
        }
      for (int a = 0; a < 5; ++a) rhs[a] = -gs[15 + a];
      if (!cholesky_solve5(A, rhs, dl)) break;
      double e1[3], e2[3];
      basis(M.a, e1, e2);
      double na[3];
      for (int c = 0; c < 3; ++c) na[c] = M.a[c] + dl[0] * e1[c] + dl[1] * e2[c];
      const double nl = std::sqrt(na[0] * na[0] + na[1] * na[1] + na[2] * na[2]);
      for (int c = 0; c < 3; ++c) {
        M.a[c] = na[c] / nl;
        M.c[c] = M.c[c] + dl[2] * e1[c] + dl[3] * e2[c];
      }
      M.r = M.r + dl[4];
      if (!(M.r > 0.0)) {
        stop = true;
        break;
      }
    }
  }
  // final inliers under the final model: count and centroid of x
  long long k = 0;
  Comp mu[3];
  if (M.r > 0.0) {
    const ModelF F = model_f32(M);
    float (*v3)[3] = reinterpret_cast<float(*)[3]>(vbuf.data());
    for (int ck = 0; ck < n_chunks; ++ck) {
      for (int t = 0; t < kRefitBlock; ++t) {
        float* g = v3[t];
        g[0] = g[1] = g[2] = 0.0f;
        for (int j = 0; j < kRefitPerThread; ++j) {
          const int i = ck * kRefitChunk + t + kRefitBlock * j;
          if (i >= n) break;
          float w[3], hh, rho;
          const float d = dist_f32(F, &X[3 * (size_t)i], w, hh, rho);
          if (d < thr && d > -thr) {
            for (int c = 0; c < 3; ++c) g[c] = g[c] + X[3 * (size_t)i + c];
            ++k;
          }
        }
      }
      Comp part[3];
      chunk_reduce<3>(v3, part);
      for (int c = 0; c < 3; ++c) comp_add(mu[c], part[c]);
    }
  }
  if (!(M.r > 0.0) || k < 3) {
    invalid((double)k);
    return;
  }
  const double kd = (double)k;
  double mean[3];
  for (int c = 0; c < 3; ++c) mean[c] = ((double)mu[c].s + (double)mu[c].e) / kd;
  const double ht = ((mean[0] - M.c[0]) * M.a[0] + (mean[1] - M.c[1]) * M.a[1] +
                     (mean[2] - M.c[2]) * M.a[2]);
  double ray[3] = {M.a[0], M.a[1], M.a[2]};
  int big = 0;
  for (int c = 1; c < 3; ++c)
    if (std::fabs(ray[c]) > std::fabs(ray[big])) big = c;
  if (ray[big] < 0)
    for (int c = 0; c < 3; ++c) ray[c] = -ray[c];
  for (int c = 0; c < 3; ++c) {
    out[c] = (double)f.m[c] + M.c[c] + ht * M.a[c];
    out[3 + c] = ray[c];
  }
  out[6] = M.r;
  out[7] = kd;
}

}  // namespace
}  // namespace oracle

using namespace oracle;

extern "C" {

int oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  U4 c = {{ctr[0], ctr[1], ctr[2], ctr[3]}};
  const U4 r = philox4x32_10(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
  return 0;
}

int oracle_draw3(uint32_t x0, uint32_t x1, uint32_t x2, uint32_t n, int32_t out[3]) {
  if (n < 3) return -1;
  draw3(x0, x1, x2, n, out);
  return 0;
}

int oracle_eig3(const double a[6], double w[3], double V[9]) {
  double VV[3][3];
  eig3(a, w, VV);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) V[3 * i + j] = VV[i][j];
  return 0;
}

// Seed frame of one cluster: m[3], e1[3], e2[3] (as float) and axis[3].
int oracle_cyl_seed(const float* xyz, int n, float frame_f[9], double axis[3]) {
  const Frame f = seed_frame(xyz, n);
  for (int k = 0; k < 3; ++k) {
    frame_f[k] = f.m[k];
    frame_f[3 + k] = f.e1[k];
    frame_f[6 + k] = f.e2[k];
    axis[k] = f.axis[k];
  }
  return 0;
}

// Hypothesis circle of one (cluster, h): cu, cv, r, valid (as float) and the
// predicate parameters A, B, L, U.
int oracle_cyl_hypothesis(const float* xyz, int n, int cluster, int h, float thr,
                          float max_radius, uint64_t seed, int tag, int32_t idx[3],
                          float circle[4], float prm[4]) {
  if (n < 3) return -1;
  const Frame f = seed_frame(xyz, n);
  const Hyp hy = hypothesis(xyz, n, cluster, h, f, thr, max_radius, (uint32_t)seed,
                            (uint32_t)(seed >> 32), (uint32_t)tag, idx);
  circle[0] = hy.cu;
  circle[1] = hy.cv;
  circle[2] = hy.r;
  circle[3] = hy.valid ? 1.0f : 0.0f;
  prm[0] = hy.A;
  prm[1] = hy.B;
  prm[2] = hy.L;
  prm[3] = hy.U;
  return 0;
}

// The whole batch, same outputs as sloam_cyl_plan_exec.  n_threads > 1 runs
// clusters in parallel (OpenMP); results do not depend on it.
// proposal (nullable): 8 floats per cluster (axis, centre, radius, valid);
// a valid one is hypothesis 0.
// fp64_refit: refit in FP64 (refit_fp64) instead of the GPU's FP32
// arithmetic (refit_fp32).
int oracle_cyl_fit_batch(const float* xyz, const int32_t* offsets, int n_clusters, int n_hyp,
                         float thr, float max_radius, int refit_rounds, int refit_iters, uint64_t seed,
                         int tag, const float* proposal, int32_t* counts, int32_t* samples,
                         uint32_t* mask, int32_t* best_out, double* model, int n_threads,
                         int fp64_refit) {
  if (n_clusters < 0 || n_hyp < 1) return -1;
  const int64_t n_points = offsets[n_clusters];
  if (mask) std::memset(mask, 0, sizeof(uint32_t) * (size_t)((n_points + 31) / 32));
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads > 0 ? n_threads : 1)
  for (int c = 0; c < n_clusters; ++c) {
    const int off = offsets[c];
    const int n = offsets[c + 1] - off;
    const float* P = xyz + 3 * (size_t)off;
    const Frame f = seed_frame(P, n);
    const float* pc = (proposal && proposal[8 * (size_t)c + 7] > 0.0f) ? proposal + 8 * (size_t)c
                                                                         : nullptr;
    std::vector<float> U((size_t)n), V((size_t)n), S((size_t)n);
    for (int i = 0; i < n; ++i) to_local(f, P + 3 * i, U[i], V[i], S[i]);
    int best = -1, best_count = -1;
    for (int h = 0; h < n_hyp; ++h) {
      int count = 0;
      int32_t idx[3] = {-1, -1, -1};
      if (n >= 3) {
        const Hyp hy =
            hypothesis(P, n, c, h, f, thr, max_radius, k0, k1, (uint32_t)tag, idx, pc);
        for (int i = 0; i < n; ++i) count += inlier(U[i], V[i], S[i], hy) ? 1 : 0;
      }
      if (counts) counts[(size_t)c * n_hyp + h] = count;
      if (samples)
        for (int k = 0; k < 3; ++k) samples[3 * ((size_t)c * n_hyp + h) + k] = idx[k];
      if (count > best_count) {
        best_count = count;
        best = h;
      }
    }
    if (n < 3) best = -1;
    if (best_out) best_out[c] = best;
    double* out = model + (size_t)c * 8;
    if (best < 0) {
      for (int i = 0; i < 6; ++i) out[i] = 0.0;
      out[5] = 1.0;
      out[6] = -1.0;
      out[7] = 0.0;
      continue;
    }
    int32_t idx[3];
    const Hyp hb = hypothesis(P, n, c, best, f, thr, max_radius, k0, k1, (uint32_t)tag, idx, pc);
    if (mask) {
      for (int i = 0; i < n; ++i) {
        if (!inlier(U[i], V[i], S[i], hb)) continue;
        const int64_t g = off + i;
#pragma omp atomic
        mask[g >> 5] |= (1u << (g & 31));
      }
    }
    if (fp64_refit)
      refit_fp64(P, n, f, hb, refit_rounds, refit_iters, thr, out);
    else
      refit_fp32(P, n, f, hb, best_count, refit_rounds, refit_iters, thr, out);
  }
  return 0;
}

// FP64 point-to-cylinder distance of every point of a cluster to hypothesis h
// (reference formula, cylinder.cpp:196-203, with root = circle centre lifted
// to 3-D and ray = seed axis).  Used to bound the FP32 predicate's error.
int oracle_cyl_distances_fp64(const float* xyz, int n, int cluster, int h, float thr,
                              float max_radius, uint64_t seed, int tag, double* dist) {
  if (n < 3) return -1;
  const Frame f = seed_frame(xyz, n);
  int32_t idx[3];
  const Hyp hy = hypothesis(xyz, n, cluster, h, f, thr, max_radius, (uint32_t)seed,
                            (uint32_t)(seed >> 32), (uint32_t)tag, idx);
  if (!hy.valid) return 1;
  double root[3], ray[3];
  for (int c = 0; c < 3; ++c) {
    root[c] = (double)f.m[c] + (double)hy.cu * (double)f.e1[c] + (double)hy.cv * (double)f.e2[c];
    // the FP32 basis defines the plane in which (u, v) live; its normal is the ray
    ray[c] = 0.0;
  }
  const double e1[3] = {f.e1[0], f.e1[1], f.e1[2]};
  const double e2[3] = {f.e2[0], f.e2[1], f.e2[2]};
  ray[0] = e1[1] * e2[2] - e1[2] * e2[1];
  ray[1] = e1[2] * e2[0] - e1[0] * e2[2];
  ray[2] = e1[0] * e2[1] - e1[1] * e2[0];
  for (int i = 0; i < n; ++i) {
    const double p[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    const double pr[3] = {p[0] - root[0], p[1] - root[1], p[2] - root[2]};
    const double rr = ray[0] * ray[0] + ray[1] * ray[1] + ray[2] * ray[2];
    const double t = (pr[0] * ray[0] + pr[1] * ray[1] + pr[2] * ray[2]) / rr;
    const double q[3] = {p[0] - (root[0] + t * ray[0]), p[1] - (root[1] + t * ray[1]),
                         p[2] - (root[2] + t * ray[2])};
    dist[i] = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) - (double)hy.r;
  }
  return 0;
}

}  // extern "C"
