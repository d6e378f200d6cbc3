// TEST INFRASTRUCTURE ONLY -- CPU oracle of the reference's cylinder axis fit:
// the RANSAC line fit over trellis vertex medians of Cylinder::computeModel
// (sloam/src/objects/cylinder.cpp:116-139: SACSegmentation, SACMODEL_LINE,
// SAC_RANSAC, threshold 0.25, optimizeCoefficients).
//
// PCL 1.10 is not available here; its semantics are restated from its
// published source (recalled, parity against PCL itself unpinned):
//   * model RNG boost::mt19937 seeded 12345; rnd() = uniform_int(0, INT_MAX),
//     which over mt19937 is mt() >> 1;
//   * drawIndexSample: for i < 2: swap(shuf[i], shuf[i + rnd() % (n - i)]),
//     the shuffled index array persisting across draws; up to 1000 draws per
//     sample until isSampleGood (all of x, y, z of the two points differ);
//   * computeModelCoefficients fails when all |coordinate differences| <=
//     FLT_EPSILON (counts as a skipped sample, not an iteration);
//   * coefficients (p0, normalise(p1 - p0)) in float; inlier test
//     ||(p0 - p) x d||^2 (float) < thr^2 (double);
//   * RANSAC: while iterations < k and skipped < 10 * max_iterations; on a
//     strictly better count k = log(1 - 0.99) / log(1 - w^2) with
//     1 - w^2 clamped to [eps, 1 - eps]; stop after iterations > max (50);
//   * optimisation: float centroid and scatter matrix of the inliers, the
//     eigenvector of the largest eigenvalue by the cross product of two rows of
//     (C - lambda I) (the largest of the three products), re-selected inliers.
// Here lambda comes from an FP64 Jacobi solve instead of PCL's
// closed-form float eigen33 roots.
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle_common.hpp"

using namespace oracle;

namespace {

// MT19937 (Matsumoto & Nishimura 1998), 32-bit
struct MT {
  uint32_t s[624];
  int idx;
  explicit MT(uint32_t seed) {
    s[0] = seed;
    for (int i = 1; i < 624; ++i) s[i] = 1812433253u * (s[i - 1] ^ (s[i - 1] >> 30)) + (uint32_t)i;
    idx = 624;
  }
  uint32_t operator()() {
    if (idx >= 624) {
      for (int i = 0; i < 624; ++i) {
        const uint32_t y = (s[i] & 0x80000000u) | (s[(i + 1) % 624] & 0x7fffffffu);
        s[i] = s[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
      }
      idx = 0;
    }
    uint32_t y = s[idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }
};

// ||(p0 - q) x d||^2 in float, d re-normalised first as the line model's
// distance functions do
struct LineEval {
  float p[3], d[3];
};

inline LineEval line_eval(const float c[6]) {
  LineEval e;
  const float l = std::sqrt((c[3] * c[3] + c[4] * c[4]) + c[5] * c[5]);
  for (int i = 0; i < 3; ++i) {
    e.p[i] = c[i];
    e.d[i] = c[3 + i] / l;
  }
  return e;
}

inline float sqr_line_dist(const float* q, const LineEval& e) {
  const float ax = e.p[0] - q[0], ay = e.p[1] - q[1], az = e.p[2] - q[2];
  const float x = ay * e.d[2] - az * e.d[1];
  const float y = az * e.d[0] - ax * e.d[2];
  const float z = ax * e.d[1] - ay * e.d[0];
  return (x * x + y * y) + z * z;
}

}  // namespace

extern "C" {

int oracle_mt19937(uint32_t seed, int n, uint32_t* out) {
  MT mt(seed);
  for (int i = 0; i < n; ++i) out[i] = mt();
  return 0;
}

// One set (tree): n points (3 floats).  coeff[6] float, inliers mask (n
// bytes), stats[4] = {ok, inliers, iterations, skipped}.
int oracle_line_ransac(const float* P, int n, double thr, int max_iter, double prob,
                       uint32_t seed, int optimize, float coeff[6], uint8_t* inl,
                       int32_t stats[4]) {
  stats[0] = stats[1] = stats[2] = stats[3] = 0;
  for (int i = 0; i < 6; ++i) coeff[i] = 0.f;
  if (inl) std::memset(inl, 0, n);
  if (n < 2) return 0;
  MT mt(seed);
  std::vector<int> shuf(n);
  for (int i = 0; i < n; ++i) shuf[i] = i;
  const double sqr_thr = thr * thr;
  const double log_p = std::log(1.0 - prob);
  const double one_over = 1.0 / (double)n;
  double k = DBL_MAX;
  int iterations = 0, skipped = 0, best = 0;
  const int max_skip = max_iter * 10;
  bool have = false;
  float model[6] = {0, 0, 0, 0, 0, 0};
  while (iterations < k && skipped < max_skip) {
    int s0 = -1, s1 = -1;
    bool good = false;
    for (int t = 0; t < 1000; ++t) {
      for (int i = 0; i < 2; ++i) {
        const int j = i + (int)((mt() >> 1) % (uint32_t)(n - i));
        std::swap(shuf[i], shuf[j]);
      }
      s0 = shuf[0];
      s1 = shuf[1];
      const float* a = P + 3 * s0;
      const float* b = P + 3 * s1;
      if (a[0] != b[0] && a[1] != b[1] && a[2] != b[2]) {
        good = true;
        break;
      }
    }
    if (!good) break;
    const float* a = P + 3 * s0;
    const float* b = P + 3 * s1;
    if (std::fabs(a[0] - b[0]) <= FLT_EPSILON && std::fabs(a[1] - b[1]) <= FLT_EPSILON &&
        std::fabs(a[2] - b[2]) <= FLT_EPSILON) {
      ++skipped;
      continue;
    }
    float c[6] = {a[0], a[1], a[2], b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const float nn = std::sqrt((c[3] * c[3] + c[4] * c[4]) + c[5] * c[5]);
    c[3] = c[3] / nn;
    c[4] = c[4] / nn;
    c[5] = c[5] / nn;
    const LineEval ev = line_eval(c);
    int cnt = 0;
    for (int i = 0; i < n; ++i) cnt += (double)sqr_line_dist(P + 3 * i, ev) < sqr_thr;
    if (cnt > best) {
      best = cnt;
      std::memcpy(model, c, sizeof(model));
      have = true;
      const double w = (double)best * one_over;
      double p_no = 1.0 - w * w;  // pow(w, sample size 2)
      p_no = std::fmax(DBL_EPSILON, p_no);
      p_no = std::fmin(1.0 - DBL_EPSILON, p_no);
      k = log_p / std::log(p_no);
    }
    ++iterations;
    if (iterations > max_iter) break;
  }
  stats[2] = iterations;
  stats[3] = skipped;
  if (!have) return 0;
  std::vector<int> in;
  {
    const LineEval ev = line_eval(model);
    for (int i = 0; i < n; ++i)
      if ((double)sqr_line_dist(P + 3 * i, ev) < sqr_thr) in.push_back(i);
  }
  if (optimize && (int)in.size() > 2) {
    float cx = 0, cy = 0, cz = 0;
    for (int i : in) {
      cx += P[3 * i];
      cy += P[3 * i + 1];
      cz += P[3 * i + 2];
    }
    const float m = (float)in.size();
    cx = cx / m;
    cy = cy / m;
    cz = cz / m;
    float C[6] = {0, 0, 0, 0, 0, 0};  // xx xy xz yy yz zz
    for (int i : in) {
      const float x = P[3 * i] - cx, y = P[3 * i + 1] - cy, z = P[3 * i + 2] - cz;
      C[3] += y * y;
      C[4] += y * z;
      C[5] += z * z;
      C[0] += x * x;
      C[1] += x * y;
      C[2] += x * z;
    }
    const double Cd[6] = {C[0], C[1], C[2], C[3], C[4], C[5]};
    double w[3], V[3][3];
    eig3(Cd, w, V);
    const float lam = (float)w[argmax3(w)];
    // eigenvector by cross products of the rows of (C - lam I), scaled to
    // the matrix's largest magnitude
    float mx = 0.f;
    for (int i = 0; i < 6; ++i) mx = std::fmax(mx, std::fabs(C[i]));
    if (mx <= FLT_MIN) mx = 1.0f;
    const float r0[3] = {C[0] / mx - lam / mx, C[1] / mx, C[2] / mx};
    const float r1[3] = {C[1] / mx, C[3] / mx - lam / mx, C[4] / mx};
    const float r2[3] = {C[2] / mx, C[4] / mx, C[5] / mx - lam / mx};
    auto cross = [](const float a[3], const float b[3], float o[3]) {
      o[0] = a[1] * b[2] - a[2] * b[1];
      o[1] = a[2] * b[0] - a[0] * b[2];
      o[2] = a[0] * b[1] - a[1] * b[0];
    };
    float v1[3], v2[3], v3[3];
    cross(r0, r1, v1);
    cross(r0, r2, v2);
    cross(r1, r2, v3);
    const float l1 = (v1[0] * v1[0] + v1[1] * v1[1]) + v1[2] * v1[2];
    const float l2 = (v2[0] * v2[0] + v2[1] * v2[1]) + v2[2] * v2[2];
    const float l3 = (v3[0] * v3[0] + v3[1] * v3[1]) + v3[2] * v3[2];
    const float* v = v3;
    float l = l3;
    if (l1 >= l2 && l1 >= l3) {
      v = v1;
      l = l1;
    } else if (l2 >= l1 && l2 >= l3) {
      v = v2;
      l = l2;
    }
    const float sl = std::sqrt(l);
    model[0] = cx;
    model[1] = cy;
    model[2] = cz;
    model[3] = v[0] / sl;
    model[4] = v[1] / sl;
    model[5] = v[2] / sl;
    in.clear();
    const LineEval ev = line_eval(model);
    for (int i = 0; i < n; ++i)
      if ((double)sqr_line_dist(P + 3 * i, ev) < sqr_thr) in.push_back(i);
  }
  std::memcpy(coeff, model, sizeof(model));
  if (inl)
    for (int i : in) inl[i] = 1;
  stats[0] = 1;
  stats[1] = (int)in.size();
  return 0;
}

}  // extern "C"
