// TEST INFRASTRUCTURE ONLY -- CPU oracle of the ground-plane path.
//
// 1. Binning + per-cell plane fit, restating sloam::sloam::binGroundPoints
//    (sloam/src/core/sloam.cpp:330-386) and Plane::Plane / computeModel
//    (sloam/src/objects/plane.cpp:3-17, 96-128) with euclideanDist2D /
//    computeCentroid (sloam/include/helpers/utils.h:9-28):
//      r = sqrt(float(dx^2) + float(dy^2)) with the squares taken in double
//      (pow_2 takes a double and returns float); theta = atan2f; bins in
//      double with PIDEF = 3.14159265; cells with more than 1/retain points
//      keep their floor(size * retain) lowest points by z (std::sort in the
//      reference: ties unspecified, here input order), other cells keep all
//      points in input order; centroid = float running sums / (double) n;
//      normal = left singular vector of the smallest singular value of the
//      centred 3 x N matrix (JacobiSVD) == smallest eigenvector of its scatter
//      matrix, computed here by FP64 Jacobi; d = -n . centroid.
//    The reference leaves the normal's sign to Eigen's SVD; here it points up
//    (z > 0), the orientation in which computeModels' angle check accepts a
//    level plane (sloam.cpp:403-406).
// 2. Point-level plane RANSAC (project-defined; see sloam_amd/csrc/planefit.hip).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "oracle_common.hpp"

using namespace oracle;

namespace {

void orient_up(double n[3]) {
  if (n[2] < 0.0) {
    n[0] = -n[0];
    n[1] = -n[1];
    n[2] = -n[2];
  }
  if (n[2] == 0.0) {
    int big = 0;
    for (int i = 1; i < 3; ++i)
      if (std::fabs(n[i]) > std::fabs(n[big])) big = i;
    if (n[big] < 0)
      for (int i = 0; i < 3; ++i) n[i] = -n[i];
  }
}

void smallest_eigvec(const double S[6], double n[3]) {
  double w[3], V[3][3];
  eig3(S, w, V);
  const int k = argmin3(w);
  for (int i = 0; i < 3; ++i) n[i] = V[i][k];
}

struct PlFrame {
  float m[3];
};

struct PlHyp {
  float nx, ny, nz, d;
  bool valid;
};

PlHyp plane_hyp(const float a[3], const float b[3], const float c[3]) {
  PlHyp h;
  const float ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
  const float vx = c[0] - a[0], vy = c[1] - a[1], vz = c[2] - a[2];
  float nx = uy * vz - uz * vy;
  float ny = uz * vx - ux * vz;
  float nz = ux * vy - uy * vx;
  const float len = std::sqrt(nx * nx + ny * ny + nz * nz);
  h.valid = len > 0.0f;
  if (h.valid) {
    nx = nx / len;
    ny = ny / len;
    nz = nz / len;
    h.nx = nx;
    h.ny = ny;
    h.nz = nz;
    h.d = -((nx * a[0] + ny * a[1]) + nz * a[2]);
  } else {
    h.nx = h.ny = h.nz = 0.f;
    h.d = INFINITY;
  }
  return h;
}

inline bool plane_inlier(float x, float y, float z, const PlHyp& h, float thr) {
  const float t = std::fmaf(h.nx, x, std::fmaf(h.ny, y, std::fmaf(h.nz, z, h.d)));
  return std::fabs(t) < thr;
}

}  // namespace

extern "C" {

// pts: n x 4 floats.  info: cells x 4 int32 {valid, kept, binned, 0};
// model: cells x 8 doubles {plane[4], centroid[3], 0}; features: cells x nf x 4.
int oracle_ground_planes(const float* pts, int n, int R, int T, double min_d, double max_d,
                         double retain, int nf, const float origin[3], int32_t* info,
                         double* model, float* features) {
  const int n_cells = R * T;
  std::vector<std::vector<int>> cells(n_cells);
  for (int i = 0; i < n; ++i) {
    const float* p = pts + 4 * (size_t)i;
    if (!(std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]))) continue;
    const float dx = origin[0] - p[0], dy = origin[1] - p[1];
    const float a = (float)((double)dx * (double)dx);
    const float b = (float)((double)dy * (double)dy);
    const double r = (double)std::sqrt(a + b);
    if (!(r < max_d && r > min_d)) continue;
    const double th = (double)std::atan2(p[1] - origin[1], p[0] - origin[0]);
    int rb = (int)std::floor(r / (max_d / (double)R));
    int tb = (int)std::floor((3.14159265 + th) / (2 * 3.14159265 / (double)T));
    rb = std::max(0, std::min(rb, R - 1));
    tb = std::max(0, std::min(tb, T - 1));
    cells[rb * T + tb].push_back(i);
  }
  const double retain_num = 1.0 / retain;
  for (int c = 0; c < n_cells; ++c) {
    std::vector<int>& L = cells[c];
    const int s = (int)L.size();
    int k = s;
    if (retain_num < (double)s) {
      k = std::min(s, (int)((double)s / retain_num));
      std::stable_sort(L.begin(), L.end(), [&](int x, int y) { return pts[4 * (size_t)x + 2] < pts[4 * (size_t)y + 2]; });
      L.resize(k);
    }
    int32_t* inf = info + 4 * (size_t)c;
    double* out = model + 8 * (size_t)c;
    const bool valid = k >= nf && k > 0;
    inf[0] = valid ? 1 : 0;
    inf[1] = k;
    inf[2] = s;
    inf[3] = 0;
    if (!valid) {
      for (int i = 0; i < 8; ++i) out[i] = 0.0;
      continue;
    }
    float cx = 0.f, cy = 0.f, cz = 0.f;
    for (int i : L) {
      cx += pts[4 * (size_t)i];
      cy += pts[4 * (size_t)i + 1];
      cz += pts[4 * (size_t)i + 2];
    }
    cx = (float)((double)cx / (double)k);
    cy = (float)((double)cy / (double)k);
    cz = (float)((double)cz / (double)k);
    double S[6] = {0, 0, 0, 0, 0, 0};
    for (int i : L) {
      const double x = (double)(pts[4 * (size_t)i] - cx);
      const double y = (double)(pts[4 * (size_t)i + 1] - cy);
      const double z = (double)(pts[4 * (size_t)i + 2] - cz);
      S[0] += x * x;
      S[1] += x * y;
      S[2] += x * z;
      S[3] += y * y;
      S[4] += y * z;
      S[5] += z * z;
    }
    double nrm[3];
    smallest_eigvec(S, nrm);
    orient_up(nrm);
    out[0] = nrm[0];
    out[1] = nrm[1];
    out[2] = nrm[2];
    out[3] = -(nrm[0] * (double)cx + nrm[1] * (double)cy + nrm[2] * (double)cz);
    out[4] = cx;
    out[5] = cy;
    out[6] = cz;
    out[7] = 0.0;
    for (int j = 0; j < nf; ++j)
      std::memcpy(features + 4 * ((size_t)c * nf + j), pts + 4 * (size_t)L[j], 16);
  }
  return 0;
}

int oracle_plane_fit_batch(const float* xyz, const int32_t* offsets, int n_sets, int n_hyp,
                           float thr, uint64_t seed, int tag, int32_t* counts, int32_t* samples,
                           uint32_t* mask, int32_t* best_out, double* model, int n_threads) {
  const int64_t n_points = offsets[n_sets];
  if (mask) std::memset(mask, 0, sizeof(uint32_t) * (size_t)((n_points + 31) / 32));
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads > 0 ? n_threads : 1)
  for (int c = 0; c < n_sets; ++c) {
    const int off = offsets[c];
    const int n = offsets[c + 1] - off;
    const float* P = xyz + 3 * (size_t)off;
    PlFrame f;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], P[3 * i + k]);
        hi[k] = std::max(hi[k], P[3 * i + k]);
      }
    for (int k = 0; k < 3; ++k) f.m[k] = n > 0 ? 0.5f * (lo[k] + hi[k]) : 0.0f;
    std::vector<float> X(3 * (size_t)n);
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) X[3 * i + k] = P[3 * i + k] - f.m[k];
    auto hyp = [&](int h, int32_t idx[3]) {
      const U4 ctr = {{(uint32_t)h, (uint32_t)c, (uint32_t)tag, 0u}};
      const U4 x = philox4x32_10(ctr, k0, k1);
      draw3(x.v[0], x.v[1], x.v[2], (uint32_t)n, idx);
      return plane_hyp(&X[3 * idx[0]], &X[3 * idx[1]], &X[3 * idx[2]]);
    };
    int best = -1, best_count = -1;
    for (int h = 0; h < n_hyp; ++h) {
      int32_t idx[3] = {-1, -1, -1};
      int count = 0;
      if (n >= 3) {
        const PlHyp hy = hyp(h, idx);
        for (int i = 0; i < n; ++i) count += plane_inlier(X[3 * i], X[3 * i + 1], X[3 * i + 2], hy, thr);
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
    double* out = model + 8 * (size_t)c;
    for (int i = 0; i < 8; ++i) out[i] = 0.0;
    if (best < 0) continue;
    int32_t idx[3];
    const PlHyp hb = hyp(best, idx);
    double mu[3] = {0, 0, 0};
    std::vector<int> in;
    for (int i = 0; i < n; ++i) {
      if (!plane_inlier(X[3 * i], X[3 * i + 1], X[3 * i + 2], hb, thr)) continue;
      in.push_back(i);
      if (mask) {
        const int64_t g = off + i;
#pragma omp atomic
        mask[g >> 5] |= (1u << (g & 31));
      }
      for (int k = 0; k < 3; ++k) mu[k] += (double)P[3 * i + k] - (double)f.m[k];
    }
    const double kk = (double)in.size();
    if (in.size() < 3 || !hb.valid) {
      out[7] = kk;
      continue;
    }
    for (int k = 0; k < 3; ++k) mu[k] /= kk;
    double S[6] = {0, 0, 0, 0, 0, 0};
    for (int i : in) {
      const double x = ((double)P[3 * i] - (double)f.m[0]) - mu[0];
      const double y = ((double)P[3 * i + 1] - (double)f.m[1]) - mu[1];
      const double z = ((double)P[3 * i + 2] - (double)f.m[2]) - mu[2];
      S[0] += x * x;
      S[1] += x * y;
      S[2] += x * z;
      S[3] += y * y;
      S[4] += y * z;
      S[5] += z * z;
    }
    double nrm[3];
    smallest_eigvec(S, nrm);
    orient_up(nrm);
    const double cxyz[3] = {(double)f.m[0] + mu[0], (double)f.m[1] + mu[1], (double)f.m[2] + mu[2]};
    const double dd = -(nrm[0] * cxyz[0] + nrm[1] * cxyz[1] + nrm[2] * cxyz[2]);
    double cnt = 0;
    for (int i = 0; i < n; ++i) {
      const double t = nrm[0] * (double)P[3 * i] + nrm[1] * (double)P[3 * i + 1] +
                       nrm[2] * (double)P[3 * i + 2] + dd;
      if (std::fabs(t) < (double)thr) cnt += 1;
    }
    out[0] = nrm[0];
    out[1] = nrm[1];
    out[2] = nrm[2];
    out[3] = dd;
    out[4] = cxyz[0];
    out[5] = cxyz[1];
    out[6] = cxyz[2];
    out[7] = cnt;
  }
  return 0;
}

}  // extern "C"
