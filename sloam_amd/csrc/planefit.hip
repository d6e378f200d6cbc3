// Ground planes for gfx950: the reference's binned per-cell plane fit and a
// batched point-level plane RANSAC.
//
// 1. sloam_ground_planes -- sloam::sloam::binGroundPoints + Plane::Plane /
//    Plane::computeModel (sloam/src/core/sloam.cpp:330-386, 391-400;
//    sloam/src/objects/plane.cpp:3-17, 96-128; utils.h:9-28):
//      bin      one thread per point: 2-D radius in float (pow_2 squares in
//               double and rounds to float), angle atan2f, radius / angle bins
//               in double; keys (cell, index) and (cell, z) for two stable
//               radix sorts (rocPRIM) that give each cell's points in input
//               order and in increasing z (ties: input order)
//      fit      one workgroup per cell: the bottom floor(size / retainNum)
//               points by z when retainNum < size (else all points, input
//               order); float centroid summed sequentially in that order as
//               the reference does; FP64 covariance of the float-centred
//               points; FP64 Jacobi: normal = smallest eigenvector (the
//               reference's JacobiSVD U[:,2]); features = the first
//               numGroundFeatures points
// 2. sloam_plane_plan_* -- point-level plane RANSAC (no counterpart in the
//    reference, which fits each cell by SVD without RANSAC): Philox4x32-10
//    samples of 3 points, plane through them in FP32, inlier |n.p + d| < thr
//    counted by wavefront ballot; best = max count, lowest index; FP64 PCA
//    refit over the inliers.  Predicate: Plane::distance(const PointT&)
//    (sloam/src/objects/plane.cpp:136-151) for a unit normal.
#include <cstring>

#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <new>
#include <numeric>
#include <vector>

namespace {

constexpr int kGB = 256;
constexpr double kPiDef = 3.14159265;  // PIDEF, sloam/include/helpers/definitions.h:28
constexpr int kCentroidTile = 2048;

// unsigned key in the order of the floats' <, -0 and +0 equal (so that their
// ties keep input order, as std::stable_sort with < does)
__device__ inline uint32_t orderable(float f) {
  const uint32_t u = __float_as_uint(f == 0.0f ? 0.0f : f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

struct GroundCfgDev {
  int R, T, n_cells, nf;
  double min_d, max_d, retain_num;
  float ox, oy;
};

// points per workgroup of ground_bin; cells counted in LDS up to kBinLds cells
constexpr int kBinPts = 2048;
constexpr int kBinLds = 4096;

__global__ __launch_bounds__(kGB) void ground_bin_kernel(
    const float4* __restrict__ pts, int cap, const int32_t* __restrict__ d_n, GroundCfgDev cfg,
    unsigned long long* __restrict__ key_idx, unsigned long long* __restrict__ key_z,
    int32_t* __restrict__ vals_idx, int32_t* __restrict__ vals_z, int32_t* __restrict__ counts) {
  // per-workgroup cell counts in LDS, one global atomic per non-empty cell
  // (all points of a scan fall in few cells: per-point global atomics on the
  // same counters serialise)
  __shared__ int hcount[kBinLds];
  const bool lds = cfg.n_cells <= kBinLds;
  if (lds)
    for (int c = threadIdx.x; c < cfg.n_cells; c += kGB) hcount[c] = 0;
  __syncthreads();
  const int n = min(d_n ? *d_n : cap, cap);
  const int base = blockIdx.x * kBinPts;
  for (int j = threadIdx.x; j < kBinPts; j += kGB) {
    const int i = base + j;
    if (i >= n) break;  // the sorts stop at n
    // points outside the grid get cell n_cells: after every cell in both sorts
    unsigned long long cell = (unsigned long long)cfg.n_cells;
    const float4 p = pts[i];
    const float z = p.z;
    if (isfinite(p.x) && isfinite(p.y) && isfinite(p.z)) {
      const float dx = cfg.ox - p.x, dy = cfg.oy - p.y;
      const float a = (float)((double)dx * (double)dx);
      const float b = (float)((double)dy * (double)dy);
      const double r = (double)sqrtf(a + b);
      if (r < cfg.max_d && r > cfg.min_d) {
        const double th = (double)atan2f(p.y - cfg.oy, p.x - cfg.ox);
        int rb = (int)floor(r / (cfg.max_d / (double)cfg.R));
        int tb = (int)floor((kPiDef + th) / (2 * kPiDef / (double)cfg.T));
        rb = max(0, min(rb, cfg.R - 1));
        tb = max(0, min(tb, cfg.T - 1));
        cell = (unsigned long long)(rb * cfg.T + tb);
        if (lds)
          atomicAdd(&hcount[(int)cell], 1);
        else
          atomicAdd(counts + cell, 1);
      }
    }
    // (cell) and (cell, z): the stable sort keeps input order within equal keys
    key_idx[i] = cell;
    key_z[i] = (cell << 32) | (unsigned long long)orderable(z);
    vals_idx[i] = i;
    vals_z[i] = i;
  }
  __syncthreads();
  if (lds)
    for (int c = threadIdx.x; c < cfg.n_cells; c += kGB) {
      const int v = hcount[c];
      if (v) atomicAdd(counts + c, v);
    }
}

__global__ __launch_bounds__(kGB) void ground_offsets_kernel(const int32_t* __restrict__ counts,
                                                             int n_cells,
                                                             int32_t* __restrict__ offsets) {
  // one workgroup; n_cells is small (radii x theta bins)
  __shared__ int carry;
  __shared__ int wtot[kGB / 64];
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int base = 0; base < n_cells; base += kGB) {
    const int i = base + (int)threadIdx.x;
    const int v = i < n_cells ? counts[i] : 0;
    int s = v;
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(s, o, 64);
      if (lane >= o) s += t;
    }
    if (lane == 63) wtot[wave] = s;
    __syncthreads();
    int woff = 0;
    for (int w = 0; w < wave; ++w) woff += wtot[w];
    const int excl = carry + woff + s - v;
    if (i < n_cells) offsets[i] = excl;
    __syncthreads();
    if (threadIdx.x == kGB - 1) carry = excl + v;
    __syncthreads();
  }
  if (threadIdx.x == 0) offsets[n_cells] = carry;
}

template <int K>
__device__ inline void gsum(double (&v)[K], double* lds) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = sl_wave_sum(v[k]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) lds[wave * K + k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double t = lds[k];
    for (int w = 1; w < kGB / 64; ++w) t += lds[w * K + k];
    v[k] = t;
  }
  __syncthreads();
}

// info: {valid, kept, binned, 0}; model: plane[4], centroid[3], unused.
// offsets (nullable): the cells' exclusive prefix of counts; without it
// (few cells) each workgroup adds the counts of the cells before its own.
__global__ __launch_bounds__(kGB) void ground_fit_kernel(
    const float4* __restrict__ pts, const int32_t* __restrict__ counts,
    const int32_t* __restrict__ offsets,
    const int32_t* __restrict__ by_idx, const int32_t* __restrict__ by_z, GroundCfgDev cfg,
    int4* __restrict__ info, double* __restrict__ model, float4* __restrict__ features,
    int32_t* __restrict__ kept_index, int32_t* __restrict__ kept_offsets) {
  __shared__ float tile[3][kCentroidTile];
  __shared__ double red[(kGB / 64) * 6];
  __shared__ float cen[3];
  const int c = blockIdx.x;
  int off, s;
  if (offsets) {
    off = offsets[c];
    s = offsets[c + 1] - off;
  } else {
    off = 0;
    for (int i = 0; i < c; ++i) off += counts[i];
    s = counts[c];
  }
  int k = s;
  const int32_t* list = by_idx + off;
  if (cfg.retain_num < (double)s) {
    k = (int)((double)s / cfg.retain_num);
    k = min(k, s);
    list = by_z + off;
  }
  const bool valid = k >= cfg.nf && k > 0;
  double* out = model + 8 * (size_t)c;
  if (kept_index) {
    // the cell's kept points, in the order used (slots past k hold -1)
    for (int j = threadIdx.x; j < s; j += kGB) kept_index[off + j] = j < k ? list[j] : -1;
    if (threadIdx.x == 0) {
      kept_offsets[c] = off;
      if (c == cfg.n_cells - 1) kept_offsets[cfg.n_cells] = off + s;
    }
  }
  if (!valid) {
    if (threadIdx.x == 0) {
      info[c] = make_int4(0, k, s, 0);
      for (int i = 0; i < 8; ++i) out[i] = 0.0;
    }
    return;
  }
  // centroid: sequential float sums in list order (the reference's
  // computeCentroid), staged through LDS tiles
  float sx = 0.f, sy = 0.f, sz = 0.f;
  for (int t0 = 0; t0 < k; t0 += kCentroidTile) {
    const int tn = min(kCentroidTile, k - t0);
    __syncthreads();
    for (int j = threadIdx.x; j < tn; j += kGB) {
      const float4 p = pts[list[t0 + j]];
      tile[0][j] = p.x;
      tile[1][j] = p.y;
      tile[2][j] = p.z;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int j = 0; j < tn; ++j) {
        sx = sx + tile[0][j];
        sy = sy + tile[1][j];
        sz = sz + tile[2][j];
      }
    }
  }
  if (threadIdx.x == 0) {
    cen[0] = (float)((double)sx / (double)k);
    cen[1] = (float)((double)sy / (double)k);
    cen[2] = (float)((double)sz / (double)k);
  }
  __syncthreads();
  const float cx = cen[0], cy = cen[1], cz = cen[2];
  double S[6] = {0, 0, 0, 0, 0, 0};
  for (int j = threadIdx.x; j < k; j += kGB) {
    const float4 p = pts[list[j]];
    const double x = (double)(p.x - cx), y = (double)(p.y - cy), z = (double)(p.z - cz);
    S[0] += x * x;
    S[1] += x * y;
    S[2] += x * z;
    S[3] += y * y;
    S[4] += y * z;
    S[5] += z * z;
  }
  gsum<6>(S, red);
  for (int j = threadIdx.x; j < cfg.nf; j += kGB) features[(size_t)c * cfg.nf + j] = pts[list[j]];
  if (threadIdx.x != 0) return;
  double w[3], V[3][3], nrm[3];
  sl_eig3(S, w, V);
  const int kmin = sl_argmin3(w);
  nrm[0] = kmin == 0 ? V[0][0] : (kmin == 1 ? V[0][1] : V[0][2]);
  nrm[1] = kmin == 0 ? V[1][0] : (kmin == 1 ? V[1][1] : V[1][2]);
  nrm[2] = kmin == 0 ? V[2][0] : (kmin == 1 ? V[2][1] : V[2][2]);
  // orientation: upward (z > 0); a horizontal normal falls back to the
  // largest-|component|-positive rule
  if (nrm[2] < 0.0) {
    nrm[0] = -nrm[0];
    nrm[1] = -nrm[1];
    nrm[2] = -nrm[2];
  }
  if (nrm[2] == 0.0) sl_sign_normalise(nrm);
  const double d = -(nrm[0] * (double)cx + nrm[1] * (double)cy + nrm[2] * (double)cz);
  out[0] = nrm[0];
  out[1] = nrm[1];
  out[2] = nrm[2];
  out[3] = d;
  out[4] = (double)cx;
  out[5] = (double)cy;
  out[6] = (double)cz;
  out[7] = 0.0;
  info[c] = make_int4(1, k, s, 0);
}

// ---------------------------------------------------------------------------
// batched plane RANSAC
// ---------------------------------------------------------------------------
constexpr int kPlB = 256;
constexpr int kPlW = kPlB / 64;
constexpr int kPlHB = 8;
constexpr int kPlTile = 2048;

struct PlFrame {
  float m[3];
  int32_t n;
};

struct PlHyp {
  float nx, ny, nz, d;  // unit normal and offset in the local frame (p - m)
  int valid;
};

__device__ inline PlHyp plane_hyp(const float a[3], const float b[3], const float c[3]) {
  PlHyp h;
  const float ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
  const float vx = c[0] - a[0], vy = c[1] - a[1], vz = c[2] - a[2];
  float nx = uy * vz - uz * vy;
  float ny = uz * vx - ux * vz;
  float nz = ux * vy - uy * vx;
  const float len = sqrtf(nx * nx + ny * ny + nz * nz);
  h.valid = len > 0.0f ? 1 : 0;  // false for a degenerate triangle or NaN
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
    h.d = INFINITY;  // |t| < thr never holds
  }
  return h;
}

__device__ inline PlHyp plane_hyp_for(const float* __restrict__ P, int n, int set, int hidx,
                                      const PlFrame& f, uint32_t k0, uint32_t k1, uint32_t tag,
                                      int32_t idx[3]) {
  SlU4 ctr = {(uint32_t)hidx, (uint32_t)set, tag, 0u};
  const SlU4 x = sl_philox4x32_10(ctr, k0, k1);
  sl_draw3(x.x, x.y, x.z, (uint32_t)n, &idx[0], &idx[1], &idx[2]);
  float q[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int c = 0; c < 3; ++c) q[j][c] = P[3 * (size_t)idx[j] + c] - f.m[c];
  return plane_hyp(q[0], q[1], q[2]);
}

__device__ inline bool plane_inlier(float x, float y, float z, const PlHyp& h, float thr) {
  const float t = __builtin_fmaf(h.nx, x, __builtin_fmaf(h.ny, y, __builtin_fmaf(h.nz, z, h.d)));
  return fabsf(t) < thr;
}

__global__ __launch_bounds__(kPlB) void plane_seed_kernel(const float* __restrict__ xyz,
                                                          const int32_t* __restrict__ offsets,
                                                          int n_sets, PlFrame* __restrict__ frames) {
  __shared__ float red[kPlW][6];
  const int c = blockIdx.x;
  if (c >= n_sets) return;
  const int off = offsets[c];
  const int n = offsets[c + 1] - off;
  const float* P = xyz + 3 * (size_t)off;
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = threadIdx.x; i < n; i += kPlB)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float q = P[3 * (size_t)i + k];
      mn[k] = fminf(mn[k], q);
      mx[k] = fmaxf(mx[k], q);
    }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      mn[k] = fminf(mn[k], __shfl_xor(mn[k], o, 64));
      mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], o, 64));
    }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    for (int k = 0; k < 3; ++k) {
      red[wave][k] = mn[k];
      red[wave][3 + k] = mx[k];
    }
  __syncthreads();
  if (threadIdx.x != 0) return;
  PlFrame f;
  for (int k = 0; k < 3; ++k) {
    float a = red[0][k], b = red[0][3 + k];
    for (int w = 1; w < kPlW; ++w) {
      a = fminf(a, red[w][k]);
      b = fmaxf(b, red[w][3 + k]);
    }
    f.m[k] = n > 0 ? 0.5f * (a + b) : 0.0f;
  }
  f.n = n;
  frames[c] = f;
}

__global__ __launch_bounds__(kPlB) void plane_score_kernel(
    const float* __restrict__ xyz, const int32_t* __restrict__ offsets,
    const PlFrame* __restrict__ frames, const int4* __restrict__ work, int n_hyp, int hcap,
    float thr, uint32_t k0, uint32_t k1, uint32_t tag, int32_t* __restrict__ counts,
    int32_t* __restrict__ samples) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float4* hp = reinterpret_cast<float4*>(smem_raw);
  int* cnt = reinterpret_cast<int*>(hp + hcap);
  float* sx = reinterpret_cast<float*>(cnt + hcap);
  float* sy = sx + kPlTile;
  float* sz = sy + kPlTile;
  const int4 w = work[blockIdx.x];
  const int c = w.x, h0 = w.y, hc = w.z;
  const int off = offsets[c];
  const int n = offsets[c + 1] - off;
  const float* P = xyz + 3 * (size_t)off;
  const PlFrame f = frames[c];
  for (int t = threadIdx.x; t < hcap; t += kPlB) {
    float4 prm = make_float4(0.f, 0.f, 0.f, INFINITY);
    if (t < hc) {
      int32_t idx[3] = {-1, -1, -1};
      if (n >= 3) {
        const PlHyp h = plane_hyp_for(P, n, c, h0 + t, f, k0, k1, tag, idx);
        prm = make_float4(h.nx, h.ny, h.nz, h.d);
      }
      if (samples) {
        int32_t* o = samples + 3 * ((size_t)c * n_hyp + h0 + t);
        o[0] = idx[0];
        o[1] = idx[1];
        o[2] = idx[2];
      }
    }
    hp[t] = prm;
    cnt[t] = 0;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int hpw = hcap / kPlW;
  for (int t0 = 0; t0 < n; t0 += kPlTile) {
    const int tn = min(kPlTile, n - t0);
    const int tn64 = (tn + 63) & ~63;
    __syncthreads();
    for (int i = threadIdx.x; i < tn64; i += kPlB) {
      float x = NAN, y = 0.f, z = 0.f;
      if (i < tn) {
        const float* q = P + 3 * (size_t)(t0 + i);
        x = q[0] - f.m[0];
        y = q[1] - f.m[1];
        z = q[2] - f.m[2];
      }
      sx[i] = x;
      sy[i] = y;
      sz[i] = z;
    }
    __syncthreads();
    for (int hb = 0; hb < hpw; hb += kPlHB) {
      const int hbase = wave * hpw + hb;
      if (hbase >= hc) break;
      float NX[kPlHB], NY[kPlHB], NZ[kPlHB], D[kPlHB];
      int cc[kPlHB];
#pragma unroll
      for (int j = 0; j < kPlHB; ++j) {
        const float4 q = hp[hbase + j];
        NX[j] = q.x;
        NY[j] = q.y;
        NZ[j] = q.z;
        D[j] = q.w;
        cc[j] = 0;
      }
      for (int i0 = 0; i0 < tn64; i0 += 64) {
        const int i = i0 + lane;
        const float x = sx[i], y = sy[i], z = sz[i];
#pragma unroll
        for (int j = 0; j < kPlHB; ++j) {
          const float t = __builtin_fmaf(NX[j], x, __builtin_fmaf(NY[j], y,
                                                                  __builtin_fmaf(NZ[j], z, D[j])));
          cc[j] += __builtin_popcountll(__builtin_amdgcn_ballot_w64(fabsf(t) < thr));
        }
      }
      if (lane == 0)
#pragma unroll
        for (int j = 0; j < kPlHB; ++j) cnt[hbase + j] += cc[j];
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < hc; t += kPlB) counts[(size_t)c * n_hyp + h0 + t] = cnt[t];
}

// model: plane[4] (unit normal, z >= 0 orientation), centroid[3], inliers of
// the refitted plane
__global__ __launch_bounds__(kPlB) void plane_refit_kernel(
    const float* __restrict__ xyz, const int32_t* __restrict__ offsets,
    const PlFrame* __restrict__ frames, const int32_t* __restrict__ counts, int n_hyp, float thr,
    uint32_t k0, uint32_t k1, uint32_t tag, uint32_t* __restrict__ mask,
    int32_t* __restrict__ best_out, double* __restrict__ model) {
  __shared__ double red[kPlW * 6];
  __shared__ long long redk[kPlW];
  __shared__ PlHyp best_h;
  const int c = blockIdx.x;
  const int off = offsets[c];
  const int n = offsets[c + 1] - off;
  const float* P = xyz + 3 * (size_t)off;
  const PlFrame f = frames[c];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double* out = model + 8 * (size_t)c;
  long long key = -1;
  for (int h = threadIdx.x; h < n_hyp; h += kPlB) {
    const long long kk = ((long long)counts[(size_t)c * n_hyp + h] << 32) |
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
  for (int w = 1; w < kPlW; ++w) key = redk[w] > key ? redk[w] : key;
  const int best = (n >= 3) ? (int)(0xFFFFFFFFu - (uint32_t)(key & 0xFFFFFFFFll)) : -1;
  if (threadIdx.x == 0) {
    if (best >= 0) {
      int32_t idx[3];
      best_h = plane_hyp_for(P, n, c, best, f, k0, k1, tag, idx);
    } else {
      best_h.valid = 0;
      best_h.nx = best_h.ny = best_h.nz = 0.f;
      best_h.d = INFINITY;
    }
    if (best_out) best_out[c] = best;
  }
  __syncthreads();
  const PlHyp h = best_h;
  // mask + centroid (FP64, relative to m)
  double acc[4] = {0, 0, 0, 0};
  const int gend = off + n;
  for (int gb = (off & ~63); gb < gend; gb += kPlB) {
    const int g = gb + (int)threadIdx.x;
    bool in = false;
    double d3[3] = {0, 0, 0};
    if (g >= off && g < gend) {
      const float* q = xyz + 3 * (size_t)g;
      const float x = q[0] - f.m[0], y = q[1] - f.m[1], z = q[2] - f.m[2];
      in = plane_inlier(x, y, z, h, thr);
      if (in) {
        d3[0] = (double)q[0] - (double)f.m[0];
        d3[1] = (double)q[1] - (double)f.m[1];
        d3[2] = (double)q[2] - (double)f.m[2];
      }
    }
    const uint64_t b = __builtin_amdgcn_ballot_w64(in);
    if (mask && lane == 0 && b) {
      const int w0 = (gb + wave * 64) >> 5;
      if ((uint32_t)b) atomicOr(mask + w0, (uint32_t)b);
      if ((uint32_t)(b >> 32)) atomicOr(mask + w0 + 1, (uint32_t)(b >> 32));
    }
    if (in) {
      acc[0] += d3[0];
      acc[1] += d3[1];
      acc[2] += d3[2];
      acc[3] += 1.0;
    }
  }
  {
    double a4[4] = {acc[0], acc[1], acc[2], acc[3]};
#pragma unroll
    for (int k = 0; k < 4; ++k) a4[k] = sl_wave_sum(a4[k]);
    __syncthreads();
    if (lane == 0)
      for (int k = 0; k < 4; ++k) red[wave * 4 + k] = a4[k];
    __syncthreads();
    for (int k = 0; k < 4; ++k) {
      double t = red[k];
      for (int w = 1; w < kPlW; ++w) t += red[w * 4 + k];
      acc[k] = t;
    }
    __syncthreads();
  }
  const double k = acc[3];
  if (k < 3.0 || !h.valid) {
    if (threadIdx.x == 0) {
      for (int i = 0; i < 7; ++i) out[i] = 0.0;
      out[7] = k;
    }
    return;
  }
  const double mu[3] = {acc[0] / k, acc[1] / k, acc[2] / k};
  double S[6] = {0, 0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < n; i += kPlB) {
    const float* q = P + 3 * (size_t)i;
    if (!plane_inlier(q[0] - f.m[0], q[1] - f.m[1], q[2] - f.m[2], h, thr)) continue;
    const double x = ((double)q[0] - (double)f.m[0]) - mu[0];
    const double y = ((double)q[1] - (double)f.m[1]) - mu[1];
    const double z = ((double)q[2] - (double)f.m[2]) - mu[2];
    S[0] += x * x;
    S[1] += x * y;
    S[2] += x * z;
    S[3] += y * y;
    S[4] += y * z;
    S[5] += z * z;
  }
  {
#pragma unroll
    for (int kk = 0; kk < 6; ++kk) S[kk] = sl_wave_sum(S[kk]);
    __syncthreads();
    if (lane == 0)
      for (int kk = 0; kk < 6; ++kk) red[wave * 6 + kk] = S[kk];
    __syncthreads();
    for (int kk = 0; kk < 6; ++kk) {
      double t = red[kk];
      for (int w = 1; w < kPlW; ++w) t += red[w * 6 + kk];
      S[kk] = t;
    }
  }
  double w3[3], V[3][3], nrm[3];
  sl_eig3(S, w3, V);
  const int kmin = sl_argmin3(w3);
  nrm[0] = kmin == 0 ? V[0][0] : (kmin == 1 ? V[0][1] : V[0][2]);
  nrm[1] = kmin == 0 ? V[1][0] : (kmin == 1 ? V[1][1] : V[1][2]);
  nrm[2] = kmin == 0 ? V[2][0] : (kmin == 1 ? V[2][1] : V[2][2]);
  if (nrm[2] < 0.0) {
    nrm[0] = -nrm[0];
    nrm[1] = -nrm[1];
    nrm[2] = -nrm[2];
  }
  if (nrm[2] == 0.0) sl_sign_normalise(nrm);
  const double cxyz[3] = {(double)f.m[0] + mu[0], (double)f.m[1] + mu[1], (double)f.m[2] + mu[2]};
  const double dd = -(nrm[0] * cxyz[0] + nrm[1] * cxyz[1] + nrm[2] * cxyz[2]);
  // inliers of the refitted plane (FP64)
  double cnt[1] = {0.0};
  for (int i = threadIdx.x; i < n; i += kPlB) {
    const float* q = P + 3 * (size_t)i;
    const double t = nrm[0] * (double)q[0] + nrm[1] * (double)q[1] + nrm[2] * (double)q[2] + dd;
    if (fabs(t) < (double)thr) cnt[0] += 1.0;
  }
  cnt[0] = sl_wave_sum(cnt[0]);
  __syncthreads();
  if (lane == 0) red[wave] = cnt[0];
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0];
    for (int w = 1; w < kPlW; ++w) t += red[w];
    out[0] = nrm[0];
    out[1] = nrm[1];
    out[2] = nrm[2];
    out[3] = dd;
    out[4] = cxyz[0];
    out[5] = cxyz[1];
    out[6] = cxyz[2];
    out[7] = t;
  }
}

}  // namespace

struct sloam_plane_plan {
  sloam_ctx* ctx = nullptr;
  sloam_plane_fit_cfg cfg{};
  int32_t n_sets = 0;
  int64_t n_points = 0;
  int hcap = 0;
  int n_work = 0;
  int4* d_work = nullptr;
  PlFrame* d_frames = nullptr;
  int32_t* d_counts = nullptr;
};

static int pl_round_up(int x, int m) { return (x + m - 1) / m * m; }

extern "C" {

int sloam_ground_planes(sloam_ctx* ctx, const float* d_points, int32_t capacity,
                        const int32_t* d_count, const sloam_ground_cfg* cfg, int32_t* d_cell_info,
                        double* d_cell_model, float* d_features, int32_t* d_kept_offsets,
                        int32_t* d_kept_index) {
  if (!ctx || !cfg || !d_cell_info || !d_cell_model || capacity < 0) return SLOAM_ERR_INVALID_ARG;
  if (cfg->radii_bins < 1 || cfg->theta_bins < 1 || !(cfg->retain_thresh > 0.0) ||
      cfg->num_features < 0)
    return SLOAM_ERR_INVALID_ARG;
  const int n_cells = cfg->radii_bins * cfg->theta_bins;
  if (n_cells > 65535) return SLOAM_ERR_TOO_LARGE;
  if (cfg->num_features > 0 && !d_features) return SLOAM_ERR_INVALID_ARG;
  if (capacity > 0 && !d_points) return SLOAM_ERR_INVALID_ARG;
  if ((d_kept_index != nullptr) != (d_kept_offsets != nullptr)) return SLOAM_ERR_INVALID_ARG;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  GroundCfgDev dc;
  dc.R = cfg->radii_bins;
  dc.T = cfg->theta_bins;
  dc.n_cells = n_cells;
  dc.nf = cfg->num_features;
  dc.min_d = cfg->min_dist;
  dc.max_d = cfg->max_dist;
  dc.retain_num = 1.0 / cfg->retain_thresh;
  dc.ox = cfg->origin[0];
  dc.oy = cfg->origin[1];
  const int cap = std::max(capacity, 1);
  // scratch: 2 key arrays + a sort buffer each, 2 value arrays + a sort
  // buffer each, counts, offsets, sort histogram
  const size_t kb = sizeof(unsigned long long) * (size_t)cap;
  const size_t vb = sizeof(int32_t) * (size_t)cap;
  const size_t cb = sizeof(int32_t) * (size_t)(n_cells + 1);
  const size_t hb = sizeof(int32_t) * sloam_radix_hist_ints(cap);
  auto al = [](size_t x) { return (x + 255) / 256 * 256; };
  const size_t need = 4 * al(kb) + 4 * al(vb) + 2 * al(cb) + al(hb);
  void* base = nullptr;
  int st = sloam_ctx_scratch(ctx, 1, need, &base);
  if (st) return st;
  char* q = (char*)base;
  auto take = [&](size_t b) {
    char* r = q;
    q += al(b);
    return (void*)r;
  };
  unsigned long long* kA = (unsigned long long*)take(kb);
  unsigned long long* kB = (unsigned long long*)take(kb);
  unsigned long long* kT = (unsigned long long*)take(kb);
  unsigned long long* kT2 = (unsigned long long*)take(kb);
  int32_t* vA = (int32_t*)take(vb);
  int32_t* vB = (int32_t*)take(vb);
  int32_t* vT = (int32_t*)take(vb);
  int32_t* vT2 = (int32_t*)take(vb);
  // the two sorted value arrays (where each sort's last pass leaves them)
  int32_t* by_idx = vA;
  int32_t* by_z = vB;
  int32_t* counts = (int32_t*)take(cb);
  int32_t* offs = (int32_t*)take(cb);
  int32_t* hist = (int32_t*)take(hb);
  st = sloam_fill_u32(ctx, (uint32_t*)counts, (size_t)n_cells, 0u);
  if (st) return st;
  {
    SloamKernelScope ks(ctx, "ground_bin");
    if (capacity > 0)
      hipLaunchKernelGGL(ground_bin_kernel, dim3((capacity + kBinPts - 1) / kBinPts), dim3(kGB), 0,
                         ctx->stream, reinterpret_cast<const float4*>(d_points), capacity,
                         d_count, dc, kA, kB, vA, vB, counts);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  if (capacity > 0) {
    SloamKernelScope ks(ctx, "ground_sort");
    // cell ids 0..n_cells (n_cells: outside the grid) in cell_bits bits; the
    // (cell, z) keys hold the orderable z in the low 32 bits
    int cell_bits = 1;
    while ((1 << cell_bits) <= n_cells) ++cell_bits;
    unsigned long long* sorted_keys = nullptr;
    st = sloam_radix_sort(ctx, kA, vA, kT, vT, capacity, 0, cell_bits, d_count, hist, &sorted_keys,
                          &by_idx);
    if (st) return st;
    st = sloam_radix_sort(ctx, kB, vB, kT2, vT2, capacity, 0, 32 + cell_bits, d_count, hist,
                          &sorted_keys,
                          &by_z);
    if (st) return st;
  }
  {
    SloamKernelScope ks(ctx, "ground_fit");
    // a few hundred cells: each fit workgroup adds the counts before its cell
    const bool prefix = n_cells > kGB;
    if (prefix)
      hipLaunchKernelGGL(ground_offsets_kernel, dim3(1), dim3(kGB), 0, ctx->stream, counts,
                         n_cells, offs);
    hipLaunchKernelGGL(ground_fit_kernel, dim3(n_cells), dim3(kGB), 0, ctx->stream,
                       reinterpret_cast<const float4*>(d_points), counts,
                       prefix ? offs : nullptr, by_idx, by_z, dc,
                       reinterpret_cast<int4*>(d_cell_info), d_cell_model,
                       reinterpret_cast<float4*>(d_features), d_kept_index, d_kept_offsets);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}

int sloam_plane_plan_create(sloam_ctx* ctx, const int32_t* h_offsets, int32_t n_sets,
                            const sloam_plane_fit_cfg* cfg, sloam_plane_plan** out) {
  if (!ctx || !h_offsets || !cfg || !out || n_sets < 0) return SLOAM_ERR_INVALID_ARG;
  *out = nullptr;
  if (cfg->n_hyp < 1 || cfg->n_hyp > (1 << 20) || !(cfg->thr > 0.0f)) return SLOAM_ERR_INVALID_ARG;
  if (h_offsets[0] != 0) return SLOAM_ERR_INVALID_ARG;
  for (int i = 0; i < n_sets; ++i)
    if (h_offsets[i + 1] < h_offsets[i]) return SLOAM_ERR_INVALID_ARG;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  sloam_plane_plan* p = new (std::nothrow) sloam_plane_plan();
  if (!p) return SLOAM_ERR_ALLOC;
  p->ctx = ctx;
  p->cfg = *cfg;
  p->n_sets = n_sets;
  p->n_points = h_offsets[n_sets];
  p->hcap = pl_round_up(std::min(256, pl_round_up(cfg->n_hyp, kPlW * kPlHB)), kPlW * kPlHB);
  std::vector<int> order(n_sets);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (h_offsets[a + 1] - h_offsets[a]) > (h_offsets[b + 1] - h_offsets[b]);
  });
  std::vector<int4> work;
  for (int c : order)
    for (int h0 = 0; h0 < cfg->n_hyp; h0 += p->hcap)
      work.push_back(make_int4(c, h0, std::min(p->hcap, cfg->n_hyp - h0), 0));
  p->n_work = (int)work.size();
  int st = SLOAM_OK;
  if (p->n_work > 0) {
    st = sloam_dev_alloc(ctx, (void**)&p->d_work, sizeof(int4) * work.size());
    if (!st) {
      hipError_t e = hipMemcpy(p->d_work, work.data(), sizeof(int4) * work.size(),
                               hipMemcpyHostToDevice);
      if (e != hipSuccess) st = sloam_record_hip_error(ctx, e, "hipMemcpy(plane work)");
    }
  }
  if (!st && n_sets > 0) st = sloam_dev_alloc(ctx, (void**)&p->d_frames, sizeof(PlFrame) * n_sets);
  if (!st && n_sets > 0)
    st = sloam_dev_alloc(ctx, (void**)&p->d_counts, sizeof(int32_t) * (size_t)n_sets * cfg->n_hyp);
  if (st) {
    sloam_plane_plan_destroy(p);
    return st;
  }
  *out = p;
  return SLOAM_OK;
}

void sloam_plane_plan_destroy(sloam_plane_plan* p) {
  if (!p) return;
  sloam_dev_free(p->d_work);
  sloam_dev_free(p->d_frames);
  sloam_dev_free(p->d_counts);
  delete p;
}

int sloam_plane_plan_exec(sloam_plane_plan* p, const float* d_xyz, const int32_t* d_offsets,
                          int32_t* d_counts, int32_t* d_samples, uint32_t* d_mask,
                          int32_t* d_best, double* d_model) {
  if (!p || !d_offsets || !d_model || (p->n_points > 0 && !d_xyz)) return SLOAM_ERR_INVALID_ARG;
  sloam_ctx* ctx = p->ctx;
  if (p->n_sets == 0) return SLOAM_OK;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  const uint32_t k0 = (uint32_t)p->cfg.seed, k1 = (uint32_t)(p->cfg.seed >> 32);
  const uint32_t tag = (uint32_t)p->cfg.stream_tag;
  int32_t* counts = d_counts ? d_counts : p->d_counts;
  if (d_mask) {
    int st = sloam_fill_u32(ctx, d_mask, (size_t)((p->n_points + 31) / 32), 0u);
    if (st) return st;
  }
  {
    SloamKernelScope ks(ctx, "plane_seed");
    hipLaunchKernelGGL(plane_seed_kernel, dim3(p->n_sets), dim3(kPlB), 0, ctx->stream, d_xyz,
                       d_offsets, p->n_sets, p->d_frames);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  if (p->n_work > 0) {
    const size_t lds = (size_t)p->hcap * (sizeof(float4) + sizeof(int)) + 3 * sizeof(float) * kPlTile;
    SloamKernelScope ks(ctx, "plane_score");
    hipLaunchKernelGGL(plane_score_kernel, dim3(p->n_work), dim3(kPlB), lds, ctx->stream, d_xyz,
                       d_offsets, p->d_frames, p->d_work, p->cfg.n_hyp, p->hcap, p->cfg.thr, k0,
                       k1, tag, counts, d_samples);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  {
    SloamKernelScope ks(ctx, "plane_refit");
    hipLaunchKernelGGL(plane_refit_kernel, dim3(p->n_sets), dim3(kPlB), 0, ctx->stream, d_xyz,
                       d_offsets, p->d_frames, counts, p->cfg.n_hyp, p->cfg.thr, k0, k1, tag,
                       d_mask, d_best, d_model);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}

}  // extern "C"
