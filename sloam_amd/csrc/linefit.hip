// Batched RANSAC line fit over trellis vertex medians (gfx950): the axis fit
// of the reference's Cylinder::computeModel
// (sloam/src/objects/cylinder.cpp:116-139, PCL SACSegmentation with
// SACMODEL_LINE / SAC_RANSAC, threshold 0.25, optimized coefficients).
//
// One wavefront per point set (tree), its points staged in LDS.  Lane 0 runs
// the sequential part of the sample-consensus loop exactly as the library
// defines it -- MT19937 seeded 12345 (state in LDS; the state after seeding
// and the first twist is the same for every set, so the host computes it once
// per seed and each workgroup copies it), rnd() = mt() >> 1, a persistent swap-shuffle of the
// index array, up to 1000 draws until the sample is good (all three
// coordinates differ), the adaptive iteration bound k = log(1 - p) /
// log(1 - w^2) -- while the whole wave scores each line with one ballot per
// 64 points.  The final fit refines the line through the inliers' float
// centroid along the largest eigenvector of their float scatter matrix
// (cross product of two rows of C - lambda I), then re-selects inliers.
#include "common.hpp"

#include <cfloat>
#include <cmath>
#include <mutex>

// MT19937 state after seeding with g_mt_seed and one twist (per device)
__device__ uint32_t g_mt_state[624];

namespace {

constexpr int kLineMax = 4096;  // points per set (LDS index array)
constexpr int kStage = 1024;    // sets up to this size are read from LDS

struct LineCfgDev {
  double sqr_thr;
  double log_p;
  int max_iter;
  int optimize;
  uint32_t seed;
};

__device__ inline uint32_t mt_next(uint32_t* __restrict__ mt, int* __restrict__ idx) {
  if (*idx >= 624) {
    for (int i = 0; i < 624; ++i) {
      const uint32_t y = (mt[i] & 0x80000000u) | (mt[i + 1 < 624 ? i + 1 : 0] & 0x7fffffffu);
      const int j = i + 397 < 624 ? i + 397 : i + 397 - 624;
      mt[i] = mt[j] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    *idx = 0;
  }
  uint32_t y = mt[(*idx)++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

struct LineEvalD {
  float p[3], d[3];
};

__device__ inline LineEvalD line_eval(const float c[6]) {
  LineEvalD e;
  const float l = sqrtf((c[3] * c[3] + c[4] * c[4]) + c[5] * c[5]);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    e.p[i] = c[i];
    e.d[i] = c[3 + i] / l;
  }
  return e;
}

__device__ inline float sqr_line_dist(const float* q, const LineEvalD& e) {
  const float ax = e.p[0] - q[0], ay = e.p[1] - q[1], az = e.p[2] - q[2];
  const float x = ay * e.d[2] - az * e.d[1];
  const float y = az * e.d[0] - ax * e.d[2];
  const float z = ax * e.d[1] - ay * e.d[0];
  return (x * x + y * y) + z * z;
}

__device__ inline int wave_count(const float* __restrict__ P, int n, const LineEvalD& e,
                                 double sqr_thr) {
  const int lane = threadIdx.x;
  int cnt = 0;
  for (int i0 = 0; i0 < n; i0 += 64) {
    const int i = i0 + lane;
    bool in = false;
    if (i < n) in = (double)sqr_line_dist(P + 3 * (size_t)i, e) < sqr_thr;
    cnt += __builtin_popcountll(__builtin_amdgcn_ballot_w64(in));
  }
  return cnt;
}

// stats: {ok, inliers, iterations, skipped}
__global__ __launch_bounds__(64) void line_ransac_kernel(const float* __restrict__ xyz,
                                                         const int32_t* __restrict__ offsets,
                                                         int n_sets,
                                                         const int32_t* __restrict__ d_n_sets,
                                                         LineCfgDev cfg,
                                                         float* __restrict__ coeff_out,
                                                         int32_t* __restrict__ stats_out,
                                                         uint8_t* __restrict__ inl_out) {
  __shared__ uint32_t mt[624];
  __shared__ int mt_idx;
  __shared__ int shuf[kLineMax];
  __shared__ int sel[3];
  __shared__ int in_list[kLineMax];
  __shared__ float model_s[6];
  __shared__ float stage[3 * kStage];
  const int s = blockIdx.x;
  if (s >= n_sets) return;
  if (d_n_sets && s >= *d_n_sets) return;
  const int lane = threadIdx.x;
  const int off = offsets[s];
  const int n = offsets[s + 1] - off;
  const float* P = xyz + 3 * (size_t)off;
  if (n > 0 && n <= kStage) {
    for (int i = lane; i < 3 * n; i += 64) stage[i] = P[i];
    P = stage;  // flat pointer into LDS
  }
  float* co = coeff_out + 6 * (size_t)s;
  int32_t* so = stats_out + 4 * (size_t)s;
  if (inl_out)
    for (int i = lane; i < n; i += 64) inl_out[off + i] = 0;
  if (n < 2 || n > kLineMax) {
    if (lane < 6) co[lane] = 0.f;
    if (lane < 4) so[lane] = 0;
    return;
  }
  for (int i = lane; i < n; i += 64) shuf[i] = i;
  for (int i = lane; i < 624; i += 64) mt[i] = g_mt_state[i];
  if (lane == 0) mt_idx = 0;
  __syncthreads();
  const double one_over = 1.0 / (double)n;
  double k = DBL_MAX;
  int iterations = 0, skipped = 0, best = 0;
  const int max_skip = cfg.max_iter * 10;
  bool have = false;
  float model[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  while (iterations < k && skipped < max_skip) {
    if (lane == 0) {
      int good = 0, s0 = -1, s1 = -1;
      int idx = mt_idx;
      for (int t = 0; t < 1000; ++t) {
        for (int i = 0; i < 2; ++i) {
          const int j = i + (int)((mt_next(mt, &idx) >> 1) % (uint32_t)(n - i));
          const int tmp = shuf[i];
          shuf[i] = shuf[j];
          shuf[j] = tmp;
        }
        s0 = shuf[0];
        s1 = shuf[1];
        const float* a = P + 3 * (size_t)s0;
        const float* b = P + 3 * (size_t)s1;
        if (a[0] != b[0] && a[1] != b[1] && a[2] != b[2]) {
          good = 1;
          break;
        }
      }
      mt_idx = idx;
      sel[0] = s0;
      sel[1] = s1;
      sel[2] = good;
    }
    __syncthreads();
    const int s0 = sel[0], s1 = sel[1], good = sel[2];
    __syncthreads();
    if (!good) break;
    const float* a = P + 3 * (size_t)s0;
    const float* b = P + 3 * (size_t)s1;
    if (fabsf(a[0] - b[0]) <= FLT_EPSILON && fabsf(a[1] - b[1]) <= FLT_EPSILON &&
        fabsf(a[2] - b[2]) <= FLT_EPSILON) {
      ++skipped;
      continue;
    }
    float c[6] = {a[0], a[1], a[2], b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const float nn = sqrtf((c[3] * c[3] + c[4] * c[4]) + c[5] * c[5]);
    c[3] = c[3] / nn;
    c[4] = c[4] / nn;
    c[5] = c[5] / nn;
    const int cnt = wave_count(P, n, line_eval(c), cfg.sqr_thr);
    if (cnt > best) {
      best = cnt;
#pragma unroll
      for (int i = 0; i < 6; ++i) model[i] = c[i];
      have = true;
      const double w = (double)best * one_over;
      double p_no = 1.0 - w * w;  // pow(w, sample size 2)
      p_no = fmax(DBL_EPSILON, p_no);
      p_no = fmin(1.0 - DBL_EPSILON, p_no);
      k = cfg.log_p / log(p_no);
    }
    ++iterations;
    if (iterations > cfg.max_iter) break;
  }
  if (!have) {
    if (lane < 6) co[lane] = 0.f;
    if (lane == 0) {
      so[0] = 0;
      so[1] = 0;
      so[2] = iterations;
      so[3] = skipped;
    }
    return;
  }
  // inliers of the best line, in index order (LDS list)
  int n_in = 0;
  {
    const LineEvalD e = line_eval(model);
    for (int i0 = 0; i0 < n; i0 += 64) {
      const int i = i0 + lane;
      bool in = false;
      if (i < n) in = (double)sqr_line_dist(P + 3 * (size_t)i, e) < cfg.sqr_thr;
      const uint64_t bits = __builtin_amdgcn_ballot_w64(in);
      if (in) in_list[n_in + __builtin_popcountll(bits & ((1ull << lane) - 1ull))] = i;
      n_in += __builtin_popcountll(bits);
    }
  }
  __syncthreads();
  if (cfg.optimize && n_in > 2) {
    if (lane == 0) {
      float cx = 0.f, cy = 0.f, cz = 0.f;
      for (int j = 0; j < n_in; ++j) {
        const float* q = P + 3 * (size_t)in_list[j];
        cx = cx + q[0];
        cy = cy + q[1];
        cz = cz + q[2];
      }
      const float m = (float)n_in;
      cx = cx / m;
      cy = cy / m;
      cz = cz / m;
      float C[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      for (int j = 0; j < n_in; ++j) {
        const float* q = P + 3 * (size_t)in_list[j];
        const float x = q[0] - cx, y = q[1] - cy, z = q[2] - cz;
        C[3] = C[3] + y * y;
        C[4] = C[4] + y * z;
        C[5] = C[5] + z * z;
        C[0] = C[0] + x * x;
        C[1] = C[1] + x * y;
        C[2] = C[2] + x * z;
      }
      const double Cd[6] = {C[0], C[1], C[2], C[3], C[4], C[5]};
      double w[3], V[3][3];
      sl_eig3(Cd, w, V);
      const int km = sl_argmax3(w);
      const float lam = (float)(km == 0 ? w[0] : (km == 1 ? w[1] : w[2]));
      float mx = 0.f;
      for (int i = 0; i < 6; ++i) mx = fmaxf(mx, fabsf(C[i]));
      if (mx <= FLT_MIN) mx = 1.0f;
      const float r0[3] = {C[0] / mx - lam / mx, C[1] / mx, C[2] / mx};
      const float r1[3] = {C[1] / mx, C[3] / mx - lam / mx, C[4] / mx};
      const float r2[3] = {C[2] / mx, C[4] / mx, C[5] / mx - lam / mx};
      float v1[3], v2[3], v3[3];
      v1[0] = r0[1] * r1[2] - r0[2] * r1[1];
      v1[1] = r0[2] * r1[0] - r0[0] * r1[2];
      v1[2] = r0[0] * r1[1] - r0[1] * r1[0];
      v2[0] = r0[1] * r2[2] - r0[2] * r2[1];
      v2[1] = r0[2] * r2[0] - r0[0] * r2[2];
      v2[2] = r0[0] * r2[1] - r0[1] * r2[0];
      v3[0] = r1[1] * r2[2] - r1[2] * r2[1];
      v3[1] = r1[2] * r2[0] - r1[0] * r2[2];
      v3[2] = r1[0] * r2[1] - r1[1] * r2[0];
      const float l1 = (v1[0] * v1[0] + v1[1] * v1[1]) + v1[2] * v1[2];
      const float l2 = (v2[0] * v2[0] + v2[1] * v2[1]) + v2[2] * v2[2];
      const float l3 = (v3[0] * v3[0] + v3[1] * v3[1]) + v3[2] * v3[2];
      float vx = v3[0], vy = v3[1], vz = v3[2], l = l3;
      if (l1 >= l2 && l1 >= l3) {
        vx = v1[0];
        vy = v1[1];
        vz = v1[2];
        l = l1;
      } else if (l2 >= l1 && l2 >= l3) {
        vx = v2[0];
        vy = v2[1];
        vz = v2[2];
        l = l2;
      }
      const float sl = sqrtf(l);
      model_s[0] = cx;
      model_s[1] = cy;
      model_s[2] = cz;
      model_s[3] = vx / sl;
      model_s[4] = vy / sl;
      model_s[5] = vz / sl;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 6; ++i) model[i] = model_s[i];
    n_in = wave_count(P, n, line_eval(model), cfg.sqr_thr);
  }
  if (inl_out) {
    const LineEvalD e = line_eval(model);
    for (int i = lane; i < n; i += 64)
      inl_out[off + i] = (double)sqr_line_dist(P + 3 * (size_t)i, e) < cfg.sqr_thr ? 1 : 0;
  }
  if (lane < 6) co[lane] = model[lane];
  if (lane == 0) {
    so[0] = 1;
    so[1] = n_in;
    so[2] = iterations;
    so[3] = skipped;
  }
}

}  // namespace

int sloam_fit_lines_dev(sloam_ctx* ctx, const float* d_xyz, const int32_t* d_offsets,
                        int32_t n_sets, const int32_t* d_n_sets, const sloam_line_fit_cfg* cfg,
                        float* d_coeff, int32_t* d_stats, uint8_t* d_inliers) {
  if (!ctx || !cfg || n_sets < 0) return SLOAM_ERR_INVALID_ARG;
  if (n_sets == 0) return SLOAM_OK;
  if (!d_xyz || !d_offsets || !d_coeff || !d_stats) return SLOAM_ERR_INVALID_ARG;
  if (!(cfg->threshold > 0.0) || cfg->max_iterations < 0 || !(cfg->probability > 0.0) ||
      !(cfg->probability < 1.0))
    return SLOAM_ERR_INVALID_ARG;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  LineCfgDev dc;
  dc.sqr_thr = cfg->threshold * cfg->threshold;
  dc.log_p = log(1.0 - cfg->probability);
  dc.max_iter = cfg->max_iterations;
  dc.optimize = cfg->optimize;
  dc.seed = cfg->seed;
  {
    // the seeded and twisted generator state, uploaded when the seed changes
    static std::mutex mu;
    static uint32_t host_state[64][624];
    static int64_t host_seed[64];
    static bool init = false;
    std::lock_guard<std::mutex> lock(mu);
    if (!init) {
      for (int d = 0; d < 64; ++d) host_seed[d] = -1;
      init = true;
    }
    if (ctx->device < 0 || ctx->device >= 64) return SLOAM_ERR_INVALID_ARG;
    if (host_seed[ctx->device] != (int64_t)cfg->seed) {
      // an earlier upload from this buffer may still be in flight
      SLOAM_HIP(ctx, hipDeviceSynchronize());
      uint32_t* mt = host_state[ctx->device];
      mt[0] = cfg->seed;
      for (int i = 1; i < 624; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
      for (int i = 0; i < 624; ++i) {
        const uint32_t y = (mt[i] & 0x80000000u) | (mt[i + 1 < 624 ? i + 1 : 0] & 0x7fffffffu);
        const int j = i + 397 < 624 ? i + 397 : i + 397 - 624;
        mt[i] = mt[j] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
      }
      SLOAM_HIP(ctx, hipMemcpyToSymbolAsync(HIP_SYMBOL(g_mt_state), mt, sizeof(uint32_t) * 624, 0,
                                            hipMemcpyHostToDevice, ctx->stream));
      host_seed[ctx->device] = (int64_t)cfg->seed;
    }
  }
  {
    SloamKernelScope ks(ctx, "line_ransac");
    hipLaunchKernelGGL(line_ransac_kernel, dim3(n_sets), dim3(64), 0, ctx->stream, d_xyz,
                       d_offsets, n_sets, d_n_sets, dc, d_coeff, d_stats, d_inliers);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}

extern "C" {

int sloam_fit_lines(sloam_ctx* ctx, const float* d_xyz, const int32_t* d_offsets, int32_t n_sets,
                    const sloam_line_fit_cfg* cfg, float* d_coeff, int32_t* d_stats,
                    uint8_t* d_inliers) {
  return sloam_fit_lines_dev(ctx, d_xyz, d_offsets, n_sets, nullptr, cfg, d_coeff, d_stats,
                             d_inliers);
}

}  // extern "C"
