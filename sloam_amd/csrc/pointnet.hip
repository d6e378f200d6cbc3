// Learned cylinder proposal for gfx950: a PointNet over each cluster, BF16
// weights on MFMA with FP32 accumulation.
//
//   pn_prep     one workgroup per cluster: up to 512 points at indices
//               floor(j n / m), shifted by their bounding-box centre, x and y
//               scaled by 1 / (larger horizontal half extent), z by
//               1 / (vertical half extent); float4 (w = 1 real, 0 padding)
//   pn_fwd      one workgroup (8 waves) per cluster, 128 points at a time:
//               layer 1 (3 -> 64, bias as a fourth input of 1) on
//               v_mfma_f32_16x16x4_f32, layers 2 (64 -> 128) and
//               3 (128 -> 256) on v_mfma_f32_16x16x32_bf16 with the BF16
//               activations in LDS and each wave's weight fragments in VGPRs
//               (wave w owns output channels 16w.. of layer 2, computed
//               transposed so that a lane's four results are consecutive
//               channels of one point and leave in one 8-byte store, and
//               columns 32w.. of layer 3), masked max-pool with the first
//               maximising point kept per channel; persistent: a fixed grid
//               loops over the clusters with the weight fragments loaded once
//   pn_head     256 -> 128 -> 7 on the VALU for 8 clusters per workgroup
//   pn_bwd      one workgroup per cluster: loss, head gradients, and the
//               max-pool gradient routed to its argmax points only (at most
//               256 of them), back through layers 3, 2, 1 from the
//               activations the forward pass saved; the layer-2 products
//               (dW2 = dz2' H1 and dz1 = dz2 W2) on v_mfma_f32_16x16x32_bf16
//               over tiles of 32 active points, dz2 entering as two BF16
//               parts
//   pn_reduce   batch mean of the per-cluster gradients (fixed order)
//   pn_adam     Adam update of the FP32 master weights
#include "common.hpp"

#include <algorithm>
#include <cmath>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kP = SLOAM_PN_POINTS;  // points per cluster
constexpr int kC1 = 64, kC2 = 128, kC3 = 256, kC4 = 128, kOut = 7;
constexpr int kChunk = 128;          // points per pass of the forward kernel
constexpr int kS1 = kC1 + 8;         // LDS row strides (bf16), 16-byte multiples
constexpr int kS2 = kC2 + 8;
constexpr int kFB = 256;             // backward / prep workgroup
constexpr int kFW = 512;             // forward workgroup (8 waves)
constexpr int kHC = 8;               // clusters per head workgroup

// parameter offsets (weight [out][in], then bias)
constexpr int kW1 = 0, kB1 = kW1 + kC1 * 3;
constexpr int kW2 = kB1 + kC1, kB2 = kW2 + kC2 * kC1;
constexpr int kW3 = kB2 + kC2, kB3 = kW3 + kC3 * kC2;
constexpr int kW4 = kB3 + kC3, kB4 = kW4 + kC4 * kC3;
constexpr int kW5 = kB4 + kC4, kB5 = kW5 + kOut * kC4;
constexpr int kNP = kB5 + kOut;
static_assert(kW4 % 4 == 0 && kW5 % 4 == 0 && kC3 % 128 == 0 && kC4 % 64 == 0 && kFB == 2 * kC4,
              "head: weight half-rows read as batches of 16 float4 by 2 x kC4 threads");
static_assert(kNP == SLOAM_PN_PARAMS, "parameter layout");

// saved activations per cluster (floats): f[256], argmax[256] (int bits),
// h4[128], out[8]
constexpr int kAF = 0, kAI = kAF + kC3, kAH = kAI + kC3, kAO = kAH + kC4, kAct = kAO + 8;

// normalised p' = ((x - cx) sh, (y - cy) sh, (z - cz) sz)
struct PnMeta {
  float cx, cy, cz, sh, sz;
  int m, n, pad;
};

struct LossCfg {
  float wa, wc, wr, inv_batch;
};

// round to nearest even
__device__ inline unsigned short bf_bits(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float bf_round(float x) { return __uint_as_float((unsigned)bf_bits(x) << 16); }
// the same rounding by the gfx950 conversion instruction (v_cvt_pk_bf16_f32,
// round to nearest even) for the forward pass's activations
__device__ inline unsigned short bf_bits_hw(float x) {
  return __builtin_bit_cast(unsigned short, (__bf16)x);
}
__device__ inline float bf_val(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

// 8 consecutive FP32 weights -> BF16 fragment
__device__ inline bf16x8 frag_from_f32(const float* p) {
  const float4 a = *reinterpret_cast<const float4*>(p);
  const float4 b = *reinterpret_cast<const float4*>(p + 4);
  u32x4 u;
  u.x = (unsigned)bf_bits(a.x) | ((unsigned)bf_bits(a.y) << 16);
  u.y = (unsigned)bf_bits(a.z) | ((unsigned)bf_bits(a.w) << 16);
  u.z = (unsigned)bf_bits(b.x) | ((unsigned)bf_bits(b.y) << 16);
  u.w = (unsigned)bf_bits(b.z) | ((unsigned)bf_bits(b.w) << 16);
  return __builtin_bit_cast(bf16x8, u);
}

// 8 consecutive BF16 activations from LDS (16-byte aligned)
__device__ inline bf16x8 frag_from_lds(const unsigned short* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(p));
}

// ---------------------------------------------------------------------------
// d_n (nullable): the cluster count in device memory, at most n_clusters
// (the grid's size); likewise in pn_fwd and pn_head
__device__ inline int pn_count(int n_clusters, const int32_t* d_n) {
  return d_n ? min(n_clusters, max(*d_n, 0)) : n_clusters;
}

__global__ __launch_bounds__(kFB) void pn_prep_kernel(const float* __restrict__ xyz,
                                                      const int32_t* __restrict__ offsets,
                                                      int n_clusters, const int32_t* __restrict__ d_n,
                                                      float4* __restrict__ X,
                                                      PnMeta* __restrict__ meta) {
  __shared__ float red[kFB / 64][6];
  const int c = blockIdx.x;
  if (c >= pn_count(n_clusters, d_n)) return;
  const int off = offsets[c];
  const int n = offsets[c + 1] - off;
  const int m = n < kP ? n : kP;
  const float* P = xyz + 3 * (size_t)off;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int j = threadIdx.x; j < m; j += kFB) {
    const int i = (int)(((long long)j * n) / m);
    for (int k = 0; k < 3; ++k) {
      const float q = P[3 * (size_t)i + k];
      lo[k] = fminf(lo[k], q);
      hi[k] = fmaxf(hi[k], q);
    }
  }
  for (int o = 32; o > 0; o >>= 1)
    for (int k = 0; k < 3; ++k) {
      lo[k] = fminf(lo[k], __shfl_xor(lo[k], o, 64));
      hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], o, 64));
    }
  if (lane == 0)
    for (int k = 0; k < 3; ++k) {
      red[wave][k] = lo[k];
      red[wave][3 + k] = hi[k];
    }
  __syncthreads();
  float cen[3], half[3];
  for (int k = 0; k < 3; ++k) {
    float a = red[0][k], b = red[0][3 + k];
    for (int w = 1; w < kFB / 64; ++w) {
      a = fminf(a, red[w][k]);
      b = fmaxf(b, red[w][3 + k]);
    }
    cen[k] = m > 0 ? 0.5f * (a + b) : 0.0f;
    half[k] = m > 0 ? 0.5f * (b - a) : 0.0f;
  }
  // horizontal and vertical scales: the cross-section keeps its own
  // resolution however tall the trunk is
  const float sh = 1.0f / fmaxf(fmaxf(half[0], half[1]), 1e-3f);
  const float sz = 1.0f / fmaxf(half[2], 1e-3f);
  float4* Xc = X + (size_t)c * kP;
  for (int j = threadIdx.x; j < kP; j += kFB) {
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (j < m) {
      const int i = (int)(((long long)j * n) / m);
      v = make_float4((P[3 * (size_t)i] - cen[0]) * sh, (P[3 * (size_t)i + 1] - cen[1]) * sh,
                      (P[3 * (size_t)i + 2] - cen[2]) * sz, 1.0f);
    }
    Xc[j] = v;
  }
  if (threadIdx.x == 0) {
    PnMeta mt;
    mt.cx = cen[0];
    mt.cy = cen[1];
    mt.cz = cen[2];
    mt.sh = sh;
    mt.sz = sz;
    mt.m = m;
    mt.n = n;
    mt.pad = 0;
    meta[c] = mt;
  }
}

// ---------------------------------------------------------------------------
// forward.  act (nullable): saved f / argmax / h4 / out; H1, H2 (nullable):
// the BF16 activations of layers 1 and 2 of every point; prop (nullable):
// proposal records in the world frame.
// ---------------------------------------------------------------------------
// kTrain: also the first maximising point of every channel (for the
// backward pass) and the BF16 activations H1 / H2; inference keeps the
// maxima only.
template <bool kTrain>
__global__ __launch_bounds__(kFW) void pn_fwd_kernel(
    const float* __restrict__ prm, const float4* __restrict__ X, const PnMeta* __restrict__ meta,
    int n_clusters, const int32_t* __restrict__ d_n, float* __restrict__ act,
    unsigned short* __restrict__ H1, unsigned short* __restrict__ H2) {
  __shared__ __align__(16) unsigned short h1s[kChunk * kS1];
  __shared__ __align__(16) unsigned short h2s[kChunk * kS2];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int lr = lane & 15, lg = lane >> 4;

  // weight fragments of this wave (lane -> output channel lr of its tile,
  // k = 8 lg + e): layer 2 (A operand) channels 16w.., layer 3 (B operand)
  // columns 32w..
  bf16x8 w2f[2], w3f[2][4];
#pragma unroll
  for (int ks = 0; ks < 2; ++ks)
    w2f[ks] = frag_from_f32(prm + kW2 + (16 * wave + lr) * kC1 + 32 * ks + 8 * lg);
#pragma unroll
  for (int ct = 0; ct < 2; ++ct)
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
      w3f[ct][ks] = frag_from_f32(prm + kW3 + (32 * wave + 16 * ct + lr) * kC2 + 32 * ks + 8 * lg);
  // layer-2 biases of the channels 16w + 4lg + e this lane's results hold
  float b2v[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) b2v[e] = bf_round(prm[kB2 + 16 * wave + 4 * lg + e]);
  float b3v[2];
#pragma unroll
  for (int ct = 0; ct < 2; ++ct) b3v[ct] = bf_round(prm[kB3 + 32 * wave + 16 * ct + lr]);
  // layer 1 on v_mfma_f32_16x16x4f32, transposed: A[channel i][k] with
  // k = (x, y, z, 1) -> (w0, w1, w2, bias); lane -> i = lr, k = lg; one
  // fragment per 16-channel tile
  float w1f[kC1 / 16];
#pragma unroll
  for (int ct = 0; ct < kC1 / 16; ++ct) {
    const int ch = 16 * ct + lr;
    w1f[ct] = bf_round(lg < 3 ? prm[kW1 + 3 * ch + lg] : prm[kB1 + ch]);
  }

  // persistent over clusters: the weight fragments stay in registers
  const int n_used = pn_count(n_clusters, d_n);
  for (int c = blockIdx.x; c < n_used; c += gridDim.x) {
  const int m = meta[c].m;
  // training: biased outputs, ReLU floor 0; inference: raw products
  float best[2] = {kTrain ? 0.f : -INFINITY, kTrain ? 0.f : -INFINITY};
  int bidx[2] = {-1, -1};
  const float4* Xc = X + (size_t)c * kP;
  for (int base = 0; base < m; base += kChunk) {
    // layer 1, transposed: wave w takes points 16w.. of the chunk;
    // D[channel 16ct + 4lg + e][point 16w + lr], B[k][point] = (x, y, z, 1)
    {
      const float4 xp = Xc[base + 16 * wave + lr];
      const float xb = lg == 0 ? xp.x : (lg == 1 ? xp.y : (lg == 2 ? xp.z : 1.0f));
      const int i = 16 * wave + lr;
#pragma unroll
      for (int ct = 0; ct < kC1 / 16; ++ct) {
        const f32x4 d = __builtin_amdgcn_mfma_f32_16x16x4f32(w1f[ct], xb,
                                                             f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        unsigned short h[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) h[e] = bf_bits_hw(fmaxf(d[e], 0.f));
        const uint2 v = make_uint2((unsigned)h[0] | ((unsigned)h[1] << 16),
                                   (unsigned)h[2] | ((unsigned)h[3] << 16));
        const int j = 16 * ct + 4 * lg;
        *reinterpret_cast<uint2*>(h1s + i * kS1 + j) = v;
        if (kTrain) *reinterpret_cast<uint2*>(H1 + ((size_t)c * kP + base + i) * kC1 + j) = v;
      }
    }
    __syncthreads();
    // layer 2, transposed: D[channel 16w + 4lg + e][point 16pt + lr] for the
    // chunk's eight point tiles
    {
      f32x4 acc[kChunk / 16];
#pragma unroll
      for (int pt = 0; pt < kChunk / 16; ++pt) acc[pt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int pt = 0; pt < kChunk / 16; ++pt) {
          const bf16x8 b = frag_from_lds(h1s + (16 * pt + lr) * kS1 + 32 * ks + 8 * lg);
          acc[pt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2f[ks], b, acc[pt], 0, 0, 0);
        }
      const int j = 16 * wave + 4 * lg;
#pragma unroll
      for (int pt = 0; pt < kChunk / 16; ++pt) {
        const int i = 16 * pt + lr;
        unsigned short h[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          h[e] = bf_bits_hw(fmaxf(acc[pt][e] + b2v[e], 0.f));
        }
        const uint2 v = make_uint2((unsigned)h[0] | ((unsigned)h[1] << 16),
                                   (unsigned)h[2] | ((unsigned)h[3] << 16));
        *reinterpret_cast<uint2*>(h2s + i * kS2 + j) = v;
        if (kTrain) *reinterpret_cast<uint2*>(H2 + ((size_t)c * kP + base + i) * kC2 + j) = v;
      }
    }
    __syncthreads();
    // layer 3 + max-pool: the chunk's points x this wave's 32 channels, two
    // row tiles x two column tiles in flight
#pragma unroll
    for (int rp = 0; rp < kChunk / 16; rp += 2) {
      f32x4 acc[2][2];
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) acc[r][ct] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        bf16x8 a[2];
#pragma unroll
        for (int r = 0; r < 2; ++r)
          a[r] = frag_from_lds(h2s + (16 * (rp + r) + lr) * kS2 + 32 * ks + 8 * lg);
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
          for (int ct = 0; ct < 2; ++ct)
            acc[r][ct] =
                __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[r], w3f[ct][ks], acc[r][ct], 0, 0, 0);
      }
      if (kTrain) {
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
          for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int p = base + 16 * (rp + r) + 4 * lg + e;
              const float y = acc[r][ct][e] + b3v[ct];
              if (p < m && y > best[ct]) {
                best[ct] = y;
                bidx[ct] = p;
              }
            }
      } else if (base + kChunk <= m) {
        // maxima of the products; the bias is added after the maximum (the
        // rounded sum is monotone in the product)
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
          for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int e = 0; e < 4; ++e) best[ct] = fmaxf(best[ct], acc[r][ct][e]);
      } else {
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
          for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int p = base + 16 * (rp + r) + 4 * lg + e;
              if (p < m) best[ct] = fmaxf(best[ct], acc[r][ct][e]);
            }
      }
    }
    __syncthreads();
  }
  // max over the lanes holding the same channel (groups lg), first point on ties
#pragma unroll
  for (int ct = 0; ct < 2; ++ct) {
    float v = best[ct];
    int ix = bidx[ct];
#pragma unroll
    for (int o = 16; o < 64; o <<= 1) {
      const float v2 = __shfl_xor(v, o, 64);
      if (kTrain) {
        const int i2 = __shfl_xor(ix, o, 64);
        if (v2 > v || (v2 == v && i2 >= 0 && (ix < 0 || i2 < ix))) {
          v = v2;
          ix = i2;
        }
      } else {
        v = fmaxf(v, v2);
      }
    }
    if (lg == 0) {
      const int ch = 32 * wave + 16 * ct + lr;
      float* a = act + (size_t)c * kAct;
      if (kTrain) {
        a[kAF + ch] = v;
        a[kAI + ch] = __int_as_float(ix);
      } else {
        a[kAF + ch] = fmaxf(v + b3v[ct], 0.f);  // ReLU of the max of the biased outputs
      }
    }
  }
  }
}

// ---------------------------------------------------------------------------
// head: 256 -> 128 (ReLU) -> 7 for kHC clusters per workgroup (each weight row
// read once per workgroup), proposals in the world frame (prop nullable)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kFB) void pn_head_kernel(const float* __restrict__ prm,
                                                      const PnMeta* __restrict__ meta,
                                                      int n_clusters,
                                                      const int32_t* __restrict__ d_n,
                                                      float* __restrict__ act,
                                                      float* __restrict__ prop) {
  __shared__ float fs[kHC][kC3];
  __shared__ float h4s[kHC][kC4];
  __shared__ float outs[kHC][8];
  const int t = threadIdx.x;
  const int c0 = blockIdx.x * kHC;
  const int nc = min(kHC, pn_count(n_clusters, d_n) - c0);
  if (nc <= 0) return;
  for (int i = t; i < nc * kC3; i += kFB)
    fs[i / kC3][i % kC3] = act[(size_t)(c0 + i / kC3) * kAct + kAF + i % kC3];
  __syncthreads();
  // layer 4 on all 256 threads: output row o, half of k per thread (fused
  // multiply-adds), the two halves added in LDS
  {
    const int o = t & (kC4 - 1);
    const int half = t / kC4;
    float z[kHC];
#pragma unroll
    for (int j = 0; j < kHC; ++j) z[j] = 0.f;
    const float4* w = reinterpret_cast<const float4*>(prm + kW4 + o * kC3 + half * (kC3 / 2));
    for (int q0 = 0; q0 < kC3 / 8; q0 += 16) {
      float4 wv[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) wv[r] = w[q0 + r];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int k = half * (kC3 / 2) + 4 * (q0 + r);
        const float wk[4] = {bf_round(wv[r].x), bf_round(wv[r].y), bf_round(wv[r].z),
                             bf_round(wv[r].w)};
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
          for (int j = 0; j < kHC; ++j) z[j] = __builtin_fmaf(wk[e], fs[j][k + e], z[j]);
      }
    }
    if (half == 1)
#pragma unroll
      for (int j = 0; j < kHC; ++j) h4s[j][o] = z[j];
    __syncthreads();
    if (half == 0) {
      const float bias = bf_round(prm[kB4 + o]);
#pragma unroll
      for (int j = 0; j < kHC; ++j) {
        const float v = bias + (z[j] + h4s[j][o]);
        h4s[j][o] = v > 0.f ? v : 0.f;
      }
    }
  }
  __syncthreads();
  if (t < kOut * kHC) {
    const int j = t / kOut, o = t % kOut;
    if (j < nc) {
      float z = bf_round(prm[kB5 + o]);
      const float4* w = reinterpret_cast<const float4*>(prm + kW5 + o * kC4);
      for (int q0 = 0; q0 < kC4 / 4; q0 += 16) {
        float4 wv[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) wv[r] = w[q0 + r];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int q = q0 + r;
          z = __builtin_fmaf(bf_round(wv[r].x), h4s[j][4 * q], z);
          z = __builtin_fmaf(bf_round(wv[r].y), h4s[j][4 * q + 1], z);
          z = __builtin_fmaf(bf_round(wv[r].z), h4s[j][4 * q + 2], z);
          z = __builtin_fmaf(bf_round(wv[r].w), h4s[j][4 * q + 3], z);
        }
      }
      outs[j][o] = z;
    }
  }
  __syncthreads();
  for (int i = t; i < nc * kC4; i += kFB)
    act[(size_t)(c0 + i / kC4) * kAct + kAH + i % kC4] = h4s[i / kC4][i % kC4];
  for (int i = t; i < nc * 8; i += kFB)
    act[(size_t)(c0 + i / 8) * kAct + kAO + i % 8] = (i % 8) < kOut ? outs[i / 8][i % 8] : 0.f;
  if (prop && t < nc) {
    const int c = c0 + t;
    const PnMeta mt = meta[c];
    const int m = mt.m;
    const float* o = outs[t];
    float* pr = prop + (size_t)c * SLOAM_PN_PROPOSAL_STRIDE;
    // back to the world frame: directions and offsets divide by the scales
    const float vx = o[0] / mt.sh, vy = o[1] / mt.sh, vz = o[2] / mt.sz;
    const float nv = sqrtf(vx * vx + vy * vy + vz * vz);
    const float r = o[6] / mt.sh;
    const bool ok = m >= 3 && nv > 0.f && r > 0.f && isfinite(r) && isfinite(nv);
    pr[0] = ok ? vx / nv : 0.f;
    pr[1] = ok ? vy / nv : 0.f;
    pr[2] = ok ? vz / nv : 1.f;
    pr[3] = mt.cx + o[3] / mt.sh;
    pr[4] = mt.cy + o[4] / mt.sh;
    pr[5] = mt.cz + o[5] / mt.sz;
    pr[6] = ok ? r : 0.f;
    pr[7] = ok ? 1.f : 0.f;
  }
}

// ---------------------------------------------------------------------------
// backward.  G: per-cluster gradients [B][kNP].
//
// The layer-2 products run on BF16 MFMA per tile of kQT active points.  dz2
// is FP32; it enters as hi = bf16(dz2) plus lo = bf16(dz2 - hi), 16
// significant bits (relative error below 2^-16), against H1 and the
// BF16-rounded W2, which are BF16 values already, so the products keep
// FP32-level accuracy with FP32 accumulation.
// ---------------------------------------------------------------------------
constexpr int kQT = 32;        // active points per tile (the K of one MFMA)
constexpr int kSQ = kQT + 8;   // row stride (bf16) of the [channel][point] tiles
constexpr int kSJ = kC2 + 8;   // row stride (bf16) of the [point][channel] tiles
// four waves: wave w owns dz1 channels 16w.. and dW2 rows 32w..; two
// threads per layer-2 channel fill the dz2 tile
static_assert(kFB == 256 && kC1 == 4 * 16 && kC2 == 4 * 32 && kFB == 2 * kC2 && kQT == 32,
              "backward tile layout");

__global__ __launch_bounds__(kFB) void pn_bwd_kernel(
    const float* __restrict__ prm, const float4* __restrict__ X, const PnMeta* __restrict__ meta,
    const float* __restrict__ act, const unsigned short* __restrict__ H1,
    const unsigned short* __restrict__ H2, const double* __restrict__ targets, LossCfg lc,
    int n_clusters, float* __restrict__ G, float* __restrict__ loss_out) {
  __shared__ float f[kC3], dz3[kC3], h4[kC4], dz4[kC4], dout[8];
  __shared__ int fidx[kC3];
  __shared__ int slot_of[kP];
  __shared__ int plist[kC3];
  __shared__ int ccount[kC3 + 1], cstart[kC3 + 1], corder[kC3];
  __shared__ int nact;
  // one tile of active points: dz2 as [point][channel] (A of dz1 = dz2 W2)
  // and [channel][point] (A of dW2 = dz2' H1), hi and lo parts; H1 as
  // [channel][point] (B of dW2); the points
  __shared__ __align__(16) unsigned short zq_hi[kQT * kSJ], zq_lo[kQT * kSJ];
  __shared__ __align__(16) unsigned short zj_hi[kC2 * kSQ], zj_lo[kC2 * kSQ];
  __shared__ __align__(16) unsigned short h1t[kC1 * kSQ];
  __shared__ float4 xq[kQT];
  __shared__ float db2s[kFB];
  const int c = blockIdx.x;
  if (c >= n_clusters) return;
  const int t = threadIdx.x;
  const float* a = act + (size_t)c * kAct;
  float* Gc = G + (size_t)c * kNP;
  const unsigned short* H1c = H1 + (size_t)c * kP * kC1;
  const unsigned short* H2c = H2 + (size_t)c * kP * kC2;
  const float4* Xc = X + (size_t)c * kP;
  for (int k = t; k < kC3; k += kFB) {
    f[k] = a[kAF + k];
    fidx[k] = __float_as_int(a[kAI + k]);
  }
  if (t < kC4) h4[t] = a[kAH + t];
  for (int p = t; p < kP; p += kFB) slot_of[p] = -1;
  if (t <= kC3) ccount[t] = 0;
  __syncthreads();
  // loss and its gradient w.r.t. the 7 outputs
  if (t == 0) {
    const PnMeta mt = meta[c];
    const double* tg = targets + (size_t)c * SLOAM_CYL_MODEL_STRIDE;
    float o[7];
    for (int k = 0; k < 7; ++k) o[k] = a[kAO + k];
    float L = 0.f;
    float d[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (tg[6] > 0.0 && mt.m >= 3) {
      // the target axis line in the normalised frame, its point nearest the
      // box centre, and the radius at the horizontal scale
      const double bx = tg[3] * mt.sh, by = tg[4] * mt.sh, bz = tg[5] * mt.sz;
      const double an = sqrt(bx * bx + by * by + bz * bz);
      const double ax = bx / an, ay = by / an, az = bz / an;
      const double rx = (tg[0] - mt.cx) * mt.sh, ry = (tg[1] - mt.cy) * mt.sh,
                   rz = (tg[2] - mt.cz) * mt.sz;
      const double pr = rx * ax + ry * ay + rz * az;
      const float cs[3] = {(float)(rx - pr * ax), (float)(ry - pr * ay), (float)(rz - pr * az)};
      const float A[3] = {(float)ax, (float)ay, (float)az};
      const float rs = (float)(tg[6] * mt.sh);
      const float nv = sqrtf(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]) + 1e-12f;
      const float ah[3] = {o[0] / nv, o[1] / nv, o[2] / nv};
      const float dd = ah[0] * A[0] + ah[1] * A[1] + ah[2] * A[2];
      L = lc.wa * (1.f - dd * dd);
      for (int k = 0; k < 3; ++k) d[k] = lc.wa * (-2.f * dd * (A[k] - dd * ah[k]) / nv);
      for (int k = 0; k < 3; ++k) {
        const float e = o[3 + k] - cs[k];
        L += lc.wc * e * e;
        d[3 + k] = lc.wc * 2.f * e;
      }
      const float er = o[6] - rs;
      L += lc.wr * er * er;
      d[6] = lc.wr * 2.f * er;
    }
    for (int k = 0; k < 7; ++k) dout[k] = d[k] * lc.inv_batch;
    dout[7] = 0.f;
    if (loss_out) loss_out[c] = L;
  }
  __syncthreads();
  // head: W5, b5, then dz4 = relu'(h4) * W5^T dout
  for (int i = t; i < kOut * kC4; i += kFB) Gc[kW5 + i] = dout[i / kC4] * h4[i % kC4];
  if (t < kOut) Gc[kB5 + t] = dout[t];
  if (t < kC4) {
    float s = 0.f;
    for (int o = 0; o < kOut; ++o) s += bf_round(prm[kW5 + o * kC4 + t]) * dout[o];
    dz4[t] = h4[t] > 0.f ? s : 0.f;
  }
  __syncthreads();
  for (int i = t; i < kC4 * kC3; i += kFB) Gc[kW4 + i] = dz4[i / kC3] * f[i % kC3];
  if (t < kC4) Gc[kB4 + t] = dz4[t];
  // df, dz3 (the max-pool passes the gradient to the argmax point)
  for (int k = t; k < kC3; k += kFB) {
    float s = 0.f;
    for (int o = 0; o < kC4; ++o) s += bf_round(prm[kW4 + o * kC3 + k]) * dz4[o];
    dz3[k] = fidx[k] >= 0 ? s : 0.f;
  }
  __syncthreads();
  // active points, in point order, and each one's channels in channel order
  for (int k = t; k < kC3; k += kFB)
    if (fidx[k] >= 0) slot_of[fidx[k]] = 0;
  __syncthreads();
  if (t == 0) {
    int s = 0;
    for (int p = 0; p < kP; ++p)
      if (slot_of[p] == 0) {
        slot_of[p] = s;
        plist[s] = p;
        ++s;
      } else {
        slot_of[p] = -1;
      }
    nact = s;
  }
  __syncthreads();
  const int P = nact;
  for (int k = t; k < kC3; k += kFB)
    if (fidx[k] >= 0) atomicAdd(&ccount[slot_of[fidx[k]]], 1);
  __syncthreads();
  if (t == 0) {
    int s = 0;
    for (int q = 0; q < P; ++q) {
      cstart[q] = s;
      s += ccount[q];
    }
    cstart[P] = s;
  }
  __syncthreads();
  for (int k = t; k < kC3; k += kFB) {
    if (fidx[k] < 0) continue;
    int before = 0;
    for (int k2 = 0; k2 < k; ++k2) before += (fidx[k2] == fidx[k]) ? 1 : 0;
    corder[cstart[slot_of[fidx[k]]] + before] = k;
  }
  __syncthreads();
  // layer 3 weights: dW3[ch][k] = dz3[ch] h2[argmax ch][k]
  for (int i = t; i < kC3 * kC2; i += kFB) {
    const int ch = i / kC2, k = i % kC2;
    Gc[kW3 + i] = fidx[ch] >= 0 ? dz3[ch] * bf_val(H2c[(size_t)fidx[ch] * kC2 + k]) : 0.f;
  }
  if (t < kC3) Gc[kB3 + t] = dz3[t];
  // layer-2 weight fragments for dz1 = dz2 W2 (B operand: column k = 16 wave
  // + lr, rows j = 32 ks + 8 lg + e), BF16-rounded as in the forward pass
  const int lane = t & 63, wave = t >> 6, lr = lane & 15, lg = lane >> 4;
  bf16x8 w2b[kC2 / 32];
#pragma unroll
  for (int ks = 0; ks < kC2 / 32; ++ks) {
    unsigned short h[8];
#pragma unroll
    for (int e = 0; e < 8; ++e)
      h[e] = bf_bits(prm[kW2 + (32 * ks + 8 * lg + e) * kC1 + 16 * wave + lr]);
    u32x4 u;
    u.x = (unsigned)h[0] | ((unsigned)h[1] << 16);
    u.y = (unsigned)h[2] | ((unsigned)h[3] << 16);
    u.z = (unsigned)h[4] | ((unsigned)h[5] << 16);
    u.w = (unsigned)h[6] | ((unsigned)h[7] << 16);
    w2b[ks] = __builtin_bit_cast(bf16x8, u);
  }
  // dW2 accumulators: rows j = 32 wave + 16 jt + 4 lg + e, columns 16 kt + lr
  f32x4 dw2[2][kC1 / 16];
#pragma unroll
  for (int jt = 0; jt < 2; ++jt)
#pragma unroll
    for (int kt = 0; kt < kC1 / 16; ++kt) dw2[jt][kt] = f32x4{0.f, 0.f, 0.f, 0.f};
  // layer-1 sums (x, y, z, bias) of channel 16 wave + lr over this lane's rows
  float s1[4] = {0.f, 0.f, 0.f, 0.f};
  // dB2 of channel jj over the tile rows of parity qpar
  const int jj = t & (kC2 - 1), qpar = t >> 7;
  float db2 = 0.f;
  for (int q0 = 0; q0 < P; q0 += kQT) {
    // dz2 of the tile's points (zero beyond P): the max-pool's channels of
    // each point back through W3, then ReLU'
    for (int qi = qpar; qi < kQT; qi += kFB / kC2) {
      const int q = q0 + qi;
      float v = 0.f;
      if (q < P) {
        float sm = 0.f;
        for (int r = cstart[q]; r < cstart[q + 1]; ++r) {
          const int ch = corder[r];
          sm += bf_round(prm[kW3 + ch * kC2 + jj]) * dz3[ch];
        }
        const float h2 = bf_val(H2c[(size_t)plist[q] * kC2 + jj]);
        v = h2 > 0.f ? sm : 0.f;
      }
      db2 += v;
      const unsigned short hi = bf_bits(v);
      const unsigned short lo = bf_bits(v - bf_val(hi));
      zq_hi[qi * kSJ + jj] = hi;
      zq_lo[qi * kSJ + jj] = lo;
      zj_hi[jj * kSQ + qi] = hi;
      zj_lo[jj * kSQ + qi] = lo;
    }
    for (int i = t; i < kQT * kC1; i += kFB) {
      const int qi = i / kC1, k = i % kC1, q = q0 + qi;
      h1t[k * kSQ + qi] = q < P ? H1c[(size_t)plist[q] * kC1 + k] : (unsigned short)0;
    }
    if (t < kQT) xq[t] = q0 + t < P ? Xc[plist[q0 + t]] : make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    // dW2[j][k] += sum_q dz2[q][j] h1[q][k]
#pragma unroll
    for (int jt = 0; jt < 2; ++jt) {
      const int j0 = 32 * wave + 16 * jt;
      const bf16x8 ah = frag_from_lds(zj_hi + (j0 + lr) * kSQ + 8 * lg);
      const bf16x8 al = frag_from_lds(zj_lo + (j0 + lr) * kSQ + 8 * lg);
#pragma unroll
      for (int kt = 0; kt < kC1 / 16; ++kt) {
        const bf16x8 bh = frag_from_lds(h1t + (16 * kt + lr) * kSQ + 8 * lg);
        dw2[jt][kt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, dw2[jt][kt], 0, 0, 0);
        dw2[jt][kt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, dw2[jt][kt], 0, 0, 0);
      }
    }
    // dz1[q][k] = relu'(h1[q][k]) sum_j dz2[q][j] W2[j][k] for this wave's
    // channels k, then the layer-1 sums over the tile's points
#pragma unroll
    for (int qt = 0; qt < kQT / 16; ++qt) {
      f32x4 d = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < kC2 / 32; ++ks) {
        const bf16x8 ah = frag_from_lds(zq_hi + (16 * qt + lr) * kSJ + 32 * ks + 8 * lg);
        const bf16x8 al = frag_from_lds(zq_lo + (16 * qt + lr) * kSJ + 32 * ks + 8 * lg);
        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, w2b[ks], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, w2b[ks], d, 0, 0, 0);
      }
      const int k = 16 * wave + lr;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int qi = 16 * qt + 4 * lg + e;
        const float g = bf_val(h1t[k * kSQ + qi]) > 0.f ? d[e] : 0.f;
        const float4 x = xq[qi];
        s1[0] += g * x.x;
        s1[1] += g * x.y;
        s1[2] += g * x.z;
        s1[3] += g;
      }
    }
    __syncthreads();  // the next tile overwrites the tiles
  }
#pragma unroll
  for (int jt = 0; jt < 2; ++jt)
#pragma unroll
    for (int kt = 0; kt < kC1 / 16; ++kt)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        Gc[kW2 + (32 * wave + 16 * jt + 4 * lg + e) * kC1 + 16 * kt + lr] = dw2[jt][kt][e];
  db2s[t] = db2;
  // layer 1: the four lane groups holding channel k, added in a fixed tree
#pragma unroll
  for (int cc = 0; cc < 4; ++cc) {
    float v = s1[cc];
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    s1[cc] = v;
  }
  if (lg == 0) {
    const int k = 16 * wave + lr;
    Gc[kW1 + 3 * k] = s1[0];
    Gc[kW1 + 3 * k + 1] = s1[1];
    Gc[kW1 + 3 * k + 2] = s1[2];
    Gc[kB1 + k] = s1[3];
  }
  __syncthreads();
  if (t < kC2) Gc[kB2 + t] = db2s[t] + db2s[t + kC2];
}

// sum over the batch of the per-cluster gradients (already scaled by 1/B),
// in two fixed-order levels: kRedSlices consecutive slices of the batch
// (grid y), each summed in cluster order with four loads in flight, then the
// slices in order
constexpr int kRedSlices = 8;

__global__ __launch_bounds__(256) void pn_reduce_part_kernel(const float* __restrict__ G, int B,
                                                             float* __restrict__ part) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kNP) return;
  const int sl = blockIdx.y;
  const int c0 = (int)((int64_t)B * sl / kRedSlices), c1 = (int)((int64_t)B * (sl + 1) / kRedSlices);
  float s = 0.f;
  int c = c0;
  for (; c + 4 <= c1; c += 4) {
    const float v0 = G[(size_t)c * kNP + i], v1 = G[(size_t)(c + 1) * kNP + i];
    const float v2 = G[(size_t)(c + 2) * kNP + i], v3 = G[(size_t)(c + 3) * kNP + i];
    s += v0;
    s += v1;
    s += v2;
    s += v3;
  }
  for (; c < c1; ++c) s += G[(size_t)c * kNP + i];
  part[(size_t)sl * kNP + i] = s;
}

__global__ __launch_bounds__(256) void pn_reduce_kernel(const float* __restrict__ part,
                                                        float* __restrict__ grad) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kNP) return;
  float s = 0.f;
  for (int sl = 0; sl < kRedSlices; ++sl) s += part[(size_t)sl * kNP + i];
  grad[i] = s;
}

__global__ __launch_bounds__(256) void pn_adam_kernel(float* __restrict__ p,
                                                      float* __restrict__ m1,
                                                      float* __restrict__ m2,
                                                      const float* __restrict__ g, float lr,
                                                      float b1, float b2, float eps, float c1,
                                                      float c2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kNP) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// persistent forward grid: two 8-wave workgroups per CU (the kernel's
// occupancy), at most one per cluster
int fwd_grid(sloam_ctx* ctx, int n_clusters) {
  int cus = 256;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) !=
          hipSuccess ||
      cus < 1)
    cus = 256;
  (void)hipGetLastError();
  return std::max(1, std::min(n_clusters, 2 * cus));
}

struct PnBuffers {
  float4* X;
  PnMeta* meta;
  float* act;
  unsigned short* H1;
  unsigned short* H2;
  float* G;
  float* part;  // kRedSlices partial sums of the gradient
  float* grad;
};

int pn_buffers(sloam_ctx* ctx, int B, bool train, PnBuffers* b) {
  auto al = [](size_t x) { return (x + 255) / 256 * 256; };
  const size_t nB = (size_t)(B > 0 ? B : 1);
  size_t need = al(sizeof(float4) * kP * nB) + al(sizeof(PnMeta) * nB) +
                al(sizeof(float) * kAct * nB);
  if (train)
    need += al(2 * (size_t)kP * kC1 * nB) +
            al(2 * (size_t)kP * kC2 * nB) + al(sizeof(float) * kNP * nB) +
            al(sizeof(float) * kNP * kRedSlices) + al(sizeof(float) * kNP);
  void* base = nullptr;
  const int st = sloam_ctx_scratch(ctx, 3, need, &base);
  if (st) return st;
  char* q = (char*)base;
  auto take = [&](size_t bytes) {
    char* r = q;
    q += al(bytes);
    return (void*)r;
  };
  *b = PnBuffers{};
  b->X = (float4*)take(sizeof(float4) * kP * nB);
  b->meta = (PnMeta*)take(sizeof(PnMeta) * nB);
  b->act = (float*)take(sizeof(float) * kAct * nB);
  if (train) {
    b->H1 = (unsigned short*)take(2 * (size_t)kP * kC1 * nB);
    b->H2 = (unsigned short*)take(2 * (size_t)kP * kC2 * nB);
    b->G = (float*)take(sizeof(float) * kNP * nB);
    b->part = (float*)take(sizeof(float) * kNP * kRedSlices);
    b->grad = (float*)take(sizeof(float) * kNP);
  }
  return SLOAM_OK;
}

int pn_check_cfg(const sloam_pn_train_cfg* cfg) {
  if (!cfg) return SLOAM_ERR_INVALID_ARG;
  if (!(cfg->w_axis >= 0.f) || !(cfg->w_centre >= 0.f) || !(cfg->w_radius >= 0.f))
    return SLOAM_ERR_INVALID_ARG;
  return SLOAM_OK;
}

// prep + forward (saving activations) + backward + reduce into b.grad / d_grad
int pn_loss_grad(sloam_ctx* ctx, const float* d_params, const float* d_xyz,
                 const int32_t* d_offsets, int B, const double* d_targets,
                 const sloam_pn_train_cfg* cfg, float* d_loss, float* d_grad, PnBuffers* b) {
  int st = pn_buffers(ctx, B, true, b);
  if (st) return st;
  LossCfg lc{cfg->w_axis, cfg->w_centre, cfg->w_radius, 1.0f / (float)B};
  {
    SloamKernelScope ks(ctx, "pn_prep");
    hipLaunchKernelGGL(pn_prep_kernel, dim3(B), dim3(kFB), 0, ctx->stream, d_xyz, d_offsets, B,
                       (const int32_t*)nullptr, b->X, b->meta);
  }
  {
    SloamKernelScope ks(ctx, "pn_fwd");
    hipLaunchKernelGGL(pn_fwd_kernel<true>, dim3(fwd_grid(ctx, B)), dim3(kFW), 0, ctx->stream,
                       d_params, b->X, b->meta, B, (const int32_t*)nullptr, b->act, b->H1, b->H2);
    hipLaunchKernelGGL(pn_head_kernel, dim3((B + kHC - 1) / kHC), dim3(kFB), 0, ctx->stream,
                       d_params, b->meta, B, (const int32_t*)nullptr, b->act, (float*)nullptr);
  }
  {
    SloamKernelScope ks(ctx, "pn_bwd");
    hipLaunchKernelGGL(pn_bwd_kernel, dim3(B), dim3(kFB), 0, ctx->stream, d_params, b->X,
                       b->meta, b->act, b->H1, b->H2, d_targets, lc, B, b->G, d_loss);
  }
  {
    SloamKernelScope ks(ctx, "pn_reduce");
    hipLaunchKernelGGL(pn_reduce_part_kernel, dim3((kNP + 255) / 256, kRedSlices), dim3(256), 0,
                       ctx->stream, b->G, B, b->part);
    hipLaunchKernelGGL(pn_reduce_kernel, dim3((kNP + 255) / 256), dim3(256), 0, ctx->stream,
                       b->part, d_grad ? d_grad : b->grad);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}

}  // namespace

extern "C" {

static int pn_infer(sloam_ctx* ctx, const float* d_params, const float* d_xyz,
                    const int32_t* d_offsets, int32_t n_clusters, const int32_t* d_n,
                    float* d_proposal) {
  if (!ctx || n_clusters < 0) return SLOAM_ERR_INVALID_ARG;
  if (n_clusters == 0) return SLOAM_OK;
  if (!d_params || !d_xyz || !d_offsets || !d_proposal) return SLOAM_ERR_INVALID_ARG;
  if (((uintptr_t)d_params & 15) != 0) return SLOAM_ERR_INVALID_ARG;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  PnBuffers b;
  int st = pn_buffers(ctx, n_clusters, false, &b);
  if (st) return st;
  {
    SloamKernelScope ks(ctx, "pn_prep");
    hipLaunchKernelGGL(pn_prep_kernel, dim3(n_clusters), dim3(kFB), 0, ctx->stream, d_xyz,
                       d_offsets, n_clusters, d_n, b.X, b.meta);
  }
  {
    SloamKernelScope ks(ctx, "pn_fwd");
    hipLaunchKernelGGL(pn_fwd_kernel<false>, dim3(fwd_grid(ctx, n_clusters)), dim3(kFW), 0,
                       ctx->stream, d_params, b.X, b.meta, n_clusters, d_n, b.act,
                       (unsigned short*)nullptr, (unsigned short*)nullptr);
  }
  {
    SloamKernelScope ks(ctx, "pn_head");
    hipLaunchKernelGGL(pn_head_kernel, dim3((n_clusters + kHC - 1) / kHC), dim3(kFB), 0,
                       ctx->stream, d_params, b.meta, n_clusters, d_n, b.act, d_proposal);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}

int sloam_pn_infer(sloam_ctx* ctx, const float* d_params, const float* d_xyz,
                   const int32_t* d_offsets, int32_t n_clusters, float* d_proposal) {
  return pn_infer(ctx, d_params, d_xyz, d_offsets, n_clusters, nullptr, d_proposal);
}

int sloam_pn_infer_counted(sloam_ctx* ctx, const float* d_params, const float* d_xyz,
                           const int32_t* d_offsets, int32_t cap_clusters,
                           const int32_t* d_n_clusters, float* d_proposal) {
  if (!d_n_clusters) return SLOAM_ERR_INVALID_ARG;
  return pn_infer(ctx, d_params, d_xyz, d_offsets, cap_clusters, d_n_clusters, d_proposal);
}

int sloam_pn_loss_grad(sloam_ctx* ctx, const float* d_params, const float* d_xyz,
                       const int32_t* d_offsets, int32_t n_clusters, const double* d_targets,
                       const sloam_pn_train_cfg* cfg, float* d_loss, float* d_grad) {
  if (!ctx || n_clusters <= 0 || !d_params || !d_xyz || !d_offsets || !d_targets || !d_grad)
    return SLOAM_ERR_INVALID_ARG;
  if (((uintptr_t)d_params & 15) != 0) return SLOAM_ERR_INVALID_ARG;
  int st = pn_check_cfg(cfg);
  if (st) return st;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  PnBuffers b;
  return pn_loss_grad(ctx, d_params, d_xyz, d_offsets, n_clusters, d_targets, cfg, d_loss, d_grad,
                      &b);
}

int sloam_pn_train_step(sloam_ctx* ctx, float* d_params, float* d_m, float* d_v, int32_t step,
                        const sloam_pn_train_cfg* cfg, const float* d_xyz,
                        const int32_t* d_offsets, int32_t n_clusters, const double* d_targets,
                        float* d_loss) {
  if (!ctx || n_clusters <= 0 || step < 1 || !d_params || !d_m || !d_v || !d_xyz || !d_offsets ||
      !d_targets)
    return SLOAM_ERR_INVALID_ARG;
  if (((uintptr_t)d_params & 15) != 0) return SLOAM_ERR_INVALID_ARG;
  int st = pn_check_cfg(cfg);
  if (st) return st;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  PnBuffers b;
  st = pn_loss_grad(ctx, d_params, d_xyz, d_offsets, n_clusters, d_targets, cfg, d_loss, nullptr,
                    &b);
  if (st) return st;
  const float c1 = 1.0f - powf(cfg->beta1, (float)step);
  const float c2 = 1.0f - powf(cfg->beta2, (float)step);
  {
    SloamKernelScope ks(ctx, "pn_adam");
    hipLaunchKernelGGL(pn_adam_kernel, dim3((kNP + 255) / 256), dim3(256), 0, ctx->stream,
                       d_params, d_m, d_v, b.grad, cfg->lr, cfg->beta1, cfg->beta2, cfg->eps, c1,
                       c2);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}

}  // extern "C"
