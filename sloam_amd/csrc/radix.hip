// Stable LSD radix sort of (uint64 key, int32 value) pairs for gfx950, 8-bit
// digits.  Per digit: a per-tile histogram, one exclusive scan over
// (digit, tile) in digit-major order, and a stable scatter that ranks the keys
// of each 256-key round with wave ballots (peers of equal digit) and per-wave
// digit counts in LDS.  No memset or host synchronisation: the sequence can
// be captured in a HIP graph.
#include "common.hpp"

namespace {

constexpr int kRB = 256;                 // threads per block
constexpr int kRounds = 16;              // 256-key rounds per tile
constexpr int kTile = kRB * kRounds;     // keys per tile
constexpr int kDigits = 256;

__global__ __launch_bounds__(kRB) void rs_hist_kernel(const unsigned long long* __restrict__ keys,
                                                      int n, int shift, int n_tiles,
                                                      int32_t* __restrict__ hist) {
  __shared__ int h[kDigits];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int base = blockIdx.x * kTile;
  for (int r = 0; r < kRounds; ++r) {
    const int i = base + r * kRB + (int)threadIdx.x;
    if (i < n) atomicAdd(&h[(int)((keys[i] >> shift) & 0xFFull)], 1);
  }
  __syncthreads();
  hist[threadIdx.x * n_tiles + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of m entries in place by one workgroup
__global__ __launch_bounds__(1024) void rs_scan_kernel(int32_t* __restrict__ a, int m) {
  __shared__ int wtot[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int carry = 0;
  for (int base = 0; base < m; base += 1024) {
    const int i = base + (int)threadIdx.x;
    const int v = i < m ? a[i] : 0;
    int s = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(s, o, 64);
      if (lane >= o) s += t;
    }
    if (lane == 63) wtot[wave] = s;
    __syncthreads();
    int off = 0, all = 0;
#pragma unroll
    for (int w = 0; w < 16; ++w) {
      off += (w < wave) ? wtot[w] : 0;
      all += wtot[w];
    }
    if (i < m) a[i] = carry + off + s - v;
    carry += all;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kRB) void rs_scatter_kernel(
    const unsigned long long* __restrict__ kin, const int32_t* __restrict__ vin, int n, int shift,
    int n_tiles, const int32_t* __restrict__ hist, unsigned long long* __restrict__ kout,
    int32_t* __restrict__ vout) {
  __shared__ int off[kDigits];
  __shared__ int cnt[kRB / 64][kDigits];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  off[t] = hist[t * n_tiles + blockIdx.x];
  const int base = blockIdx.x * kTile;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < kRounds; ++r) {
    for (int w = 0; w < kRB / 64; ++w) cnt[w][t] = 0;
    __syncthreads();
    const int i = base + r * kRB + t;
    const bool act = i < n;
    unsigned long long k = 0;
    int32_t v = 0;
    int d = 0;
    if (act) {
      k = kin[i];
      if (vin) v = vin[i];
      d = (int)((k >> shift) & 0xFFull);
    }
    // lanes of this wave holding the same digit
    uint64_t peers = __builtin_amdgcn_ballot_w64(act);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1;
      const uint64_t m = __builtin_amdgcn_ballot_w64(act && bit);
      peers &= bit ? m : ~m;
    }
    const int rank = __builtin_popcountll(peers & lt);
    if (act && rank == 0) cnt[wave][d] = __builtin_popcountll(peers);
    __syncthreads();
    int dest = 0;
    if (act) {
      dest = off[d] + rank;
      for (int w = 0; w < wave; ++w) dest += cnt[w][d];
    }
    __syncthreads();
    int tot = 0;
    for (int w = 0; w < kRB / 64; ++w) tot += cnt[w][t];
    off[t] += tot;
    if (act) {
      kout[dest] = k;
      if (vout) vout[dest] = v;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void rs_copy_kernel(const unsigned long long* __restrict__ ks,
                                                      const int32_t* __restrict__ vs, int n,
                                                      unsigned long long* __restrict__ kd,
                                                      int32_t* __restrict__ vd) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  kd[i] = ks[i];
  if (vs) vd[i] = vs[i];
}

}  // namespace

size_t sloam_radix_hist_ints(int n) {
  const int n_tiles = (n + kTile - 1) / kTile;
  return (size_t)kDigits * (size_t)(n_tiles > 0 ? n_tiles : 1);
}

int sloam_radix_sort(sloam_ctx* ctx, unsigned long long* keys, int32_t* vals,
                     unsigned long long* keys_tmp, int32_t* vals_tmp, int n, int end_bit,
                     int32_t* hist) {
  if (n <= 1 || end_bit <= 0) return SLOAM_OK;
  if (end_bit > 64 || !keys || !keys_tmp || !hist || (vals && !vals_tmp))
    return SLOAM_ERR_INVALID_ARG;
  const int n_tiles = (n + kTile - 1) / kTile;
  const int m = kDigits * n_tiles;
  unsigned long long *ka = keys, *kb = keys_tmp;
  int32_t *va = vals, *vb = vals_tmp;
  int passes = 0;
  for (int shift = 0; shift < end_bit; shift += 8, ++passes) {
    hipLaunchKernelGGL(rs_hist_kernel, dim3(n_tiles), dim3(kRB), 0, ctx->stream, ka, n, shift,
                       n_tiles, hist);
    hipLaunchKernelGGL(rs_scan_kernel, dim3(1), dim3(1024), 0, ctx->stream, hist, m);
    hipLaunchKernelGGL(rs_scatter_kernel, dim3(n_tiles), dim3(kRB), 0, ctx->stream, ka,
                       vals ? va : nullptr, n, shift, n_tiles, hist, kb, vals ? vb : nullptr);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (passes & 1)
    hipLaunchKernelGGL(rs_copy_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, ka,
                       vals ? va : nullptr, n, keys, vals);
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}
