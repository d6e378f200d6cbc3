// Stable LSD radix sort of (uint64 key, int32 value) pairs for gfx950, 8-bit
// digits over the key bits [begin_bit, end_bit).  No memset and no host
// synchronisation: the sequence can be captured in a HIP graph.
//
// Up to 512 tiles of 2048 keys (n <= 1 Mi), two launches per digit:
//   rs_hist_rows   per-tile counts of the digit, table[tile][digit];
//   rs_pass        every workgroup derives its own scatter offsets from the
//                  table (digit totals over all tiles, exclusive scan over
//                  digits, counts of the tiles before its own), then ranks the
//                  keys of each 256-key round with wave ballots (peers of
//                  equal digit) and per-wave digit counts in LDS and scatters
//                  them stably.
// Larger inputs: per digit a per-tile histogram, one exclusive scan over
// (digit, tile) in digit-major order, and the same stable scatter.
//
// A device-side element count (nullable) bounds the sort below the capacity
// the grid is sized for; keys at and past it are neither read nor written.
#include "common.hpp"

namespace {

typedef unsigned long long u64;

constexpr int kRB = 256;      // threads per block
constexpr int kDigits = 256;  // 8-bit digits
// two-launch-per-digit path
constexpr int kFRounds = 8;
constexpr int kFTile = kRB * kFRounds;  // 2048 keys per tile
constexpr int kFMaxTiles = 512;
// three-kernel path
constexpr int kRounds = 16;
constexpr int kTile = kRB * kRounds;  // 4096 keys per tile

// row stride (tiles, a multiple of 4) of the digit-major count table of the
// two-launch path: 16-byte aligned rows
__host__ __device__ inline int count_stride(int n_tiles) { return (n_tiles + 3) & ~3; }

__device__ inline int eff_n(const int32_t* d_n, int cap) {
  if (!d_n) return cap;
  const int m = *d_n;
  return m < 0 ? 0 : (m < cap ? m : cap);
}

// Stable rank of this thread's key among the keys of its 256-key round, and
// the per-wave digit counts in cnt[wave][digit] (read after the caller's
// barrier).  Returns the rank among the wave's peers of equal digit.
__device__ inline int wave_rank(bool act, int d, int lane, uint64_t lt, int* cnt_wave) {
  uint64_t peers = __builtin_amdgcn_ballot_w64(act);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1;
    const uint64_t m = __builtin_amdgcn_ballot_w64(act && bit);
    peers &= bit ? m : ~m;
  }
  const int rank = __builtin_popcountll(peers & lt);
  if (act && rank == 0) cnt_wave[d] = __builtin_popcountll(peers);
  (void)lane;
  return rank;
}

// ---- two-launch-per-digit path -----------------------------------------------

__global__ __launch_bounds__(kRB) void rs_hist_rows_kernel(const u64* __restrict__ keys,
                                                           const int32_t* __restrict__ d_n,
                                                           int cap, int shift,
                                                           int32_t* __restrict__ cnt) {
  __shared__ int h[kDigits];
  const int t = threadIdx.x;
  h[t] = 0;
  __syncthreads();
  const int n = eff_n(d_n, cap);
  const int base = blockIdx.x * kFTile;
#pragma unroll
  for (int r = 0; r < kFRounds; ++r) {
    const int i = base + r * kRB + t;
    if (i < n) atomicAdd(&h[(int)((keys[i] >> shift) & 0xFFull)], 1);
  }
  __syncthreads();
  // digit-major: row t holds digit t's count of every tile
  cnt[(size_t)t * count_stride(gridDim.x) + blockIdx.x] = h[t];
}

__global__ __launch_bounds__(kRB) void rs_pass_kernel(
    const u64* __restrict__ kin, const int32_t* __restrict__ vin,
    const int32_t* __restrict__ d_n, int cap, int shift, int n_tiles,
    const int32_t* __restrict__ cnt, u64* __restrict__ kout, int32_t* __restrict__ vout) {
  __shared__ int off[kDigits];
  __shared__ int wc[kRB / 64][kDigits];
  __shared__ int wtot[kRB / 64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int n = eff_n(d_n, cap);
  const int base = blockIdx.x * kFTile;
  if (base >= n) return;  // uniform over the block
  // the tile's keys, loaded before the offsets are summed so that both
  // latencies overlap
  u64 kr[kFRounds];
  int32_t vr[kFRounds];
#pragma unroll
  for (int r = 0; r < kFRounds; ++r) {
    const int i = base + r * kRB + t;
    kr[r] = 0;
    vr[r] = 0;
    if (i < n) {
      kr[r] = kin[i];
      if (vin) vr[r] = vin[i];
    }
  }
  // offsets of this tile: digit totals (exclusive scan over digits) plus the
  // counts of the tiles before it, from digit t's row (16-byte loads; the
  // row's padding past n_tiles is never written and is masked)
  int tot = 0, before = 0;
  const int me = blockIdx.x;
  const int4* row = reinterpret_cast<const int4*>(cnt + (size_t)t * count_stride(n_tiles));
#pragma unroll 8
  for (int q = 0; q < count_stride(n_tiles) / 4; ++q) {
    const int4 v = row[q];
    const int b = 4 * q;
    const int c0 = b < n_tiles ? v.x : 0;
    const int c1 = b + 1 < n_tiles ? v.y : 0;
    const int c2 = b + 2 < n_tiles ? v.z : 0;
    const int c3 = b + 3 < n_tiles ? v.w : 0;
    tot += c0 + c1 + c2 + c3;
    before += (b < me ? c0 : 0) + (b + 1 < me ? c1 : 0) + (b + 2 < me ? c2 : 0) +
              (b + 3 < me ? c3 : 0);
  }
  int s = tot;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(s, o, 64);
    if (lane >= o) s += u;
  }
  if (lane == 63) wtot[wave] = s;
  __syncthreads();
  int wo = 0;
#pragma unroll
  for (int w = 0; w < kRB / 64; ++w) wo += w < wave ? wtot[w] : 0;
  off[t] = wo + s - tot + before;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int r = 0; r < kFRounds; ++r) {
#pragma unroll
    for (int w = 0; w < kRB / 64; ++w) wc[w][t] = 0;
    __syncthreads();
    const int i = base + r * kRB + t;
    const bool act = i < n;
    const u64 k = kr[r];
    const int32_t v = vr[r];
    const int d = act ? (int)((k >> shift) & 0xFFull) : 0;
    const int rank = wave_rank(act, d, lane, lt, wc[wave]);
    __syncthreads();
    int dest = 0;
    if (act) {
      dest = off[d] + rank;
      for (int w = 0; w < wave; ++w) dest += wc[w][d];
    }
    __syncthreads();
    int rt = 0;
#pragma unroll
    for (int w = 0; w < kRB / 64; ++w) rt += wc[w][t];
    off[t] += rt;
    if (act) {
      kout[dest] = k;
      if (vout) vout[dest] = v;
    }
    __syncthreads();
  }
}

// ---- three-kernel path ---------------------------------------------------------

__global__ __launch_bounds__(kRB) void rs_hist_kernel(const u64* __restrict__ keys,
                                                      const int32_t* __restrict__ d_n, int cap,
                                                      int shift, int n_tiles,
                                                      int32_t* __restrict__ hist) {
  __shared__ int h[kDigits];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int n = eff_n(d_n, cap);
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
    const u64* __restrict__ kin, const int32_t* __restrict__ vin, const int32_t* __restrict__ d_n,
    int cap, int shift, int n_tiles, const int32_t* __restrict__ hist, u64* __restrict__ kout,
    int32_t* __restrict__ vout) {
  __shared__ int off[kDigits];
  __shared__ int cnt[kRB / 64][kDigits];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  off[t] = hist[t * n_tiles + blockIdx.x];
  const int n = eff_n(d_n, cap);
  const int base = blockIdx.x * kTile;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < kRounds; ++r) {
    for (int w = 0; w < kRB / 64; ++w) cnt[w][t] = 0;
    __syncthreads();
    const int i = base + r * kRB + t;
    const bool act = i < n;
    u64 k = 0;
    int32_t v = 0;
    int d = 0;
    if (act) {
      k = kin[i];
      if (vin) v = vin[i];
      d = (int)((k >> shift) & 0xFFull);
    }
    const int rank = wave_rank(act, d, lane, lt, cnt[wave]);
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

__global__ __launch_bounds__(256) void rs_copy_kernel(const u64* __restrict__ ks,
                                                      const int32_t* __restrict__ vs,
                                                      const int32_t* __restrict__ d_n, int cap,
                                                      u64* __restrict__ kd,
                                                      int32_t* __restrict__ vd) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= eff_n(d_n, cap)) return;
  kd[i] = ks[i];
  if (vs) vd[i] = vs[i];
}

int n_tiles_fused(int n) { return (n + kFTile - 1) / kFTile; }

}  // namespace

size_t sloam_radix_hist_ints(int n) {
  const int tf = n_tiles_fused(n > 0 ? n : 1);
  const int t3 = (n + kTile - 1) / kTile;
  const size_t a = (size_t)kDigits * (size_t)count_stride(tf);
  const size_t b = (size_t)kDigits * (size_t)(t3 > 0 ? t3 : 1);
  return a > b ? a : b;
}

int sloam_radix_sort(sloam_ctx* ctx, unsigned long long* keys, int32_t* vals,
                     unsigned long long* keys_tmp, int32_t* vals_tmp, int n, int begin_bit,
                     int end_bit, const int32_t* d_n, int32_t* hist,
                     unsigned long long** keys_out, int32_t** vals_out) {
  if (keys_out) *keys_out = keys;
  if (vals_out) *vals_out = vals;
  if (n <= 1 || end_bit <= begin_bit) return SLOAM_OK;
  if (begin_bit < 0 || end_bit > 64 || !keys || !keys_tmp || !hist || (vals && !vals_tmp))
    return SLOAM_ERR_INVALID_ARG;
  u64 *ka = keys, *kb = keys_tmp;
  int32_t *va = vals, *vb = vals_tmp;
  int passes = 0;
  const int tf = n_tiles_fused(n);
  if (tf <= kFMaxTiles) {
    for (int shift = begin_bit; shift < end_bit; shift += 8, ++passes) {
      hipLaunchKernelGGL(rs_hist_rows_kernel, dim3(tf), dim3(kRB), 0, ctx->stream, ka, d_n, n,
                         shift, hist);
      hipLaunchKernelGGL(rs_pass_kernel, dim3(tf), dim3(kRB), 0, ctx->stream, ka,
                         vals ? va : nullptr, d_n, n, shift, tf, hist, kb, vals ? vb : nullptr);
      std::swap(ka, kb);
      std::swap(va, vb);
    }
  } else {
    const int n_tiles = (n + kTile - 1) / kTile;
    const int m = kDigits * n_tiles;
    for (int shift = begin_bit; shift < end_bit; shift += 8, ++passes) {
      hipLaunchKernelGGL(rs_hist_kernel, dim3(n_tiles), dim3(kRB), 0, ctx->stream, ka, d_n, n,
                         shift, n_tiles, hist);
      hipLaunchKernelGGL(rs_scan_kernel, dim3(1), dim3(1024), 0, ctx->stream, hist, m);
      hipLaunchKernelGGL(rs_scatter_kernel, dim3(n_tiles), dim3(kRB), 0, ctx->stream, ka,
                         vals ? va : nullptr, d_n, n, shift, n_tiles, hist, kb,
                         vals ? vb : nullptr);
      std::swap(ka, kb);
      std::swap(va, vb);
    }
  }
  if (keys_out) {
    // the caller reads the sorted pairs where the last pass left them
    *keys_out = ka;
    if (vals_out) *vals_out = vals ? va : nullptr;
  } else if (passes & 1) {
    hipLaunchKernelGGL(rs_copy_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, ka,
                       vals ? va : nullptr, d_n, n, keys, vals);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}
