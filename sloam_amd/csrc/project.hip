// Spherical range projection, label mask and ground / tree partition (gfx950).
//
// Reference: seg::Segmentation::_doProjection, _makeTensor, _mask, maskCloud
// and _destaggerCloud (sloam/src/segmentation/inference.cpp:80-300).  With the
// per-point labels given (0 other, 1 ground, 2 tree), the class of a pixel is
// the label of its nearest point.
//
// Kernels (context stream):
//   proj_zbuf     one thread per point: float range / yaw / pitch with the
//                 reference's float-double mix, pixel (row, col), and a 64-bit
//                 atomicMin of (range bits << 32 | point index) per pixel: the
//                 nearest point wins, the lowest index on equal ranges (for
//                 non-negative floats the IEEE bit pattern is monotonic)
//   proj_pixels   one thread per pixel: range image, winner, mask
//                 ((int)range == 0 -> invalid -> 0; tree 255, ground 1)
//   part_count    ground flags per block of kPartItems points
//   part_scan     exclusive scan of the block counts (one workgroup)
//   part_scatter  order-preserving ground compaction + dense tree cloud
//                 (NaN where the pixel is not tree; optional destagger)
// Points with a non-finite coordinate take no part (the reference sorts their
// NaN ranges, which is undefined behaviour).
#include "common.hpp"

#include <cmath>
#include <mutex>

namespace {

constexpr int kPB = 256;
constexpr int kPartItems = 1024;  // points per partition block (4 per thread)

struct ProjCfgDev {
  int H, W;
  float fov_down_abs, fov;
  int destagger;
};

__device__ inline bool finite3(float x, float y, float z) {
  return isfinite(x) && isfinite(y) && isfinite(z);
}

__global__ __launch_bounds__(kPB) void fill_u64_kernel(unsigned long long* __restrict__ d, int n,
                                                       unsigned long long v) {
  const int i = blockIdx.x * kPB + threadIdx.x;
  if (i < n) d[i] = v;
}

__global__ __launch_bounds__(kPB) void proj_zbuf_kernel(const float4* __restrict__ xyzi, int n,
                                                        ProjCfgDev cfg, int2* __restrict__ proj,
                                                        unsigned long long* __restrict__ keys) {
  const int i = blockIdx.x * kPB + threadIdx.x;
  if (i >= n) return;
  const float4 p = xyzi[i];
  if (!finite3(p.x, p.y, p.z)) {
    proj[i] = make_int2(-1, -1);
    return;
  }
  const float range = sqrtf(p.x * p.x + p.y * p.y + p.z * p.z);
  const float yaw = -atan2f(p.y, p.x);
  const float pitch = asinf(p.z / range);
  // float proj_x = 0.5 * (yaw / M_PI + 1.0): double arithmetic, stored as float
  float proj_x = (float)(0.5 * ((double)yaw / 3.14159265358979323846 + 1.0));
  // float proj_y = 1.0 - (pitch + |fov_down|) / fov: float quotient, double 1.0 - q
  float proj_y = (float)(1.0 - (double)((pitch + cfg.fov_down_abs) / cfg.fov));
  proj_x = proj_x * (float)cfg.W;
  proj_y = proj_y * (float)cfg.H;
  proj_x = floorf(proj_x);
  proj_x = fminf((float)cfg.W - 1.0f, proj_x);
  proj_x = fmaxf(0.0f, proj_x);
  proj_y = floorf(proj_y);
  proj_y = fminf((float)cfg.H - 1.0f, proj_y);
  proj_y = fmaxf(0.0f, proj_y);
  const int col = (int)proj_x, row = (int)proj_y;
  proj[i] = make_int2(col, row);
  const unsigned long long key =
      ((unsigned long long)__float_as_uint(range) << 32) | (unsigned long long)(uint32_t)i;
  atomicMin(keys + (size_t)row * cfg.W + col, key);
}

__global__ __launch_bounds__(kPB) void proj_pixels_kernel(
    const unsigned long long* __restrict__ keys, const uint8_t* __restrict__ label, int n_pix,
    float* __restrict__ range_img, int32_t* __restrict__ winner, uint8_t* __restrict__ mask) {
  const int p = blockIdx.x * kPB + threadIdx.x;
  if (p >= n_pix) return;
  const unsigned long long k = keys[p];
  float r = 0.f;
  int w = -1;
  if (k != ~0ull) {
    r = __uint_as_float((uint32_t)(k >> 32));
    w = (int)(uint32_t)k;
  }
  uint8_t m = 0;
  if (w >= 0 && (int)r != 0) {
    const uint8_t c = label[w];
    m = c == 2 ? 255 : (c == 1 ? 1 : 0);
  }
  if (range_img) range_img[p] = r;
  if (winner) winner[p] = w;
  mask[p] = m;
}

__device__ inline uint8_t point_mask(const int2* __restrict__ proj, const uint8_t* __restrict__ mask,
                                     int W, int i) {
  const int2 q = proj[i];
  if (q.x < 0) return 0;
  return mask[(size_t)q.y * W + q.x];
}

__global__ __launch_bounds__(kPB) void part_count_kernel(const int2* __restrict__ proj,
                                                         const uint8_t* __restrict__ mask, int n,
                                                         int W, int32_t* __restrict__ block_counts) {
  __shared__ int wsum[kPB / 64];
  const int base = blockIdx.x * kPartItems;
  int c = 0;
  for (int j = threadIdx.x; j < kPartItems; j += kPB) {
    const int i = base + j;
    if (i < n && point_mask(proj, mask, W, i) == 1) ++c;
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kPB / 64; ++w) t += wsum[w];
    block_counts[blockIdx.x] = t;
  }
}

// exclusive scan of up to any number of block counts by one workgroup
__global__ __launch_bounds__(kPB) void part_scan_kernel(int32_t* __restrict__ block_counts,
                                                        int n_blocks, int32_t* __restrict__ total) {
  __shared__ int carry;
  __shared__ int wtot[kPB / 64];
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int base = 0; base < n_blocks; base += kPB) {
    const int i = base + (int)threadIdx.x;
    const int v = i < n_blocks ? block_counts[i] : 0;
    // inclusive wave scan
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
    if (i < n_blocks) block_counts[i] = excl;
    __syncthreads();
    if (threadIdx.x == kPB - 1) carry = excl + v;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kPB) void part_scatter_kernel(
    const float4* __restrict__ xyzi, const int2* __restrict__ proj,
    const uint8_t* __restrict__ mask, int n, int W, int destagger,
    const int32_t* __restrict__ block_offsets, float4* __restrict__ ground,
    float4* __restrict__ tree) {
  __shared__ int wcount[kPB / 64];
  __shared__ int run;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int base = blockIdx.x * kPartItems;
  if (threadIdx.x == 0) run = block_offsets[blockIdx.x];
  __syncthreads();
  for (int j0 = 0; j0 < kPartItems; j0 += kPB) {
    const int i = base + j0 + (int)threadIdx.x;
    uint8_t m = 0;
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i < n) {
      p = xyzi[i];
      m = point_mask(proj, mask, W, i);
    }
    const bool g = (m == 1);
    const uint64_t bits = __builtin_amdgcn_ballot_w64(g);
    const int before = __builtin_popcountll(bits & ((1ull << lane) - 1ull));
    if (lane == 0) wcount[wave] = __builtin_popcountll(bits);
    __syncthreads();
    int woff = run;
    for (int w = 0; w < wave; ++w) woff += wcount[w];
    if (g) ground[woff + before] = p;
    if (tree && i < n) {
      const float qn = __int_as_float(0x7fc00000);
      const float4 v = (m == 255) ? p : make_float4(qn, qn, qn, 0.f);
      if (!destagger) {
        tree[i] = v;
      } else {
        // even rows shift x, y, z by +32 columns; intensity stays in place
        const int r = i / W, c = i % W;
        if ((r & 1) == 0) {
          if (c + 32 < W) {
            float4* o = tree + i + 32;
            o->x = v.x;
            o->y = v.y;
            o->z = v.z;
          }
          if (c < 32) tree[i] = v;
          else tree[i].w = v.w;
        } else {
          tree[i] = v;
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
      for (int w = 0; w < kPB / 64; ++w) t += wcount[w];
      run += t;
    }
    __syncthreads();
  }
}

// Scratch of sloam_project_partition_noalloc: the z-buffer keys, the point
// projections (when the caller passes no d_proj) and the partition's block
// counts of one call of at most kNoAllocMaxPixels pixels and kNoAllocMaxPoints
// points.  Each device holds its own copy for the life of the process; the
// host reaches it only through the address hipGetSymbolAddress returns for
// that device (noalloc_ptr).
constexpr int64_t kNoAllocMaxPixels = int64_t(1) << 19;
constexpr int64_t kNoAllocMaxPoints = int64_t(1) << 19;
constexpr size_t kNoAllocBytes = sizeof(unsigned long long) * kNoAllocMaxPixels +
                                 sizeof(int2) * kNoAllocMaxPoints +
                                 sizeof(int32_t) * (kNoAllocMaxPoints / kPartItems + 2);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};
hipError_t g_noalloc_err[kMaxDevices] = {};  // the lookup's failure, if any

// Bytes of scratch one call needs: keys (8 B / pixel), projections (8 B /
// point) when the caller gives none, block counts (4 B / block + 1).
size_t proj_scratch_bytes(int n_pix, int N, bool own_proj) {
  const int n_blocks = (N + kPartItems - 1) / kPartItems;
  return sizeof(unsigned long long) * (size_t)n_pix + (own_proj ? 0 : sizeof(int2) * (size_t)N) +
         sizeof(int32_t) * (size_t)(n_blocks + 1);
}

int proj_check(sloam_ctx* ctx, const float* d_xyzi, const uint8_t* d_label, int64_t n,
               const sloam_proj_cfg* cfg, uint8_t* d_mask, float* d_ground, int32_t* d_n_ground) {
  if (!ctx || !cfg || !d_mask || !d_n_ground || n < 0) return SLOAM_ERR_INVALID_ARG;
  if (n > 0 && (!d_xyzi || !d_label || !d_ground)) return SLOAM_ERR_INVALID_ARG;
  if (cfg->height < 1 || cfg->width < 1 || cfg->height > 4096 || cfg->width > 65536)
    return SLOAM_ERR_INVALID_ARG;
  if (n > 0x7fffffff) return SLOAM_ERR_TOO_LARGE;
  if (cfg->destagger && n != (int64_t)cfg->height * cfg->width) return SLOAM_ERR_INVALID_ARG;
  return SLOAM_OK;
}

// Enqueues the projection and partition of one scan with its scratch at
// `base` (proj_scratch_bytes bytes).  `timed`: bracket the kernels with the
// context's timers (the no-allocation variant does not, since a timer's
// first use creates events).
int proj_run(sloam_ctx* ctx, const float* d_xyzi, const uint8_t* d_label, int64_t n,
             const sloam_proj_cfg* cfg, int32_t* d_proj, float* d_range, int32_t* d_winner,
             uint8_t* d_mask, float* d_ground, int32_t* d_n_ground, float* d_tree, void* base,
             bool timed) {
  const int H = cfg->height, W = cfg->width, N = (int)n;
  const int n_pix = H * W;
  const int n_blocks = (N + kPartItems - 1) / kPartItems;
  unsigned long long* keys = (unsigned long long*)base;
  int2* proj = d_proj ? (int2*)d_proj : (int2*)(keys + n_pix);
  int32_t* counts = (int32_t*)((char*)base + sizeof(unsigned long long) * n_pix +
                               (d_proj ? 0 : sizeof(int2) * (size_t)N));
  ProjCfgDev dc;
  dc.H = H;
  dc.W = W;
  {
    // float members computed as in the reference constructor (double -> float)
    const float fov_up = (float)(cfg->fov_up_deg / 180.0 * 3.14159265358979323846);
    const float fov_down = (float)(cfg->fov_down_deg / 180.0 * 3.14159265358979323846);
    dc.fov_down_abs = fabsf(fov_down);
    dc.fov = fabsf(fov_down) + fabsf(fov_up);
  }
  dc.destagger = cfg->destagger;
  sloam_ctx* tctx = timed ? ctx : nullptr;  // a null context times nothing
  {
    SloamKernelScope ks(tctx, "proj_zbuf");
    hipLaunchKernelGGL(fill_u64_kernel, dim3((n_pix + kPB - 1) / kPB), dim3(kPB), 0, ctx->stream,
                       keys, n_pix, ~0ull);
    if (N > 0)
      hipLaunchKernelGGL(proj_zbuf_kernel, dim3((N + kPB - 1) / kPB), dim3(kPB), 0, ctx->stream,
                         reinterpret_cast<const float4*>(d_xyzi), N, dc, proj, keys);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  {
    SloamKernelScope ks(tctx, "proj_pixels");
    hipLaunchKernelGGL(proj_pixels_kernel, dim3((n_pix + kPB - 1) / kPB), dim3(kPB), 0,
                       ctx->stream, keys, d_label, n_pix, d_range, d_winner, d_mask);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  {
    SloamKernelScope ks(tctx, "partition");
    if (N > 0) {
      hipLaunchKernelGGL(part_count_kernel, dim3(n_blocks), dim3(kPB), 0, ctx->stream, proj,
                         d_mask, N, W, counts);
      hipLaunchKernelGGL(part_scan_kernel, dim3(1), dim3(kPB), 0, ctx->stream, counts, n_blocks,
                         d_n_ground);
      hipLaunchKernelGGL(part_scatter_kernel, dim3(n_blocks), dim3(kPB), 0, ctx->stream,
                         reinterpret_cast<const float4*>(d_xyzi), proj, d_mask, N, W,
                         cfg->destagger, counts, reinterpret_cast<float4*>(d_ground),
                         reinterpret_cast<float4*>(d_tree));
    } else {
      int st2 = sloam_fill_u32(ctx, (uint32_t*)d_n_ground, 1, 0u);
      if (st2) return st2;
    }
  }
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}

}  // namespace

// Looks up this device's copy of the no-allocation scratch once per device
// (the device must be current).  Called when a context is created.
int sloam_project_noalloc_init(sloam_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices) return SLOAM_OK;  // variant unavailable
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (g_noalloc_ptr[ctx->device]) return SLOAM_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_noalloc_err[ctx->device] = e;
  if (e != hipSuccess) return sloam_record_hip_error(ctx, e, "hipGetSymbolAddress(scratch)");
  g_noalloc_ptr[ctx->device] = ptr;
  return ptr ? SLOAM_OK : SLOAM_ERR_HIP;
}

extern "C" {

int sloam_project_partition(sloam_ctx* ctx, const float* d_xyzi, const uint8_t* d_label,
                            int64_t n, const sloam_proj_cfg* cfg, int32_t* d_proj, float* d_range,
                            int32_t* d_winner, uint8_t* d_mask, float* d_ground,
                            int32_t* d_n_ground, float* d_tree) {
  int st = proj_check(ctx, d_xyzi, d_label, n, cfg, d_mask, d_ground, d_n_ground);
  if (st) return st;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  void* base = nullptr;
  st = sloam_ctx_scratch(ctx, 0, proj_scratch_bytes(cfg->height * cfg->width, (int)n, d_proj),
                         &base);
  if (st) return st;
  return proj_run(ctx, d_xyzi, d_label, n, cfg, d_proj, d_range, d_winner, d_mask, d_ground,
                  d_n_ground, d_tree, base, true);
}

int sloam_project_partition_noalloc(sloam_ctx* ctx, const float* d_xyzi, const uint8_t* d_label,
                                    int64_t n, const sloam_proj_cfg* cfg, int32_t* d_proj,
                                    float* d_range, int32_t* d_winner, uint8_t* d_mask,
                                    float* d_ground, int32_t* d_n_ground, float* d_tree) {
  int st = proj_check(ctx, d_xyzi, d_label, n, cfg, d_mask, d_ground, d_n_ground);
  if (st) return st;
  if ((int64_t)cfg->height * cfg->width > kNoAllocMaxPixels || n > kNoAllocMaxPoints) {
    ctx->last_error = "sloam_project_partition_noalloc: at most 524288 pixels and 524288 points";
    return SLOAM_ERR_TOO_LARGE;
  }
  void* base = nullptr;
  hipError_t lookup = hipSuccess;
  if (ctx->device >= 0 && ctx->device < kMaxDevices) {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    base = g_noalloc_ptr[ctx->device];
    lookup = g_noalloc_err[ctx->device];
  }
  if (!base) {
    ctx->last_error = std::string("sloam_project_partition_noalloc: no scratch address for "
                                  "this device (hipGetSymbolAddress: ") +
                      hipGetErrorString(lookup) + ")";
    return SLOAM_ERR_INTERNAL;
  }
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  return proj_run(ctx, d_xyzi, d_label, n, cfg, d_proj, d_range, d_winner, d_mask, d_ground,
                  d_n_ground, d_tree, base, false);
}

}  // extern "C"
