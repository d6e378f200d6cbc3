// Tree instance front end for gfx950: organized connected components of the
// tree cloud and per-beam tree vertices (sloam/src/segmentation/trellis.cpp:
// 15-132, Instance::findClusters / findTrees / computeVertexProperties).
//
//   tr_init, tr_union   union-find over the 4-neighbour grid: a pixel (finite
//                       x) joins its left and upper neighbours when their
//                       float distance is < the cluster threshold; unions
//                       link the larger root under the smaller with
//                       atomicMin, so every root is its component's smallest
//                       pixel index, i.e. its first pixel in row-major order
//   tr_flatten .. tr_label
//                       roots ranked in pixel order (ordered block scan) give
//                       the component numbers; per-component sizes
//   tr_key + radix sort pixels of components with more than 80 points, keyed
//                       (component, bottom-up row, column); the sort is the
//                       memset-free LSD sort of radix.hip
//   tr_segflag .. tr_segscatter
//                       one segment per (component, row) = one beam of a tree
//   tr_vertex           one wave per segment: medians of x, y, z (rank
//                       selection), points within 0.2 m of the median in
//                       (z, y, x, column) order, radius of the first / last
//   tr_assemble         one workgroup: vertices per tree (> 16 kept, first
//                       56), tree / vertex / point offsets
//   tr_copy             the kept points of each vertex to the output
#include "common.hpp"

#include <cstring>

namespace {

constexpr int kTB = 256;        // pixel kernels
constexpr int kItems = 1024;    // pixels per block of the ordered scans
constexpr int kAsm = 1024;      // assemble workgroup
constexpr int kSegLds = 1024;   // beam points staged in LDS (larger beams read HBM)
constexpr int kVtxBlocks = 2048;  // waves of the grid-stride vertex / copy kernels
// sort key: component | bottom-up row | column, with rb / cb bits for the row
// and column and the component above them; the padding key (all ones below
// key_bits) sorts last and matches no real pixel

struct TrCfgDev {
  int H, W, N;
  int rb, cb, key_bits;
  unsigned long long pad;
  float thr, max_dist;
  int min_label_points, min_row_points, min_vertices, max_vertices;
  double radius_scale;
  int cap_trees, cap_vertices, cap_points;
};

// scratch counters
enum { kCntComps = 0, kCntSegs = 1, kCntN = 8 };

__device__ inline int ld_relaxed(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline int uf_find(int* L, int a) {
  int p = ld_relaxed(L + a);
  while (p != a) {
    a = p;
    p = ld_relaxed(L + a);
  }
  return a;
}

// Parents only decrease (L[x] <= x); a retry happens only when another lane
// lowered the parent meanwhile, so the loop ends.
__device__ inline void uf_union(int* L, int a, int b) {
  bool done = false;
  while (!done) {
    a = uf_find(L, a);
    b = uf_find(L, b);
    if (a < b) {
      const int old = atomicMin(L + b, a);
      done = (old == b);
      b = old;
    } else if (b < a) {
      const int old = atomicMin(L + a, b);
      done = (old == a);
      a = old;
    } else {
      done = true;
    }
  }
}

// float distance as pcl::euclideanDistance: sqrt((dx^2 + dy^2) + dz^2)
__device__ inline float dist3(float ax, float ay, float az, float bx, float by, float bz) {
  const float dx = ax - bx;
  const float dy = ay - by;
  const float dz = az - bz;
  return sqrtf(dx * dx + dy * dy + dz * dz);
}

// also zeroes the component sizes and the counters (no fill launches)
__global__ __launch_bounds__(kTB) void tr_init_kernel(const float4* __restrict__ cloud, int N,
                                                      int* __restrict__ L, int* __restrict__ size,
                                                      int* __restrict__ counters) {
  const int p = blockIdx.x * kTB + threadIdx.x;
  if (p < N) {
    L[p] = isfinite(cloud[p].x) ? p : -1;
    size[p] = 0;
  }
  if (p < kCntN) counters[p] = 0;
}

__global__ __launch_bounds__(kTB) void tr_union_kernel(const float4* __restrict__ cloud,
                                                       TrCfgDev cfg, int* L) {
  const int p = blockIdx.x * kTB + threadIdx.x;
  if (p >= cfg.N) return;
  const float4 a = cloud[p];
  if (!isfinite(a.x)) return;
  const int c = p % cfg.W, r = p / cfg.W;
  if (c > 0) {
    const float4 b = cloud[p - 1];
    if (isfinite(b.x) && dist3(a.x, a.y, a.z, b.x, b.y, b.z) < cfg.thr) uf_union(L, p, p - 1);
  }
  if (r > 0) {
    const float4 b = cloud[p - cfg.W];
    if (isfinite(b.x) && dist3(a.x, a.y, a.z, b.x, b.y, b.z) < cfg.thr)
      uf_union(L, p, p - cfg.W);
  }
}

__global__ __launch_bounds__(kTB) void tr_flatten_kernel(int N, int* L, int* __restrict__ root) {
  const int p = blockIdx.x * kTB + threadIdx.x;
  if (p < N) root[p] = ld_relaxed(L + p) < 0 ? -1 : uf_find(L, p);
}

// --- ordered compaction helpers (blocks of kItems items) --------------------

// inclusive wave scan
__device__ inline int wave_incl(int v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// exclusive scan over the workgroup (NT threads); *total gets the sum.  All
// threads must call it.
template <int NT>
__device__ inline int block_excl(int v, int* wtot, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int s = wave_incl(v, lane);
  if (lane == 63) wtot[wave] = s;
  __syncthreads();
  int off = 0, all = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) {
    const int t = wtot[w];
    off += (w < wave) ? t : 0;
    all += t;
  }
  __syncthreads();
  *total = all;
  return off + s - v;
}

__global__ __launch_bounds__(kTB) void tr_count_roots_kernel(const int* __restrict__ root, int N,
                                                             int* __restrict__ block_counts) {
  __shared__ int wsum[kTB / 64];
  const int base = blockIdx.x * kItems;
  int c = 0;
  for (int j = threadIdx.x; j < kItems; j += kTB) {
    const int p = base + j;
    if (p < N && root[p] == p) ++c;
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kTB / 64; ++w) t += wsum[w];
    block_counts[blockIdx.x] = t;
  }
}

// exclusive scan of n block counts in place by one workgroup; total to *total
__global__ __launch_bounds__(kTB) void tr_scan_kernel(int* __restrict__ counts, int n,
                                                      int* __restrict__ total) {
  __shared__ int wtot[kTB / 64];
  int carry = 0;
  for (int base = 0; base < n; base += kTB) {
    const int i = base + (int)threadIdx.x;
    const int v = i < n ? counts[i] : 0;
    int all;
    const int e = block_excl<kTB>(v, wtot, &all);
    if (i < n) counts[i] = carry + e;
    carry += all;
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kTB) void tr_rootid_kernel(const int* __restrict__ root, int N,
                                                        const int* __restrict__ block_offsets,
                                                        int* __restrict__ cid) {
  __shared__ int wcount[kTB / 64];
  __shared__ int run;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int base = blockIdx.x * kItems;
  if (threadIdx.x == 0) run = block_offsets[blockIdx.x];
  __syncthreads();
  for (int j0 = 0; j0 < kItems; j0 += kTB) {
    const int p = base + j0 + (int)threadIdx.x;
    const bool is_root = p < N && root[p] == p;
    const uint64_t bits = __builtin_amdgcn_ballot_w64(is_root);
    const int before = __builtin_popcountll(bits & ((1ull << lane) - 1ull));
    if (lane == 0) wcount[wave] = __builtin_popcountll(bits);
    __syncthreads();
    int off = run;
    for (int w = 0; w < wave; ++w) off += wcount[w];
    if (is_root) cid[p] = off + before;
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
      for (int w = 0; w < kTB / 64; ++w) t += wcount[w];
      run += t;
    }
    __syncthreads();
  }
}

// label[p] = component id of pixel p (-1: no component); size[l] counts the
// pixels of component l.  Threads walk the image column by column (a wave
// holds 64 rows of one column: mostly one trunk) and each wave adds one
// atomic per distinct label it holds, so a large component does not
// serialise on its counter.
__global__ __launch_bounds__(kTB) void tr_label_kernel(const int* __restrict__ root,
                                                       const int* __restrict__ cid, TrCfgDev cfg,
                                                       int* __restrict__ label,
                                                       int* __restrict__ size) {
  const int idx = blockIdx.x * kTB + threadIdx.x;
  int l = -1;
  if (idx < cfg.N) {
    const int p = (idx % cfg.H) * cfg.W + idx / cfg.H;
    const int r = root[p];
    l = r < 0 ? -1 : cid[r];
    label[p] = l;
  }
  bool pending = l >= 0;
  for (uint64_t m = __builtin_amdgcn_ballot_w64(pending); m != 0;
       m = __builtin_amdgcn_ballot_w64(pending)) {
    const int lv = __shfl(l, (int)__builtin_ctzll(m), 64);
    const uint64_t same = __builtin_amdgcn_ballot_w64(pending && l == lv);
    if ((threadIdx.x & 63) == (int)__builtin_ctzll(m)) atomicAdd(size + lv, __builtin_popcountll(same));
    if (l == lv) pending = false;
  }
}

__global__ __launch_bounds__(kTB) void tr_key_kernel(const int* __restrict__ label,
                                                     const int* __restrict__ size, TrCfgDev cfg,
                                                     unsigned long long* __restrict__ keys) {
  const int p = blockIdx.x * kTB + threadIdx.x;
  if (p >= cfg.N) return;
  const int l = label[p];
  const int r = p / cfg.W, c = p % cfg.W;
  unsigned long long k = cfg.pad;
  if (l >= 0 && size[l] > cfg.min_label_points) {
    k = ((unsigned long long)l << (cfg.rb + cfg.cb)) |
        ((unsigned long long)(cfg.H - 1 - r) << cfg.cb) | (unsigned long long)c;
  }
  // stored in (H-1-r, c) order: a stable sort by the label bits alone then
  // yields the order of the whole key (label, H-1-r, c)
  keys[(cfg.H - 1 - r) * cfg.W + c] = k;
}

__device__ inline bool seg_start_at(const unsigned long long* keys, int i, const TrCfgDev& cfg) {
  const unsigned long long k = keys[i];
  return k != cfg.pad && (i == 0 || (keys[i - 1] >> cfg.cb) != (k >> cfg.cb));
}

__global__ __launch_bounds__(kTB) void tr_segcount_kernel(const unsigned long long* __restrict__ keys,
                                                          int N, TrCfgDev cfg,
                                                          int* __restrict__ block_counts) {
  __shared__ int wsum[kTB / 64];
  const int base = blockIdx.x * kItems;
  int c = 0;
  for (int j = threadIdx.x; j < kItems; j += kTB) {
    const int i = base + j;
    if (i < N && seg_start_at(keys, i, cfg)) ++c;
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kTB / 64; ++w) t += wsum[w];
    block_counts[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kTB) void tr_segscatter_kernel(
    const unsigned long long* __restrict__ keys, int N, TrCfgDev cfg,
    const int* __restrict__ block_offsets, int* __restrict__ seg_start) {
  __shared__ int wcount[kTB / 64];
  __shared__ int run;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int base = blockIdx.x * kItems;
  if (threadIdx.x == 0) run = block_offsets[blockIdx.x];
  __syncthreads();
  for (int j0 = 0; j0 < kItems; j0 += kTB) {
    const int i = base + j0 + (int)threadIdx.x;
    const bool f = i < N && seg_start_at(keys, i, cfg);
    const uint64_t bits = __builtin_amdgcn_ballot_w64(f);
    const int before = __builtin_popcountll(bits & ((1ull << lane) - 1ull));
    if (lane == 0) wcount[wave] = __builtin_popcountll(bits);
    __syncthreads();
    int off = run;
    for (int w = 0; w < wave; ++w) off += wcount[w];
    if (f) seg_start[off + before] = i;
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
      for (int w = 0; w < kTB / 64; ++w) t += wcount[w];
      run += t;
    }
    __syncthreads();
  }
}

// --- vertices ----------------------------------------------------------------

struct SegOut {
  float4 med;     // median x, y, z; w unused
  double radius;
  int valid, m;
};

// (z, y, x, column) lexicographic order of two beam points
__device__ inline bool zless(const float4& a, int ca, const float4& b, int cb) {
  if (a.z != b.z) return a.z < b.z;
  if (a.y != b.y) return a.y < b.y;
  if (a.x != b.x) return a.x < b.x;
  return ca < cb;
}

__device__ inline float axis_of(const float4& p, int a) { return a == 0 ? p.x : (a == 1 ? p.y : p.z); }

__global__ __launch_bounds__(64) void tr_vertex_kernel(
    const float4* __restrict__ cloud, const unsigned long long* __restrict__ keys,
    const int* __restrict__ seg_start, const int* __restrict__ counters, TrCfgDev cfg,
    SegOut* __restrict__ seg_out, float4* __restrict__ tmp) {
  __shared__ float4 sp[kSegLds];
  __shared__ int sc[kSegLds];
  __shared__ unsigned char skeep[kSegLds];
  const int lane = threadIdx.x;
  const int S = counters[kCntSegs];
  for (int s = blockIdx.x; s < S; s += gridDim.x) {
    const int start = seg_start[s];
    const unsigned long long k0 = keys[start];
    const unsigned long long hi = k0 >> cfg.cb;
    const int row = cfg.H - 1 - (int)((k0 >> cfg.cb) & ((1ull << cfg.rb) - 1ull));
    // beam length: the run of equal (component, row) keys
    int n = 0;
    for (;;) {
      const int i = start + n + lane;
      const bool same = i < cfg.N && (keys[i] >> cfg.cb) == hi;
      const uint64_t b = __builtin_amdgcn_ballot_w64(same);
      // keys are sorted: the run is a prefix of the lanes
      const int run = (b == ~0ull) ? 64 : __builtin_ctzll(~b);
      n += run;
      if (run < 64) break;
    }
    SegOut o;
    o.med = make_float4(0.f, 0.f, 0.f, 0.f);
    o.radius = 0.0;
    o.valid = 0;
    o.m = 0;
    if (n > cfg.min_row_points) {
      const bool in_lds = n <= kSegLds;
      auto col_of = [&](int j) -> int {
        return (int)(keys[start + j] & ((1ull << cfg.cb) - 1ull));
      };
      if (in_lds) {
        for (int j = lane; j < n; j += 64) {
          const int c = col_of(j);
          sc[j] = c;
          sp[j] = cloud[row * cfg.W + c];
        }
        __syncthreads();
      }
      auto P = [&](int j) -> float4 { return in_lds ? sp[j] : cloud[row * cfg.W + col_of(j)]; };
      auto C = [&](int j) -> int { return in_lds ? sc[j] : col_of(j); };
      // medians: the element of rank floor(n / 2) by (value, index)
      const int mid = n / 2;
      float med[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        float found = 0.f;
        bool has = false;
        for (int i = lane; i < n; i += 64) {
          const float vi = axis_of(P(i), a);
          int rank = 0;
          for (int j = 0; j < n; ++j) {
            const float vj = axis_of(P(j), a);
            rank += (vj < vi || (vj == vi && j < i)) ? 1 : 0;
          }
          if (rank == mid) {
            found = vi;
            has = true;
          }
        }
        const uint64_t b = __builtin_amdgcn_ballot_w64(has);
        med[a] = __shfl(found, (int)__builtin_ctzll(b), 64);
      }
      // kept points
      int m = 0;
      for (int i0 = 0; i0 < n; i0 += 64) {
        const int i = i0 + lane;
        bool keep = false;
        if (i < n) {
          const float4 p = P(i);
          keep = dist3(p.x, p.y, p.z, med[0], med[1], med[2]) < cfg.max_dist;
          if (in_lds) skeep[i] = keep ? 1 : 0;
        }
        m += __builtin_popcountll(__builtin_amdgcn_ballot_w64(keep));
      }
      __syncthreads();
      o.med = make_float4(med[0], med[1], med[2], 0.f);
      o.m = m;
      if (m > 1) {
        float4 first = make_float4(0.f, 0.f, 0.f, 0.f), last = first;
        bool hf = false, hl = false;
        for (int i = lane; i < n; i += 64) {
          const float4 p = P(i);
          const bool keep =
              in_lds ? skeep[i] != 0
                     : dist3(p.x, p.y, p.z, med[0], med[1], med[2]) < cfg.max_dist;
          if (!keep) continue;
          const int ci = C(i);
          int rank = 0;
          for (int j = 0; j < n; ++j) {
            const float4 q = P(j);
            const bool kj = in_lds ? skeep[j] != 0
                                   : dist3(q.x, q.y, q.z, med[0], med[1], med[2]) < cfg.max_dist;
            if (kj && zless(q, C(j), p, ci)) ++rank;
          }
          tmp[start + rank] = p;
          if (rank == 0) {
            first = p;
            hf = true;
          }
          if (rank == m - 1) {
            last = p;
            hl = true;
          }
        }
        const int lf = (int)__builtin_ctzll(__builtin_amdgcn_ballot_w64(hf));
        const int ll = (int)__builtin_ctzll(__builtin_amdgcn_ballot_w64(hl));
        const float fx = __shfl(first.x, lf, 64), fy = __shfl(first.y, lf, 64),
                    fz = __shfl(first.z, lf, 64);
        const float lx = __shfl(last.x, ll, 64), ly = __shfl(last.y, ll, 64),
                    lz = __shfl(last.z, ll, 64);
        o.radius = cfg.radius_scale * (double)dist3(fx, fy, fz, lx, ly, lz);
        o.valid = 1;
      }
      __syncthreads();  // LDS reused by the next segment
    }
    if (lane == 0) seg_out[s] = o;
  }
}

// --- assembly ----------------------------------------------------------------

__device__ inline int seg_comp(const unsigned long long* keys, const int* seg_start, int s,
                               const TrCfgDev& cfg) {
  return (int)(keys[seg_start[s]] >> (cfg.rb + cfg.cb));
}

__global__ __launch_bounds__(kAsm) void tr_assemble_kernel(
    const unsigned long long* __restrict__ keys, const int* __restrict__ seg_start,
    const SegOut* __restrict__ seg_out, int* __restrict__ counters, TrCfgDev cfg,
    int* __restrict__ segP, int* __restrict__ compbase, int* __restrict__ comptotal,
    int* __restrict__ seg_vi, int* __restrict__ seg_po, int32_t* __restrict__ tree_offsets,
    int32_t* __restrict__ vertex_int, double* __restrict__ vertex_f64,
    int32_t* __restrict__ point_offsets, int32_t* __restrict__ out_counts) {
  __shared__ int wtot[kAsm / 64];
  const int S = counters[kCntSegs];
  // A: exclusive prefix of valid vertices; base of each component
  int carry = 0;
  for (int base = 0; base < S; base += kAsm) {
    const int s = base + (int)threadIdx.x;
    const int v = s < S ? seg_out[s].valid : 0;
    int all;
    const int e = block_excl<kAsm>(v, wtot, &all);
    if (s < S) {
      segP[s] = carry + e;
      const int c = seg_comp(keys, seg_start, s, cfg);
      if (s == 0 || seg_comp(keys, seg_start, s - 1, cfg) != c) compbase[c] = carry + e;
    }
    carry += all;
  }
  __syncthreads();
  // B: valid vertices per component
  for (int s = threadIdx.x; s < S; s += kAsm) {
    const int c = seg_comp(keys, seg_start, s, cfg);
    if (s == S - 1 || seg_comp(keys, seg_start, s + 1, cfg) != c)
      comptotal[c] = segP[s] + seg_out[s].valid - compbase[c];
  }
  __syncthreads();
  // C: trees, vertices, points in output order
  int ct = 0, cv = 0, cp = 0;
  for (int base = 0; base < S; base += kAsm) {
    const int s = base + (int)threadIdx.x;
    int tf = 0, vf = 0, pf = 0, c = -1, rank = 0;
    bool start = false;
    if (s < S) {
      c = seg_comp(keys, seg_start, s, cfg);
      start = (s == 0 || seg_comp(keys, seg_start, s - 1, cfg) != c);
      const bool tree_kept = comptotal[c] > cfg.min_vertices;
      rank = segP[s] - compbase[c];
      const bool vk = seg_out[s].valid && tree_kept && rank < cfg.max_vertices;
      tf = (start && tree_kept) ? 1 : 0;
      vf = vk ? 1 : 0;
      pf = vk ? seg_out[s].m : 0;
    }
    int at, av, ap;
    const int et = block_excl<kAsm>(tf, wtot, &at);
    const int ev = block_excl<kAsm>(vf, wtot, &av);
    const int ep = block_excl<kAsm>(pf, wtot, &ap);
    if (s < S) {
      const int ti = ct + et, vi = cv + ev, pi = cp + ep;
      if (tf && ti < cfg.cap_trees) tree_offsets[ti] = vi;
      if (vf && vi < cfg.cap_vertices) {
        const SegOut o = seg_out[s];
        const int row =
            cfg.H - 1 - (int)((keys[seg_start[s]] >> cfg.cb) & ((1ull << cfg.rb) - 1ull));
        vertex_int[4 * (size_t)vi + 0] = c;
        vertex_int[4 * (size_t)vi + 1] = row;
        vertex_int[4 * (size_t)vi + 2] = o.m;
        vertex_int[4 * (size_t)vi + 3] = 1;
        vertex_f64[4 * (size_t)vi + 0] = o.radius;
        vertex_f64[4 * (size_t)vi + 1] = (double)o.med.x;
        vertex_f64[4 * (size_t)vi + 2] = (double)o.med.y;
        vertex_f64[4 * (size_t)vi + 3] = (double)o.med.z;
        point_offsets[vi] = pi;
        seg_vi[s] = vi;
        seg_po[s] = pi;
      } else {
        seg_vi[s] = -1;
        seg_po[s] = 0;
      }
    }
    ct += at;
    cv += av;
    cp += ap;
  }
  if (threadIdx.x == 0) {
    const bool over = ct > cfg.cap_trees || cv > cfg.cap_vertices || cp > cfg.cap_points;
    if (!over) {
      tree_offsets[ct] = cv;
      point_offsets[cv] = cp;
    }
    out_counts[0] = ct;
    out_counts[1] = cv;
    out_counts[2] = cp;
    out_counts[3] = counters[kCntComps];
    out_counts[4] = over ? 1 : 0;
    out_counts[5] = S;
  }
}

__global__ __launch_bounds__(64) void tr_copy_kernel(const int* __restrict__ seg_start,
                                                     const int* __restrict__ counters,
                                                     const SegOut* __restrict__ seg_out,
                                                     const int* __restrict__ seg_vi,
                                                     const int* __restrict__ seg_po, TrCfgDev cfg,
                                                     const float4* __restrict__ tmp,
                                                     float4* __restrict__ points) {
  const int S = counters[kCntSegs];
  for (int s = blockIdx.x; s < S; s += gridDim.x) {
    if (seg_vi[s] < 0) continue;
    const int m = seg_out[s].m, po = seg_po[s], st = seg_start[s];
    if (po + m > cfg.cap_points) continue;
    for (int j = threadIdx.x; j < m; j += 64) points[po + j] = tmp[st + j];
  }
}

}  // namespace

extern "C" {

int sloam_trellis(sloam_ctx* ctx, const float* d_cloud, int32_t height, int32_t width,
                  const sloam_trellis_cfg* cfg, int32_t* d_labels, int32_t cap_trees,
                  int32_t cap_vertices, int32_t cap_points, int32_t* d_tree_offsets,
                  int32_t* d_vertex_int, double* d_vertex_f64, int32_t* d_point_offsets,
                  float* d_points, int32_t* d_counts) {
  if (!ctx || !cfg || !d_counts || height < 0 || width < 0) return SLOAM_ERR_INVALID_ARG;
  if (cap_trees < 0 || cap_vertices < 0 || cap_points < 0) return SLOAM_ERR_INVALID_ARG;
  if (!d_tree_offsets || !d_vertex_int || !d_vertex_f64 || !d_point_offsets ||
      (cap_points > 0 && !d_points))
    return SLOAM_ERR_INVALID_ARG;
  if (cfg->min_row_points < 0 || cfg->max_vertices < 0 || cfg->min_label_points < 0)
    return SLOAM_ERR_INVALID_ARG;
  const long long N64 = (long long)height * (long long)width;
  if (height > 65535 || width > 65535 || N64 >= (1ll << 20)) return SLOAM_ERR_TOO_LARGE;
  const int N = (int)N64;
  if (N > 0 && !d_cloud) return SLOAM_ERR_INVALID_ARG;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  TrCfgDev dc;
  dc.H = height;
  dc.W = width;
  dc.N = N;
  dc.thr = cfg->cluster_dist;
  dc.max_dist = cfg->max_dist_to_centroid;
  dc.min_label_points = cfg->min_label_points;
  dc.min_row_points = cfg->min_row_points;
  dc.min_vertices = cfg->min_vertices;
  dc.max_vertices = cfg->max_vertices;
  dc.radius_scale = cfg->radius_scale;
  dc.cap_trees = cap_trees;
  dc.cap_vertices = cap_vertices;
  dc.cap_points = cap_points;
  auto bits = [](long long v) {
    int b = 0;
    while ((1ll << b) < v) ++b;
    return b;
  };
  dc.rb = bits(height);
  dc.cb = bits(width);
  dc.key_bits = dc.rb + dc.cb + bits(N64 + 1);
  dc.pad = (1ull << dc.key_bits) - 1ull;
  const int n1 = std::max(N, 1);
  const int n_blocks = (n1 + kItems - 1) / kItems;
  const size_t sort_tmp = sizeof(int32_t) * sloam_radix_hist_ints(n1);
  auto al = [](size_t x) { return (x + 255) / 256 * 256; };
  const size_t ib = sizeof(int) * (size_t)n1;
  const size_t kb = sizeof(unsigned long long) * (size_t)n1;
  const size_t need = 12 * al(ib) + 2 * al(kb) + al(sizeof(SegOut) * (size_t)n1) +
                      al(sizeof(float4) * (size_t)n1) + 2 * al(sizeof(int) * (n_blocks + 1)) +
                      al(sizeof(int) * kCntN) + al(sort_tmp);
  void* base = nullptr;
  int st = sloam_ctx_scratch(ctx, 2, need, &base);
  if (st) return st;
  char* q = (char*)base;
  auto take = [&](size_t b) {
    char* r = q;
    q += al(b);
    return (void*)r;
  };
  int* L = (int*)take(ib);
  int* root = (int*)take(ib);
  int* cid = (int*)take(ib);
  int* label_s = (int*)take(ib);
  int* size = (int*)take(ib);
  int* seg_start = (int*)take(ib);
  int* segP = (int*)take(ib);
  int* compbase = (int*)take(ib);
  int* comptotal = (int*)take(ib);
  int* seg_vi = (int*)take(ib);
  int* seg_po = (int*)take(ib);
  (void)take(ib);
  unsigned long long* keys = (unsigned long long*)take(kb);
  unsigned long long* keys2 = (unsigned long long*)take(kb);
  unsigned long long* skeys = keys;  // the sorted keys (keys or keys2)
  SegOut* seg_out = (SegOut*)take(sizeof(SegOut) * (size_t)n1);
  float4* tmp = (float4*)take(sizeof(float4) * (size_t)n1);
  int* bc = (int*)take(sizeof(int) * (n_blocks + 1));
  int* bc2 = (int*)take(sizeof(int) * (n_blocks + 1));
  int* counters = (int*)take(sizeof(int) * kCntN);
  void* sort_buf = take(sort_tmp);
  (void)bc2;
  int* label = d_labels ? d_labels : label_s;
  const float4* cloud = reinterpret_cast<const float4*>(d_cloud);
  if (N == 0) {
    st = sloam_fill_u32(ctx, (uint32_t*)counters, kCntN, 0u);
    if (st) return st;
  }
  if (N > 0) {
    const dim3 g((N + kTB - 1) / kTB);
    {
      SloamKernelScope ks(ctx, "trellis_cc");
      hipLaunchKernelGGL(tr_init_kernel, g, dim3(kTB), 0, ctx->stream, cloud, N, L, size,
                         counters);
      hipLaunchKernelGGL(tr_union_kernel, g, dim3(kTB), 0, ctx->stream, cloud, dc, L);
      hipLaunchKernelGGL(tr_flatten_kernel, g, dim3(kTB), 0, ctx->stream, N, L, root);
      hipLaunchKernelGGL(tr_count_roots_kernel, dim3(n_blocks), dim3(kTB), 0, ctx->stream, root,
                         N, bc);
      hipLaunchKernelGGL(tr_scan_kernel, dim3(1), dim3(kTB), 0, ctx->stream, bc, n_blocks,
                         counters + kCntComps);
      hipLaunchKernelGGL(tr_rootid_kernel, dim3(n_blocks), dim3(kTB), 0, ctx->stream, root, N,
                         bc, cid);
      hipLaunchKernelGGL(tr_label_kernel, g, dim3(kTB), 0, ctx->stream, root, cid, dc, label,
                         size);
      hipLaunchKernelGGL(tr_key_kernel, g, dim3(kTB), 0, ctx->stream, label, size, dc, keys);
    }
    SLOAM_HIP(ctx, hipGetLastError());
    {
      SloamKernelScope ks(ctx, "trellis_sort");
      st = sloam_radix_sort(ctx, keys, nullptr, keys2, nullptr, N, dc.rb + dc.cb, dc.key_bits,
                            nullptr, (int32_t*)sort_buf, &skeys);
      if (st) return st;
    }
    {
      SloamKernelScope ks(ctx, "trellis_segments");
      hipLaunchKernelGGL(tr_segcount_kernel, dim3(n_blocks), dim3(kTB), 0, ctx->stream, skeys, N, dc,
                         bc);
      hipLaunchKernelGGL(tr_scan_kernel, dim3(1), dim3(kTB), 0, ctx->stream, bc, n_blocks,
                         counters + kCntSegs);
      hipLaunchKernelGGL(tr_segscatter_kernel, dim3(n_blocks), dim3(kTB), 0, ctx->stream, skeys,
                         N, dc, bc, seg_start);
    }
    {
      SloamKernelScope ks(ctx, "trellis_vertex");
      hipLaunchKernelGGL(tr_vertex_kernel, dim3(kVtxBlocks), dim3(64), 0, ctx->stream, cloud,
                         skeys, seg_start, counters, dc, seg_out, tmp);
    }
    SLOAM_HIP(ctx, hipGetLastError());
  }
  {
    SloamKernelScope ks(ctx, "trellis_assemble");
    hipLaunchKernelGGL(tr_assemble_kernel, dim3(1), dim3(kAsm), 0, ctx->stream, skeys, seg_start,
                       seg_out, counters, dc, segP, compbase, comptotal, seg_vi, seg_po,
                       d_tree_offsets, d_vertex_int, d_vertex_f64, d_point_offsets, d_counts);
    if (N > 0)
      hipLaunchKernelGGL(tr_copy_kernel, dim3(kVtxBlocks), dim3(64), 0, ctx->stream, seg_start,
                         counters, seg_out, seg_vi, seg_po, dc, tmp,
                         reinterpret_cast<float4*>(d_points));
  }
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}

}  // extern "C"
