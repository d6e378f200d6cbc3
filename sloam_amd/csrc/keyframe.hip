// Per-keyframe estimator on the GPU: sloam::sloam::RunSloam
// (sloam/src/core/sloam.cpp:388-532) for trees modelled by cylinder fits of
// their points, fed by the front end's device buffers and the tree fits, so
// that a keyframe needs no round trip to the host between its front end and
// its pose.
//
//   kf_models   one workgroup: computeModels (sloam.cpp:388-436) -- the
//               ground cells' gate (FromTwoVectors + Euler angles, height),
//               each tree's nearest accepted plane, range / height gates and
//               its Cylinder model from the fit (cylinder.cpp:3-173) --,
//               matchFeatures for trees (against the map trees) and planes
//               (against the previous keyframe's planes, sloam.cpp:258-285),
//               the tree / ground checks (sloam.cpp:497-500) and the pose
//               problem(s) of the solve the reference would run: the
//               residual points, the deduplicated landmark tables, the
//               initial pose, the free parameters and the LM work list
//   pose        the LM state set-up and the fused LM kernel (pose.hip) over
//               that problem set, its size read from device memory
//   kf_finish   one workgroup: the pose from the solve(s) (sloam.cpp:33-53,
//               174-255), projectModels and matchModels at it
//               (sloam.cpp:299-328, 438-451), the planes kept for the next
//               keyframe, and the results written into pinned host memory.
// Every model computation calls the same functions (host/geometry.h,
// host/definitions.h) as the host drop-in classes, so both evaluate the same
// expressions in the same order.  The estimator's state (the reference's
// firstScan_ and prevGPlanes_, sloam.h:98-106) lives on the device.
#include "common.hpp"

#include <algorithm>
#include <cstring>
#include <new>

#include "geometry.h"

namespace {

constexpr int kKB = 256;
constexpr int kChunk = 2048;  // residuals per LM work item (pose.hip's chunk)

// KfHead word indices (device work area and the result header)
enum : int {
  kHMode = 0,       // 0 first scan, 1 estimate, 2 RunSloam returns false
  kHNPlanes,        // accepted planes
  kHNLm,            // landmarks (valid tree models)
  kHTreeCheck,
  kHGroundCheck,
  kHTreeMatches,    // tree feature matches
  kHPlaneMatches,   // ground feature matches
  kHTreeK,          // problem index of the tree (XY-yaw) solve, or -1
  kHGroundK,        // problem index of the ground (Z-roll-pitch) solve, or -1
  kHNProblems,
  kHOverflow,       // 1: a capacity was exceeded (outputs truncated)
  kHNTrees,         // trees of the front end
  kHNMatchedLm,     // landmarks with a map match
  kHNMatchedPl,     // planes with a previous-plane match
  kHNCyl,           // cylinder table rows
  kHNPl,            // plane table rows
  kHSuccess,
  kHRan,            // RunSloam wrote its outputs (first scan or estimate)
  kHTreeIters,      // LM iterations of the tree (or joint) solve, -1: not run
  kHGroundIters,    // LM iterations of the ground solve, -1: not run
  kHWords = 32,
};

struct KfParamsDev {
  double maxLidarDist, defaultTreeRadius, maxTreeRadius, maxAxisTheta;
  double treeMatchThresh, addNewTreeThreshDist, minTreeModels, minGroundModels, minPlanes;
  int fpt, nf, two_step;
};

struct KfBufs {
  // inputs
  const int32_t* counts;
  const int32_t* tree_offsets;
  const int32_t* vertex_int;
  const double* vertex_f64;
  const int32_t* point_offsets;
  const float* points;
  int n_cells;
  const int32_t* cell_info;
  const double* cell_model;
  const float* cell_features;
  const double* fits;
  const double* pose;
  const double* map;
  const int32_t* n_map;
  // state: [0] first scan, [1] previous planes; prev: 8 doubles per plane
  int32_t* st;
  double* prev;
  // work area
  int32_t* head;
  int32_t* plane_cell;  // accepted planes -> cell
  int32_t* lm_tree;     // landmarks -> tree
  int32_t* lm_id;
  double* lm_model;     // 8 doubles per landmark: root, ray, radius, 0
  int32_t* lm_map;      // feature match: map tree or -1
  int32_t* lm_tab;      // cylinder table row of the match
  int32_t* pl_match;    // previous plane or -1
  int32_t* pl_tab;      // plane table row of the match
  int32_t* mlm;         // matched landmarks in order
  int32_t* mpl;         // matched planes in order
  // pose problem set (pose.hip layout)
  float* res;
  double* cyl;
  double* pl;
  int32_t* cyl_off;
  int32_t* pl_off;
  double* pose_in;
  int32_t* free_mask;
  int32_t* work;
  int32_t* first;
  int32_t* n_problems;
  // solve outputs
  const double* pose_out;
  const double* params_out;
  const int32_t* status;
  // results (pinned host memory, device addresses)
  int32_t* r_head;
  double* r_pose;
  double* r_delta;
  int32_t* r_tree;
  int32_t* r_id;
  int32_t* r_match;
  double* r_model;
  int cap_trees, cap_cells, cap_map, cap_res, cap_work;
};

__device__ inline SE3 load_pose(const double* p) {
  return SE3::fromUnitQuaternion(Quat(p[6], p[3], p[4], p[5]), Vector3(p[0], p[1], p[2]));
}

__device__ inline void store_pose(const SE3& tf, double* o) {
  const Vector3& t = tf.translation();
  const Quat& q = tf.unit_quaternion();
  o[0] = t[0];
  o[1] = t[1];
  o[2] = t[2];
  o[3] = q.x;
  o[4] = q.y;
  o[5] = q.z;
  o[6] = q.w;
}

__device__ inline PointT vertex_point(const double* vf, int v) {
  PointT p;
  p.x = (float)vf[4 * (size_t)v + 1];
  p.y = (float)vf[4 * (size_t)v + 2];
  p.z = (float)vf[4 * (size_t)v + 3];
  return p;
}

__device__ inline Vector4 cell_plane(const double* m, int c) {
  return Vector4(m[8 * (size_t)c], m[8 * (size_t)c + 1], m[8 * (size_t)c + 2],
                 m[8 * (size_t)c + 3]);
}

__device__ inline Vector3 cell_centroid(const double* m, int c) {
  return Vector3(m[8 * (size_t)c + 4], m[8 * (size_t)c + 5], m[8 * (size_t)c + 6]);
}

// Order-keeping compaction of one tile of kKB items (one per thread, called
// by every thread): the index of this thread's item among the flagged ones
// (or -1); *base advances by the tile's count.
__device__ inline int tile_compact(bool f, int* wcount, int* base) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t bits = __builtin_amdgcn_ballot_w64(f);
  const int before = __builtin_popcountll(bits & ((1ull << lane) - 1ull));
  if (lane == 0) wcount[wave] = __builtin_popcountll(bits);
  __syncthreads();
  int off = *base;
  for (int w = 0; w < wave; ++w) off += wcount[w];
  const int idx = f ? off + before : -1;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kKB / 64; ++w) t += wcount[w];
    *base += t;
  }
  __syncthreads();
  return idx;
}

// The same over the whole workgroup for every thread: the number of flagged
// items before this thread's (over all tiles so far), flagged or not.
__device__ inline int tile_scan(bool f, int* wcount, int* base) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t bits = __builtin_amdgcn_ballot_w64(f);
  const int before = __builtin_popcountll(bits & ((1ull << lane) - 1ull));
  if (lane == 0) wcount[wave] = __builtin_popcountll(bits);
  __syncthreads();
  int off = *base;
  for (int w = 0; w < wave; ++w) off += wcount[w];
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kKB / 64; ++w) t += wcount[w];
    *base += t;
  }
  __syncthreads();
  return off + before;
}

__device__ inline bool same_row(const double* a, const double* b, int n) {
  for (int i = 0; i < n; ++i)
    if (!(a[i] == b[i])) return false;
  return true;
}

// ---------------------------------------------------------------------------
// computeModels, matchFeatures, the checks and the pose problem set
// ---------------------------------------------------------------------------
// The scan `best = 0, bestDist = init; for i < n: if (d(i) < bestDist)
// {bestDist = d(i); best = i;}` by one wave: lane j scans i = j, j + 64, ...
// and the lanes' (distance, index) pairs are combined by the least distance,
// then the least index, which is the scan's choice.  Every lane returns best
// and its distance.
template <class Dist>
__device__ inline int wave_scan_min(int n, double init, Dist dist, double* best_dist) {
  const int lane = threadIdx.x & 63;
  double bv = init;
  int bi = -1;
  for (int i = lane; i < n; i += 64) {
    const double d = dist(i);
    if (d < bv) {
      bv = d;
      bi = i;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double ov = __shfl_xor(bv, o, 64);
    const int oi = __shfl_xor(bi, o, 64);
    if (oi >= 0 && (bi < 0 || ov < bv || (ov == bv && oi < bi))) {
      bv = ov;
      bi = oi;
    }
  }
  *best_dist = bv;
  return bi < 0 ? 0 : bi;
}

__global__ __launch_bounds__(kKB) void kf_models_kernel(KfBufs b, KfParamsDev P) {
  __shared__ int wcount[kKB / 64];
  __shared__ int base;
  __shared__ int overflow;
  const int tid = threadIdx.x;
  const SE3 pose = load_pose(b.pose);
  if (tid == 0) {
    base = 0;
    overflow = b.counts[4] != 0 || b.counts[0] > b.cap_trees || b.n_cells > b.cap_cells;
  }
  __syncthreads();
  const int n_cells = min(b.n_cells, b.cap_cells);
  const int n_trees = min(max(b.counts[0], 0), b.cap_trees);

  // ground cells gated as computeModels does (sloam.cpp:396-410)
  {
    const Matrix4 tfm = pose.inverse().matrix().transpose();
    for (int i0 = 0; i0 < n_cells; i0 += kKB) {
      const int c = i0 + tid;
      bool f = false;
      if (c < n_cells && b.cell_info[4 * (size_t)c] != 0)
        f = sloam_geom::plane_accepted(pose, tfm, cell_plane(b.cell_model, c),
                                       cell_centroid(b.cell_model, c));
      const int idx = tile_compact(f, wcount, &base);
      if (idx >= 0) b.plane_cell[idx] = c;
    }
  }
  const int n_planes = base;
  __syncthreads();
  if (tid == 0) base = 0;
  __syncthreads();

  // tree models (sloam.cpp:416-435 with the cylinder fits; cylinder.cpp:3-27)
  if (n_planes > 0) {
    for (int i0 = 0; i0 < n_trees; i0 += kKB) {
      const int t = i0 + tid;
      bool f = false;
      sloam_geom::CylModel M;
      int id = 0;
      if (t < n_trees) {
        const int v0 = b.tree_offsets[t];
        const int nv = b.tree_offsets[t + 1] - v0;
        if (nv >= 3) {
          const PointT v1 = vertex_point(b.vertex_f64, v0 + 1);
          // nearest accepted plane to vertex [1] (strict <, first wins)
          double bestDist = 100000;
          int best = 0;
          for (int i = 0; i < n_planes; ++i) {
            const double d =
                sloam_geom::plane_point_distance(cell_plane(b.cell_model, b.plane_cell[i]), v1);
            if (d < bestDist) {
              bestDist = d;
              best = i;
            }
          }
          const int gc = b.plane_cell[best];
          const Vector4 g = cell_plane(b.cell_model, gc);
          const Vector3 gcen = cell_centroid(b.cell_model, gc);
          if (sloam_geom::cyl_within_range(v1, gcen, P.maxLidarDist)) {
            const PointT vn2 = vertex_point(b.vertex_f64, v0 + nv - 2);
            const bool needs = sloam_geom::cyl_needs_line_fit(nv, v1, vn2);
            M.root = Vector3(v1.x, v1.y, v1.z);
            if (!needs) {
              M.radius = -1;
            } else {
              const double* m = b.fits + 8 * (size_t)t;
              sloam_geom::CylFitRecord fit;
              fit.ok = m[6] > 0.0 && m[7] >= 3.0;
              for (int i = 0; i < 3; ++i) {
                fit.root[i] = m[i];
                fit.ray[i] = m[3 + i];
              }
              fit.radius = m[6];
              sloam_geom::cyl_model_from_fit(v1, fit, P.defaultTreeRadius, M);
            }
            f = sloam_geom::cyl_valid(M, g, gcen, P.maxTreeRadius, P.maxAxisTheta);
            id = b.vertex_int[4 * (size_t)(v0 + 2)];
          }
        }
      }
      const int idx = tile_compact(f, wcount, &base);
      if (idx >= 0) {
        b.lm_tree[idx] = t;
        b.lm_id[idx] = id;
        double* o = b.lm_model + 8 * (size_t)idx;
        for (int i = 0; i < 3; ++i) {
          o[i] = M.root[i];
          o[3 + i] = M.ray[i];
        }
        o[6] = M.radius;
        o[7] = 0.0;
      }
    }
  }
  const int n_lm = base;
  __syncthreads();

  const int first = b.st[0];
  const int n_map = min(max(*b.n_map, 0), b.cap_map);
  const int n_prev = b.st[1];
  int mode = first ? 0 : (n_map == 0 || n_planes == 0 || n_lm == 0) ? 2 : 1;
  if (tid == 0) {
    int* h = b.head;
    for (int i = 0; i < kHWords; ++i) h[i] = 0;
    h[kHMode] = mode;
    h[kHNPlanes] = n_planes;
    h[kHNLm] = n_lm;
    h[kHTreeK] = -1;
    h[kHGroundK] = -1;
    h[kHOverflow] = overflow;
    h[kHNTrees] = n_trees;
    *b.n_problems = 0;
  }
  if (mode != 1) return;

  // matchFeatures: trees against the map (sloam.cpp:477-480), one wave per
  // landmark with the map trees across its lanes; planes against the previous
  // planes with the fixed 1.0 threshold (sloam.cpp:482)
  for (int l = tid >> 6; l < n_lm; l += kKB / 64) {
    const double* o = b.lm_model + 8 * (size_t)l;
    Vector3 pr, pd;
    sloam_geom::cyl_project(pose, Vector3(o[0], o[1], o[2]), Vector3(o[3], o[4], o[5]), pr, pd);
    double bestDist;
    const int best = wave_scan_min(
        n_map, P.treeMatchThresh + 100,
        [&](int i) {
          const double* m = b.map + 8 * (size_t)i;
          return sloam_geom::cyl_model_distance(Vector3(m[0], m[1], m[2]),
                                                Vector3(m[3], m[4], m[5]), pr, pd);
        },
        &bestDist);
    if ((tid & 63) == 0) b.lm_map[l] = bestDist < P.treeMatchThresh ? best : -1;
  }
  for (int p = tid; p < n_planes; p += kKB) {
    int match = -1;
    if (n_prev > 0) {
      const Vector3 pc = pose * cell_centroid(b.cell_model, b.plane_cell[p]);
      double bestDist = 1.0 + 100;
      int best = 0;
      for (int i = 0; i < n_prev; ++i) {
        const double* q = b.prev + 8 * (size_t)i;
        const double d = sloam_geom::plane_model_distance(Vector3(q[4], q[5], q[6]), pc);
        if (d < bestDist) {
          bestDist = d;
          best = i;
        }
      }
      if (bestDist < 1.0) match = best;
    }
    b.pl_match[p] = match;
  }
  __syncthreads();

  // landmark tables (consecutive matches of one landmark share a row;
  // host/sloam.cpp add_*_matches), built by the whole workgroup: the matched
  // items in order, then a new row wherever an item's row differs from the
  // previous matched item's (equal rows are equal to the run's first row, so
  // this is the sequential scan's outcome)
  if (tid == 0) base = 0;
  __syncthreads();
  for (int i0 = 0; i0 < n_lm; i0 += kKB) {
    const int l = i0 + tid;
    const int idx = tile_compact(l < n_lm && b.lm_map[l] >= 0, wcount, &base);
    if (idx >= 0) b.mlm[idx] = l;
  }
  const int nml = base;
  __syncthreads();
  if (tid == 0) base = 0;
  __syncthreads();
  for (int i0 = 0; i0 < nml; i0 += kKB) {
    const int i = i0 + tid;
    bool fresh = false;
    int j = 0;
    if (i < nml) {
      j = b.lm_map[b.mlm[i]];
      fresh = i == 0 || !same_row(b.map + 8 * (size_t)b.lm_map[b.mlm[i - 1]],
                                  b.map + 8 * (size_t)j, 7);
    }
    const int before = tile_scan(fresh, wcount, &base);
    if (i < nml) {
      const int row = fresh ? before : before - 1;
      b.lm_tab[b.mlm[i]] = row;
      if (fresh) {
        double* dst = b.cyl + 8 * (size_t)row;
        const double* src = b.map + 8 * (size_t)j;
        for (int c = 0; c < 7; ++c) dst[c] = src[c];
        dst[7] = 0.0;
      }
    }
  }
  const int ncyl = base;
  __syncthreads();
  if (tid == 0) base = 0;
  __syncthreads();
  for (int i0 = 0; i0 < n_planes; i0 += kKB) {
    const int p = i0 + tid;
    const int idx = tile_compact(p < n_planes && b.pl_match[p] >= 0, wcount, &base);
    if (idx >= 0) b.mpl[idx] = p;
  }
  const int nmp = base;
  __syncthreads();
  if (tid == 0) base = 0;
  __syncthreads();
  for (int i0 = 0; i0 < nmp; i0 += kKB) {
    const int i = i0 + tid;
    bool fresh = false;
    int j = 0;
    if (i < nmp) {
      j = b.pl_match[b.mpl[i]];
      fresh = i == 0 || !same_row(b.prev + 8 * (size_t)b.pl_match[b.mpl[i - 1]],
                                  b.prev + 8 * (size_t)j, 4);
    }
    const int before = tile_scan(fresh, wcount, &base);
    if (i < nmp) {
      const int row = fresh ? before : before - 1;
      b.pl_tab[b.mpl[i]] = row;
      if (fresh) {
        double* dst = b.pl + 4 * (size_t)row;
        const double* src = b.prev + 8 * (size_t)j;
        for (int c = 0; c < 4; ++c) dst[c] = src[c];
      }
    }
  }
  const int npl = base;
  __syncthreads();

  // checks and problem layout (sloam.cpp:497-532)
  __shared__ int s_res_t, s_res_g;
  if (tid == 0) {
    int* h = b.head;
    const int tm = nml * P.fpt, pm = nmp * P.nf;
    const bool treeCheck = (double)n_lm > P.minTreeModels && tm > 5 * P.fpt;
    const bool groundCheck = (double)pm > P.minGroundModels && (double)n_planes > P.minPlanes;
    h[kHTreeCheck] = treeCheck;
    h[kHGroundCheck] = groundCheck;
    h[kHTreeMatches] = tm;
    h[kHPlaneMatches] = pm;
    h[kHNMatchedLm] = nml;
    h[kHNMatchedPl] = nmp;
    h[kHNCyl] = ncyl;
    h[kHNPl] = npl;
    // the problem set: two-step (an XY-yaw solve on the tree matches, a
    // Z-roll-pitch solve on the ground matches, each only when its check
    // passes and it has residuals) or joint (one solve when both pass)
    int np = 0, tk = -1, gk = -1;
    bool run_t = false, run_g = false;
    if (P.two_step) {
      if (treeCheck && tm > 0) {
        tk = np++;
        run_t = true;
      }
      if (groundCheck && pm > 0) {
        gk = np++;
        run_g = true;
      }
    } else if (treeCheck && groundCheck) {
      tk = gk = np++;
      run_t = run_g = true;
    }
    const int res_t = run_t ? tm : 0, res_g = run_g ? pm : 0;
    if (res_t + res_g > b.cap_res) {
      h[kHOverflow] = 1;
      np = 0;
      tk = gk = -1;
      run_t = run_g = false;
    }
    h[kHTreeK] = tk;
    h[kHGroundK] = gk;
    h[kHNProblems] = np;
    s_res_t = run_t ? tm : 0;
    s_res_g = run_g ? pm : 0;
    // residual ranges, table ranges, initial poses, free parameters
    int res_off[3] = {0, 0, 0}, cyl_off[3] = {0, 0, 0}, pl_off[3] = {0, 0, 0};
    if (P.two_step) {
      int k = 0;
      if (run_t) {
        res_off[k + 1] = res_off[k] + tm;
        cyl_off[k + 1] = ncyl;
        pl_off[k + 1] = 0;
        b.free_mask[k] = (1 << 0) | (1 << 1) | (1 << 5);  // x, y, yaw
        ++k;
      }
      if (run_g) {
        res_off[k + 1] = res_off[k] + pm;
        cyl_off[k + 1] = cyl_off[k];
        pl_off[k + 1] = pl_off[k] + npl;
        b.free_mask[k] = (1 << 2) | (1 << 3) | (1 << 4);  // z, roll, pitch
        ++k;
      }
    } else if (np == 1) {
      res_off[1] = tm + pm;
      cyl_off[1] = ncyl;
      pl_off[1] = npl;
      b.free_mask[0] = 0x3f;
    }
    int nw = 0;
    for (int k = 0; k < np; ++k) {
      b.first[k] = nw;
      const int rb = res_off[k], re = res_off[k + 1];
      if (re == rb && nw < b.cap_work) {
        int* w = b.work + 4 * (size_t)nw++;
        w[0] = k;
        w[1] = rb;
        w[2] = re;
        w[3] = 0;
      }
      for (int r = rb; r < re && nw < b.cap_work; r += kChunk) {
        int* w = b.work + 4 * (size_t)nw++;
        w[0] = k;
        w[1] = r;
        w[2] = min(re, r + kChunk);
        w[3] = 0;
      }
      b.cyl_off[k] = cyl_off[k];
      b.pl_off[k] = pl_off[k];
      const Vector3& t = pose.translation();
      const Quat& q = pose.unit_quaternion();
      double* x0 = b.pose_in + 7 * (size_t)k;
      x0[0] = t[0];
      x0[1] = t[1];
      x0[2] = t[2];
      x0[3] = q.x;
      x0[4] = q.y;
      x0[5] = q.z;
      x0[6] = q.w;
    }
    b.first[np] = nw;
    b.cyl_off[np] = cyl_off[np];
    b.pl_off[np] = pl_off[np];
    *b.n_problems = np;
  }
  __syncthreads();
  // residual points: the tree features (the first features_per_tree points
  // of the tree in vertex order, zeros past its last point) of each matched
  // landmark, then the ground features of each matched plane
  const int res_t = s_res_t, res_g = s_res_g;
  for (int r = tid; r < res_t + res_g; r += kKB) {
    float x = 0.f, y = 0.f, z = 0.f;
    int32_t id;
    if (r < res_t) {
      const int s = r / P.fpt, j = r % P.fpt;
      const int l = b.mlm[s];
      const int t = b.lm_tree[l];
      const int pb = b.point_offsets[b.tree_offsets[t]];
      const int pe = b.point_offsets[b.tree_offsets[t + 1]];
      if (pb + j < pe) {
        const float* q = b.points + 4 * (size_t)(pb + j);
        x = q[0];
        y = q[1];
        z = q[2];
      }
      id = b.lm_tab[l];
    } else {
      const int rr = r - res_t;
      const int s = rr / P.nf, j = rr % P.nf;
      const int p = b.mpl[s];
      const int c = b.plane_cell[p];
      const float* q = b.cell_features + 4 * ((size_t)c * P.nf + j);
      x = q[0];
      y = q[1];
      z = q[2];
      id = -(b.pl_tab[p] + 1);
    }
    float* o = b.res + 4 * (size_t)r;
    o[0] = x;
    o[1] = y;
    o[2] = z;
    o[3] = __int_as_float(id);
  }
}

// ---------------------------------------------------------------------------
// the pose from the solves, projectModels, matchModels, next planes, results
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kKB) void kf_finish_kernel(KfBufs b, KfParamsDev P) {
  __shared__ double s_pose[7], s_delta[7];
  __shared__ int s_success;
  const int tid = threadIdx.x;
  const int* h = b.head;
  const int mode = h[kHMode];
  int32_t* r = b.r_head;
  if (mode == 2) {
    // RunSloam returns false and changes nothing (sloam.cpp:476-486)
    if (tid == 0) {
      for (int i = 0; i < kHWords; ++i) r[i] = h[i];
      r[kHRan] = 0;
      r[kHSuccess] = 0;
      r[kHTreeIters] = r[kHGroundIters] = -1;
    }
    return;
  }
  const SE3 est = load_pose(b.pose);
  if (tid == 0) {
    SE3 cur = est, delta;
    bool success = true;
    if (mode == 1) {
      if (P.two_step) {
        // TwoStepOptimizePose (sloam.cpp:33-53): x, y, yaw from the tree
        // solve, z, roll, pitch from the ground solve, each only when it ran
        // and converged
        double p0[6];
        {
          const Vector3& t = est.translation();
          const Quat& q = est.unit_quaternion();
          const double quat[4] = {q.w, q.x, q.y, q.z};
          p0[0] = t[0];
          p0[1] = t[1];
          p0[2] = t[2];
          quaternionToAngleAxis(quat, p0 + 3);
        }
        double treeOut[3] = {p0[0], p0[1], p0[5]};
        double groundOut[3] = {p0[2], p0[3], p0[4]};
        const int tk = h[kHTreeK], gk = h[kHGroundK];
        if (tk >= 0 && b.status[2 * tk] == 0) {
          treeOut[0] = b.params_out[6 * tk];
          treeOut[1] = b.params_out[6 * tk + 1];
          treeOut[2] = b.params_out[6 * tk + 5];
        }
        if (gk >= 0 && b.status[2 * gk] == 0) {
          groundOut[0] = b.params_out[6 * gk + 2];
          groundOut[1] = b.params_out[6 * gk + 3];
          groundOut[2] = b.params_out[6 * gk + 4];
        }
        const double angleAxis[3] = {groundOut[1], groundOut[2], treeOut[2]};
        double q[4];
        angleAxisToQuaternion(angleAxis, q);
        cur.setQuaternion(Quat(q[0], q[1], q[2], q[3]));
        cur.translation()[0] = treeOut[0];
        cur.translation()[1] = treeOut[1];
        cur.translation()[2] = groundOut[0];
      } else if (h[kHNProblems] == 1) {
        // OptimizePose (sloam.cpp:174-255): the pose only on convergence
        success = b.status[0] == 0;
        if (success) {
          const double* x = b.pose_out;
          delta.setQuaternion(Quat(x[6], x[3], x[4], x[5]));
          delta.translation() = Vector3(x[0], x[1], x[2]);
          cur = delta;
        }
      }
    }
    store_pose(cur, s_pose);
    store_pose(delta, s_delta);
    s_success = success;
  }
  __syncthreads();
  const SE3 cur = load_pose(s_pose);
  const int n_lm = h[kHNLm], n_planes = h[kHNPlanes];
  const int n_map = min(max(*b.n_map, 0), b.cap_map);
  // projectModels at the pose, then matchModels (sloam.cpp:299-328), one
  // wave per landmark with the map trees across its lanes
  for (int l = tid >> 6; l < n_lm; l += kKB / 64) {
    const double* o = b.lm_model + 8 * (size_t)l;
    Vector3 pr, pd;
    sloam_geom::cyl_project(cur, Vector3(o[0], o[1], o[2]), Vector3(o[3], o[4], o[5]), pr, pd);
    int match = -1;
    if (mode == 1 && n_map > 0) {
      double bestDist;
      const int best = wave_scan_min(
          n_map, P.treeMatchThresh + 100,
          [&](int i) {
            const double* m = b.map + 8 * (size_t)i;
            return sloam_geom::cyl_model_distance(Vector3(m[0], m[1], m[2]),
                                                  Vector3(m[3], m[4], m[5]), pr, pd);
          },
          &bestDist);
      if (bestDist < P.addNewTreeThreshDist) match = best;
    }
    if ((tid & 63) != 0) continue;
    double* ro = b.r_model + 8 * (size_t)l;
    for (int i = 0; i < 3; ++i) {
      ro[i] = pr[i];
      ro[3 + i] = pd[i];
    }
    ro[6] = o[6];
    ro[7] = 0.0;
    b.r_tree[l] = b.lm_tree[l];
    b.r_id[l] = b.lm_id[l];
    b.r_match[l] = match;
  }
  // the planes kept for the next keyframe, projected at the pose
  for (int p = tid; p < n_planes; p += kKB) {
    const int c = b.plane_cell[p];
    Vector4 po;
    Vector3 co;
    sloam_geom::plane_project(cur, cell_plane(b.cell_model, c), cell_centroid(b.cell_model, c), po,
                              co);
    double* q = b.prev + 8 * (size_t)p;
    for (int i = 0; i < 4; ++i) q[i] = po[i];
    for (int i = 0; i < 3; ++i) q[4 + i] = co[i];
    q[7] = 0.0;
  }
  if (tid == 0) {
    b.st[0] = 0;
    b.st[1] = n_planes;
    for (int i = 0; i < kHWords; ++i) r[i] = h[i];
    r[kHRan] = 1;
    r[kHSuccess] = s_success;
    int ti = -1, gi = -1;
    if (mode == 1) {
      if (P.two_step) {
        if (h[kHTreeK] >= 0) ti = b.status[2 * h[kHTreeK] + 1];
        if (h[kHGroundK] >= 0) gi = b.status[2 * h[kHGroundK] + 1];
      } else if (h[kHNProblems] == 1) {
        ti = b.status[1];
      }
    }
    r[kHTreeIters] = ti;
    r[kHGroundIters] = gi;
    for (int i = 0; i < 7; ++i) {
      b.r_pose[i] = s_pose[i];
      b.r_delta[i] = s_delta[i];
    }
  }
}

// Tree t's points (the front end's points of its vertices, contiguous) as
// offsets relative to the first tree's first point and packed xyz, for at
// most cap_trees trees and cap_points points; *n_out = the trees in use.
__global__ __launch_bounds__(256) void kf_tree_points_kernel(
    const int32_t* __restrict__ counts, const int32_t* __restrict__ tree_offsets,
    const int32_t* __restrict__ point_offsets, const float4* __restrict__ points, int cap_trees,
    long long cap_points, int32_t* __restrict__ off_out, float* __restrict__ xyz,
    int32_t* __restrict__ n_out) {
  const int n = min(max(counts[0], 0), cap_trees);
  const int base = point_offsets[tree_offsets[0]];
  const long long total = min((long long)point_offsets[tree_offsets[n]] - base, cap_points);
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i <= n || i < total;
       i += stride) {
    if (i <= n) off_out[i] = point_offsets[tree_offsets[i]] - base;
    if (i < total) {
      const float4 p = points[base + i];
      xyz[3 * i] = p.x;
      xyz[3 * i + 1] = p.y;
      xyz[3 * i + 2] = p.z;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *n_out = n;
}

__global__ void kf_reset_kernel(int32_t* st) {
  st[0] = 1;
  st[1] = 0;
}

}  // namespace

struct sloam_kf_estimator {
  sloam_ctx* ctx = nullptr;
  sloam_kf_params params{};
  int cap_trees = 0, cap_cells = 0, cap_map = 0, cap_res = 0, cap_work = 0;
  void* dev = nullptr;   // every device buffer, one allocation
  void* host = nullptr;  // pinned results
  void* host_dev = nullptr;
  KfBufs bufs{};         // device pointers into dev (inputs set per exec)
  double* pose_out = nullptr;
  double* params_out = nullptr;
  int32_t* status = nullptr;
  double* cost = nullptr;
  void* lm_state = nullptr;
  void* cyl_tab = nullptr;  // pose landmark tables, FP64 size
  void* pl_tab = nullptr;
  size_t r_off[7] = {};  // byte offsets of the result regions
};

extern "C" {

int sloam_kf_estimator_create(sloam_ctx* ctx, const sloam_kf_params* p, int32_t cap_trees,
                              int32_t cap_cells, int32_t cap_map, sloam_kf_estimator** out) {
  if (!ctx || !p || !out || cap_trees < 1 || cap_cells < 1 || cap_map < 1)
    return SLOAM_ERR_INVALID_ARG;
  if (p->features_per_tree < 1 || p->num_ground_features < 1 || cap_trees > (1 << 16) ||
      cap_cells > (1 << 16) || cap_map > (1 << 20) || p->pose_precision < 0 ||
      p->pose_precision > 1)
    return SLOAM_ERR_INVALID_ARG;
  *out = nullptr;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  sloam_kf_estimator* e = new (std::nothrow) sloam_kf_estimator();
  if (!e) return SLOAM_ERR_ALLOC;
  e->ctx = ctx;
  e->params = *p;
  e->cap_trees = cap_trees;
  e->cap_cells = cap_cells;
  e->cap_map = cap_map;
  e->cap_res = cap_trees * p->features_per_tree + cap_cells * p->num_ground_features;
  e->cap_work = e->cap_res / kChunk + 4;
  // device layout (bytes, each region 256-byte aligned)
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t T = (size_t)cap_trees, Cc = (size_t)cap_cells;
  const size_t o_st = take(4 * 4), o_prev = take(8 * 8 * Cc), o_head = take(4 * kHWords),
               o_plane_cell = take(4 * Cc), o_lm_tree = take(4 * T), o_lm_id = take(4 * T),
               o_lm_model = take(8 * 8 * T), o_lm_map = take(4 * T), o_lm_tab = take(4 * T),
               o_pl_match = take(4 * Cc), o_pl_tab = take(4 * Cc), o_mlm = take(4 * T),
               o_mpl = take(4 * Cc), o_res = take(16 * (size_t)e->cap_res),
               o_cyl = take(8 * 8 * T), o_pl = take(8 * 4 * Cc), o_cyl_off = take(4 * 4),
               o_pl_off = take(4 * 4), o_pose_in = take(8 * 14), o_free = take(4 * 2),
               o_work = take(16 * (size_t)e->cap_work), o_first = take(4 * 4),
               o_np = take(4), o_pose_out = take(8 * 14), o_params = take(8 * 12),
               o_status = take(4 * 4), o_cost = take(8 * 4),
               o_lm_state = take(2 * sloam_pose_state_bytes()), o_cylf = take(8 * 8 * T),
               o_plf = take(8 * 4 * Cc);
  int st = sloam_dev_alloc(ctx, &e->dev, off);
  if (st) {
    delete e;
    return st;
  }
  // results: header, pose, delta, landmark trees, ids, matches, models
  size_t roff = 0;
  auto rtake = [&](size_t bytes) {
    const size_t o = roff;
    roff += (bytes + 255) / 256 * 256;
    return o;
  };
  e->r_off[0] = rtake(4 * kHWords);
  e->r_off[1] = rtake(8 * 7);
  e->r_off[2] = rtake(8 * 7);
  e->r_off[3] = rtake(4 * T);
  e->r_off[4] = rtake(4 * T);
  e->r_off[5] = rtake(4 * T);
  e->r_off[6] = rtake(8 * 8 * T);
  hipError_t he = hipHostMalloc(&e->host, roff, hipHostMallocDefault);
  if (he == hipSuccess) he = hipHostGetDevicePointer(&e->host_dev, e->host, 0);
  if (he != hipSuccess) {
    sloam_record_hip_error(ctx, he, "hipHostMalloc(keyframe results)");
    sloam_kf_estimator_destroy(e);
    return SLOAM_ERR_ALLOC;
  }
  std::memset(e->host, 0, roff);
  char* d = static_cast<char*>(e->dev);
  char* hd = static_cast<char*>(e->host_dev);
  KfBufs& b = e->bufs;
  b.st = reinterpret_cast<int32_t*>(d + o_st);
  b.prev = reinterpret_cast<double*>(d + o_prev);
  b.head = reinterpret_cast<int32_t*>(d + o_head);
  b.plane_cell = reinterpret_cast<int32_t*>(d + o_plane_cell);
  b.lm_tree = reinterpret_cast<int32_t*>(d + o_lm_tree);
  b.lm_id = reinterpret_cast<int32_t*>(d + o_lm_id);
  b.lm_model = reinterpret_cast<double*>(d + o_lm_model);
  b.lm_map = reinterpret_cast<int32_t*>(d + o_lm_map);
  b.lm_tab = reinterpret_cast<int32_t*>(d + o_lm_tab);
  b.pl_match = reinterpret_cast<int32_t*>(d + o_pl_match);
  b.pl_tab = reinterpret_cast<int32_t*>(d + o_pl_tab);
  b.mlm = reinterpret_cast<int32_t*>(d + o_mlm);
  b.mpl = reinterpret_cast<int32_t*>(d + o_mpl);
  b.res = reinterpret_cast<float*>(d + o_res);
  b.cyl = reinterpret_cast<double*>(d + o_cyl);
  b.pl = reinterpret_cast<double*>(d + o_pl);
  b.cyl_off = reinterpret_cast<int32_t*>(d + o_cyl_off);
  b.pl_off = reinterpret_cast<int32_t*>(d + o_pl_off);
  b.pose_in = reinterpret_cast<double*>(d + o_pose_in);
  b.free_mask = reinterpret_cast<int32_t*>(d + o_free);
  b.work = reinterpret_cast<int32_t*>(d + o_work);
  b.first = reinterpret_cast<int32_t*>(d + o_first);
  b.n_problems = reinterpret_cast<int32_t*>(d + o_np);
  e->pose_out = reinterpret_cast<double*>(d + o_pose_out);
  e->params_out = reinterpret_cast<double*>(d + o_params);
  e->status = reinterpret_cast<int32_t*>(d + o_status);
  e->cost = reinterpret_cast<double*>(d + o_cost);
  e->lm_state = d + o_lm_state;
  e->cyl_tab = d + o_cylf;
  e->pl_tab = d + o_plf;
  b.pose_out = e->pose_out;
  b.params_out = e->params_out;
  b.status = e->status;
  b.r_head = reinterpret_cast<int32_t*>(hd + e->r_off[0]);
  b.r_pose = reinterpret_cast<double*>(hd + e->r_off[1]);
  b.r_delta = reinterpret_cast<double*>(hd + e->r_off[2]);
  b.r_tree = reinterpret_cast<int32_t*>(hd + e->r_off[3]);
  b.r_id = reinterpret_cast<int32_t*>(hd + e->r_off[4]);
  b.r_match = reinterpret_cast<int32_t*>(hd + e->r_off[5]);
  b.r_model = reinterpret_cast<double*>(hd + e->r_off[6]);
  b.cap_trees = cap_trees;
  b.cap_cells = cap_cells;
  b.cap_map = cap_map;
  b.cap_res = e->cap_res;
  b.cap_work = e->cap_work;
  st = sloam_kf_estimator_reset(e);
  if (st) {
    sloam_kf_estimator_destroy(e);
    return st;
  }
  *out = e;
  return SLOAM_OK;
}

void sloam_kf_estimator_destroy(sloam_kf_estimator* e) {
  if (!e) return;
  if (e->ctx) (void)hipStreamSynchronize(e->ctx->stream);
  sloam_dev_free(e->dev);
  if (e->host) (void)hipHostFree(e->host);
  delete e;
}

int sloam_kf_estimator_reset(sloam_kf_estimator* e) {
  if (!e) return SLOAM_ERR_INVALID_ARG;
  sloam_ctx* ctx = e->ctx;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  hipLaunchKernelGGL(kf_reset_kernel, dim3(1), dim3(1), 0, ctx->stream, e->bufs.st);
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}

int sloam_kf_estimator_exec(sloam_kf_estimator* e, const sloam_kf_inputs* in) {
  if (!e || !in || !in->trellis_counts || !in->tree_offsets || !in->vertex_int ||
      !in->vertex_f64 || !in->point_offsets || !in->points || !in->cell_info || !in->cell_model ||
      !in->cell_features || !in->tree_fits || !in->pose || !in->map || !in->n_map)
    return SLOAM_ERR_INVALID_ARG;
  if (in->n_cells < 0 || in->num_features != e->params.num_ground_features)
    return SLOAM_ERR_INVALID_ARG;
  sloam_ctx* ctx = e->ctx;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  KfBufs b = e->bufs;
  b.counts = in->trellis_counts;
  b.tree_offsets = in->tree_offsets;
  b.vertex_int = in->vertex_int;
  b.vertex_f64 = in->vertex_f64;
  b.point_offsets = in->point_offsets;
  b.points = in->points;
  b.n_cells = in->n_cells;
  b.cell_info = in->cell_info;
  b.cell_model = in->cell_model;
  b.cell_features = in->cell_features;
  b.fits = in->tree_fits;
  b.pose = in->pose;
  b.map = in->map;
  b.n_map = in->n_map;
  const sloam_kf_params& p = e->params;
  KfParamsDev P;
  P.maxLidarDist = p.max_lidar_dist;
  P.defaultTreeRadius = p.default_tree_radius;
  P.maxTreeRadius = p.max_tree_radius;
  P.maxAxisTheta = p.max_axis_theta;
  P.treeMatchThresh = p.tree_match_thresh;
  P.addNewTreeThreshDist = p.add_new_tree_thresh_dist;
  P.minTreeModels = p.min_tree_models;
  P.minGroundModels = p.min_ground_models;
  P.minPlanes = p.min_planes;
  P.fpt = p.features_per_tree;
  P.nf = p.num_ground_features;
  P.two_step = p.two_step != 0;
  {
    SloamKernelScope ks(ctx, "kf_models");
    hipLaunchKernelGGL(kf_models_kernel, dim3(1), dim3(kKB), 0, ctx->stream, b, P);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  // Ceres settings of the reference's solves (sloam.cpp:88-97, 219-226)
  sloam_pose_cfg c{};
  c.max_iters = 50;
  c.stop_on_convergence = 1;
  c.huber_delta = 0.1;
  c.function_tolerance = 1e-6;
  c.gradient_tolerance = 1e-10;
  c.parameter_tolerance = 1e-8;
  c.initial_radius = 1e4;
  c.min_relative_decrease = 1e-3;
  c.parametrization = P.two_step ? 1 : 0;
  c.precision = e->params.pose_precision;
  {
    SloamKernelScope ks(ctx, "kf_pose");
    const int st = sloam_pose_fused_device(
        ctx, 2, b.n_problems, b.res, b.work, b.first, b.cyl, b.cyl_off, b.pl, b.pl_off, b.pose_in,
        b.free_mask, &c, e->lm_state, e->cyl_tab, e->pl_tab, e->pose_out, e->params_out, e->status,
        e->cost);
    if (st) return st;
  }
  {
    SloamKernelScope ks(ctx, "kf_finish");
    hipLaunchKernelGGL(kf_finish_kernel, dim3(1), dim3(kKB), 0, ctx->stream, b, P);
  }
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}

int sloam_kf_tree_points(sloam_ctx* ctx, const sloam_frontend_out* fe, int32_t cap_trees,
                         int64_t cap_points, int32_t* d_offsets, float* d_xyz, int32_t* d_n) {
  if (!ctx || !fe || !d_offsets || !d_xyz || !d_n || cap_trees < 1 || cap_points < 1)
    return SLOAM_ERR_INVALID_ARG;
  if (cap_trees > fe->cap_trees || cap_points > fe->cap_points) return SLOAM_ERR_INVALID_ARG;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  const long long work = std::max<long long>((long long)cap_trees + 1, cap_points);
  const int blocks = (int)std::min<long long>((work + 255) / 256, 1024);
  hipLaunchKernelGGL(kf_tree_points_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                     fe->trellis_counts, fe->tree_offsets, fe->point_offsets,
                     reinterpret_cast<const float4*>(fe->points), cap_trees, (long long)cap_points,
                     d_offsets, d_xyz, d_n);
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}

int sloam_kf_estimator_result(const sloam_kf_estimator* e, sloam_kf_result_view* out) {
  if (!e || !out) return SLOAM_ERR_INVALID_ARG;
  const char* h = static_cast<const char*>(e->host);
  out->head = reinterpret_cast<const int32_t*>(h + e->r_off[0]);
  out->pose = reinterpret_cast<const double*>(h + e->r_off[1]);
  out->delta = reinterpret_cast<const double*>(h + e->r_off[2]);
  out->lm_tree = reinterpret_cast<const int32_t*>(h + e->r_off[3]);
  out->lm_id = reinterpret_cast<const int32_t*>(h + e->r_off[4]);
  out->lm_match = reinterpret_cast<const int32_t*>(h + e->r_off[5]);
  out->lm_model = reinterpret_cast<const double*>(h + e->r_off[6]);
  return SLOAM_OK;
}

}  // extern "C"
