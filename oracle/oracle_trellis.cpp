// TEST INFRASTRUCTURE ONLY -- CPU oracle of the tree instance front end.
//
// Restates Instance::findClusters / findTrees / computeVertexProperties
// (sloam/src/segmentation/trellis.cpp:15-132):
//   * clusters: PCL OrganizedConnectedComponentSegmentation with
//     EuclideanClusterComparator (threshold 1.0 m, not depth dependent,
//     trellis.cpp:20-28).  A pixel takes part iff its x is finite; it is joined
//     to its left and upper neighbours when sqrtf(dx*dx + dy*dy + dz*dz) < thr
//     (float).  The raster two-pass labelling with union by smallest
//     provisional label numbers the components by their first pixel in
//     row-major order; that numbering is this restatement's (PCL 1.10 is not
//     available here: its label order is parity unpinned, it only changes tree
//     order and treeId).
//   * trees (trellis.cpp:104-132): each component with more than 80 pixels;
//     rows from the last (bottom beam) to the first; the row's pixels of the
//     component in column order; rows with more than 3 points give a vertex.
//   * vertex (trellis.cpp:45-102): medians of x, y, z each at index
//     floor(n / 2) of the sorted values; points whose float distance to the
//     median point is < max_dist_to_centroid (0.2) are kept in z-ascending
//     order (the reference's last std::sort; ties by y, x, column, which is
//     libstdc++'s order for rows of at most 16 points and parity unpinned
//     above); valid iff more than one point is kept; radius = distance
//     of the first and last kept points (times radius_scale: 1 in the
//     reference's code, 0.5 reproduces its test fixtures).
//   * a tree is kept with more than 16 valid vertices, cut to the first 56.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "oracle_common.hpp"

namespace {

inline float pt_dist(const float* a, const float* b) {
  const float dx = a[0] - b[0];
  const float dy = a[1] - b[1];
  const float dz = a[2] - b[2];
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

int find_root(const std::vector<int>& run, int i) {
  while (run[i] != i) i = run[i];
  return i;
}

}  // namespace

extern "C" {

// cloud: H*W x 4 floats, row-major.  labels: H*W int32 component id, -1 where
// x is not finite.  Returns the number of components.
int oracle_trellis_clusters(const float* cloud, int H, int W, float thr, int32_t* labels) {
  const int N = H * W;
  std::vector<int> prov(N, -1), run;
  auto valid = [&](int p) { return std::isfinite(cloud[4 * (size_t)p]); };
  auto joined = [&](int a, int b) { return pt_dist(cloud + 4 * (size_t)a, cloud + 4 * (size_t)b) < thr; };
  for (int r = 0; r < H; ++r)
    for (int c = 0; c < W; ++c) {
      const int p = r * W + c;
      if (!valid(p)) continue;
      int lab = -1;
      if (c > 0 && valid(p - 1) && joined(p, p - 1)) lab = prov[p - 1];
      if (r > 0 && valid(p - W) && joined(p, p - W)) {
        if (lab < 0) {
          lab = prov[p - W];
        } else {
          const int a = find_root(run, lab), b = find_root(run, prov[p - W]);
          if (a < b)
            run[b] = a;
          else
            run[a] = b;
        }
      }
      if (lab < 0) {
        lab = (int)run.size();
        run.push_back(lab);
      }
      prov[p] = lab;
    }
  std::vector<int> map(run.size());
  int n = 0;
  for (size_t i = 0; i < run.size(); ++i)
    map[i] = run[i] == (int)i ? n++ : map[find_root(run, (int)i)];
  for (int p = 0; p < N; ++p) labels[p] = prov[p] < 0 ? -1 : map[prov[p]];
  return n;
}

// Trees from the labelled cloud.  Outputs (capacities from the caller):
//   tree_offsets [trees + 1] into vertices
//   vertex_int   [vertices][4]: treeId, row, kept points, valid (always 1)
//   vertex_f64   [vertices][4]: radius, median x, y, z
//   point_offsets[vertices + 1] into points
//   points       [points][4]: x, y, z, intensity
// counts[3] = trees, vertices, points.  Returns 0, or -1 when a capacity is
// exceeded.
int oracle_trellis_trees(const float* cloud, const int32_t* labels, int H, int W, int n_labels,
                         int min_label_points, int min_row_points, float max_dist,
                         int min_vertices, int max_vertices, double radius_scale, int cap_trees,
                         int cap_vertices, int cap_points, int32_t* tree_offsets,
                         int32_t* vertex_int, double* vertex_f64, int32_t* point_offsets,
                         float* points, int32_t* counts) {
  std::vector<int> size(n_labels, 0);
  for (int p = 0; p < H * W; ++p)
    if (labels[p] >= 0) size[labels[p]]++;
  int nt = 0, nv = 0, np = 0;
  tree_offsets[0] = 0;
  point_offsets[0] = 0;
  for (int l = 0; l < n_labels; ++l) {
    if (size[l] <= min_label_points) continue;
    struct V {
      int row;
      double radius;
      float med[3];
      std::vector<int> kept;
    };
    std::vector<V> tree;
    for (int r = H - 1; r >= 0; --r) {
      std::vector<int> beam;
      for (int c = 0; c < W; ++c)
        if (labels[r * W + c] == l) beam.push_back(r * W + c);
      if ((int)beam.size() <= min_row_points) continue;
      const int n = (int)beam.size(), mid = n / 2;
      float med[3];
      for (int k = 0; k < 3; ++k) {
        std::vector<float> v(n);
        for (int i = 0; i < n; ++i) v[i] = cloud[4 * (size_t)beam[i] + k];
        std::nth_element(v.begin(), v.begin() + mid, v.end());
        med[k] = v[mid];
      }
      // z-ascending; ties by y, then x, then column: the order the
      // reference's three successive sorts (x, y, z) leave when each is
      // stable, which libstdc++'s std::sort is for up to 16 elements (it
      // insertion-sorts short ranges)
      std::vector<int> byz = beam;
      std::sort(byz.begin(), byz.end(), [&](int a, int b) {
        const float* pa = cloud + 4 * (size_t)a;
        const float* pb = cloud + 4 * (size_t)b;
        if (pa[2] != pb[2]) return pa[2] < pb[2];
        if (pa[1] != pb[1]) return pa[1] < pb[1];
        if (pa[0] != pb[0]) return pa[0] < pb[0];
        return a < b;
      });
      V v;
      v.row = r;
      for (int p : byz)
        if (pt_dist(cloud + 4 * (size_t)p, med) < max_dist) v.kept.push_back(p);
      if (v.kept.size() <= 1) continue;
      v.radius = radius_scale * (double)pt_dist(cloud + 4 * (size_t)v.kept.front(),
                                                cloud + 4 * (size_t)v.kept.back());
      for (int k = 0; k < 3; ++k) v.med[k] = med[k];
      tree.push_back(std::move(v));
    }
    if ((int)tree.size() <= min_vertices) continue;
    if ((int)tree.size() > max_vertices) tree.resize(max_vertices);
    if (nt + 1 > cap_trees || nv + (int)tree.size() > cap_vertices) return -1;
    for (const V& v : tree) {
      if (np + (int)v.kept.size() > cap_points) return -1;
      vertex_int[4 * nv + 0] = l;
      vertex_int[4 * nv + 1] = v.row;
      vertex_int[4 * nv + 2] = (int)v.kept.size();
      vertex_int[4 * nv + 3] = 1;
      vertex_f64[4 * nv + 0] = v.radius;
      for (int k = 0; k < 3; ++k) vertex_f64[4 * nv + 1 + k] = v.med[k];
      for (int p : v.kept) {
        for (int k = 0; k < 4; ++k) points[4 * (size_t)np + k] = cloud[4 * (size_t)p + k];
        ++np;
      }
      ++nv;
      point_offsets[nv] = np;
    }
    ++nt;
    tree_offsets[nt] = nv;
  }
  counts[0] = nt;
  counts[1] = nv;
  counts[2] = np;
  return 0;
}

}  // extern "C"
