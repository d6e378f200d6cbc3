// Tree instance front end (sloam/src/segmentation/trellis.cpp) on the GPU.
#include "trellis.h"

#include <stdexcept>

#include "gpu_bridge.h"

namespace {

sloam_trellis_cfg reference_cfg(const Instance::Params& p) {
  sloam_trellis_cfg c{};
  c.cluster_dist = 1.0f;                         // trellis.cpp:23
  c.max_dist_to_centroid = p.max_dist_to_centroid;  // trellis.cpp:90
  c.min_label_points = 80;                       // trellis.cpp:109
  c.min_row_points = 3;                          // trellis.cpp:119
  c.min_vertices = 16;                           // trellis.cpp:124
  c.max_vertices = 56;                           // trellis.cpp:125
  c.radius_scale = 1.0;                          // trellis.cpp:98
  return c;
}

void require_organized(const CloudT::Ptr& pc) {
  if (!pc || (size_t)pc->width * pc->height != pc->points.size())
    throw std::runtime_error("Instance: the tree cloud must be organized (width x height)");
}

}  // namespace

Instance::Instance() { tree_id_ = 0; }

void Instance::computeGraph(const CloudT::Ptr /*cloud*/, const CloudT::Ptr tree_cloud,
                            std::vector<std::vector<TreeVertex>>& landmarks) {
  require_organized(tree_cloud);
  if (tree_cloud->points.empty()) return;
  const auto r = sloam_gpu::trellis(tree_cloud->points, (int)tree_cloud->height,
                                    (int)tree_cloud->width, reference_cfg(params_), false);
  for (size_t t = 0; t + 1 < r.tree_offsets.size(); ++t) {
    std::vector<TreeVertex> tree;
    for (int v = r.tree_offsets[t]; v < r.tree_offsets[t + 1]; ++v) {
      TreeVertex vx;
      vx.treeId = r.vertex_int[4 * (size_t)v];
      vx.prevVertexSize = 0;
      vx.beam = 0;
      vx.isValid = r.vertex_int[4 * (size_t)v + 3] != 0;
      vx.radius = r.vertex_f64[4 * (size_t)v];
      vx.coords.x = (float)r.vertex_f64[4 * (size_t)v + 1];
      vx.coords.y = (float)r.vertex_f64[4 * (size_t)v + 2];
      vx.coords.z = (float)r.vertex_f64[4 * (size_t)v + 3];
      for (int p = r.point_offsets[v]; p < r.point_offsets[v + 1]; ++p) {
        const float* q = &r.points[4 * (size_t)p];
        vx.points.push_back(PointT{q[0], q[1], q[2], q[3]});
      }
      tree.push_back(std::move(vx));
    }
    landmarks.push_back(std::move(tree));
  }
}

void Instance::findClusters(const CloudT::Ptr pc, std::vector<int>& labels,
                            std::vector<std::vector<int>>& label_indices) {
  require_organized(pc);
  labels.clear();
  label_indices.clear();
  if (pc->points.empty()) return;
  const auto r = sloam_gpu::trellis(pc->points, (int)pc->height, (int)pc->width,
                                    reference_cfg(params_), true);
  labels.assign(r.labels.begin(), r.labels.end());
  label_indices.resize(r.n_labels);
  for (size_t i = 0; i < labels.size(); ++i)
    if (labels[i] >= 0) label_indices[labels[i]].push_back((int)i);
}
