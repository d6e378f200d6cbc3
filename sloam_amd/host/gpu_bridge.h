// Host-side bridge from the drop-in classes to the C ABI: one context per
// host thread (device 0, or $SLOAM_DEVICE), device buffers for one call at a
// time, std::vector in / out.  Failures throw std::runtime_error: the drop-in
// has no CPU fallback.
#pragma once

#include <cstdint>
#include <vector>

#include "definitions.h"
#include "sloam_gpu.h"

namespace sloam_gpu {

sloam_ctx* context();
// the device the contexts of this library's host code use (0, or
// $SLOAM_DEVICE)
int deviceIndex();

struct CellFit {
  bool valid = false;
  int kept = 0;
  double plane[4] = {0, 0, 0, 0};
  double centroid[3] = {0, 0, 0};
  VectorType features;
  std::vector<int32_t> kept_index;  // input indices of the kept points, in order
};

// sloam_ground_planes over `points`; cells r-major
std::vector<CellFit> groundPlanes(const VectorType& points, const sloam_ground_cfg& cfg);

struct LineResult {
  bool ok = false;
  float coeff[6] = {0, 0, 0, 0, 0, 0};
  int inliers = 0;
};

// sloam_fit_lines over point sets (x, y, z of each PointT)
std::vector<LineResult> fitLines(const std::vector<VectorType>& sets, const sloam_line_fit_cfg& cfg);

struct PoseProblem {
  std::vector<float> points;     // 4 floats per residual (x, y, z, landmark id bits)
  std::vector<double> cyl;       // 8 doubles per cylinder
  std::vector<double> planes;    // 4 doubles per plane
  double pose[7] = {0, 0, 0, 0, 0, 0, 1};  // t, q (x, y, z, w)
  int32_t free_mask = 0x3f;
};

struct PoseResult {
  double pose[7];
  double params[6];
  int32_t termination = 1;  // 0 CONVERGENCE
  int32_t iterations = 0;
  double cost[2] = {0, 0};
};

// sloam_pose_plan over independent problems, one batch
std::vector<PoseResult> solvePoses(const std::vector<PoseProblem>& problems,
                                   const sloam_pose_cfg& cfg);

struct TrellisResult {
  std::vector<int32_t> labels;  // per pixel (when asked for)
  int n_labels = 0;
  std::vector<int32_t> tree_offsets, vertex_int, point_offsets;
  std::vector<double> vertex_f64;
  std::vector<float> points;
};

// sloam_trellis over an organized cloud (height x width PointT, row-major)
TrellisResult trellis(const VectorType& cloud, int height, int width,
                      const sloam_trellis_cfg& cfg, bool want_labels);

struct ProjectionResult {
  std::vector<int32_t> proj;   // per point: column, row (-1, -1 if non-finite)
  std::vector<uint8_t> mask;   // H*W: 255 tree, 1 ground, 0 other / invalid
  std::vector<float> ground;   // ground points (x, y, z, intensity), input order
  std::vector<float> tree;     // dense tree cloud, NaN where not tree (n x 4)
};

// sloam_project_partition of n = xyzi.size() / 4 points with per-point labels
// (0 other, 1 ground, 2 tree); `tree` is filled when want_tree
ProjectionResult projectPartition(const std::vector<float>& xyzi,
                                  const std::vector<uint8_t>& labels, const sloam_proj_cfg& cfg,
                                  bool want_tree);

}  // namespace sloam_gpu
