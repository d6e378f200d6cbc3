// Tree instance front end: the interface of
// sloam/include/segmentation/trellis.h:15-75 (class Instance).  computeGraph
// runs the organized connected components and the per-beam vertices of the
// whole tree cloud in one sloam_trellis call.
#pragma once

#include <vector>

#include "definitions.h"

class Instance {
 public:
  // trellis.h:33-39; beam_cluster_threshold, min_vertex_size,
  // min_landmark_size and min_landmark_height are read by no code path of
  // the reference and are kept for source compatibility
  struct Params {
    float beam_cluster_threshold = 0.1f;
    float max_dist_to_centroid = 0.2f;
    int min_vertex_size = 2;
    int min_landmark_size = 4;
    float min_landmark_height = 1.0f;
  };

  explicit Instance();
  const Params& params() const { return params_; }
  void set_params(const Params& params) { params_ = params; }
  void reset_tree_id() { tree_id_ = 0; }

  // trellis.cpp:134-141: clusters of the organized tree cloud (x NaN = empty)
  // and their trees, bottom-up vertices.  `cloud` is unused, as in the
  // reference.
  void computeGraph(const CloudT::Ptr cloud, const CloudT::Ptr tree_cloud,
                    std::vector<std::vector<TreeVertex>>& landmarks);

  // trellis.cpp:15-43: per-pixel component ids (-1 = empty) and the pixel
  // indices of each component
  void findClusters(const CloudT::Ptr pc, std::vector<int>& labels,
                    std::vector<std::vector<int>>& label_indices);

 private:
  Params params_;
  int tree_id_ = 0;
};
