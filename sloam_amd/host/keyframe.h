// One SLOAM keyframe on the GPU: the front end (projection + ground / tree
// partition, tree instances, ground cells; one HIP graph per scan after the
// first) and the estimator (sloam::sloam::RunSloam) fed from its results.
// The per-scan path of the reference's SLOAMNode::run
// (sloam/src/core/sloamNode.cpp:235-300) with the labels given instead of the
// segmentation network.
#pragma once

#include <cstdint>
#include <vector>

#include "definitions.h"
#include "sloam.h"
#include "sloam_gpu.h"

struct KeyframeTiming {
  double upload_ms = 0, frontend_ms = 0, download_ms = 0, estimator_ms = 0, total_ms = 0;
  double build_ms = 0;      // estimator input assembled from the front end's arrays
  SloamPhaseTimes phases;   // RunSloam's phases
  int trees = 0, planes = 0;
};

class KeyframeRunner {
 public:
  // height x width organised scans; fmParams for the estimator; use_graph
  // replays the front end as a HIP graph
  KeyframeRunner(int height, int width, const FeatureModelParams& fmParams, bool use_graph = true,
                 float fov_up_deg = 22.5f, float fov_down_deg = -22.5f);
  ~KeyframeRunner();
  KeyframeRunner(const KeyframeRunner&) = delete;
  KeyframeRunner& operator=(const KeyframeRunner&) = delete;

  // xyzi: height * width points of 4 floats (sensor frame, NaN = no return);
  // labels: 0 other, 1 ground, 2 tree.  mapModels: the map trees (world
  // frame).  Returns RunSloam's result.
  bool run(const float* xyzi, const uint8_t* labels, const SE3& poseEstimate,
           const std::vector<Cylinder>& mapModels, SloamOutput& out,
           KeyframeTiming* timing = nullptr);

  sloam::sloam& estimator() { return sloam_; }

 private:
  int H_, W_;
  sloam_frontend* fe_ = nullptr;
  sloam_frontend_out buf_{};
  sloam_frontend_host host_{};
  FeatureModelParams fm_;
  sloam::sloam sloam_;
};
