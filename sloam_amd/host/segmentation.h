// Segmentation front end: the interface of
// sloam/include/segmentation/inference.h:55-68 (seg::Segmentation) over the
// GPU projection and partition (sloam_project_partition).
//
// The reference runs an ONNX range-image network and takes each pixel's
// class from its argmax; here the per-point labels are given (0 other,
// 1 ground, 2 tree) and a pixel takes the label of its nearest point, with
// the reference's invalid-pixel rule (range < 1 m -> 0).  The mask image is a
// plain byte image (MaskImage) instead of cv::Mat.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "definitions.h"

namespace seg {

using Point = PointT;
using Cloud = CloudT;

// rows x cols CV_8U image of the reference's mask: 255 tree, 1 ground, 0
// other or invalid
struct MaskImage {
  int rows = 0, cols = 0;
  std::vector<unsigned char> bytes;
  unsigned char* data() { return bytes.data(); }
  const unsigned char* data() const { return bytes.data(); }
  unsigned char at(int row, int col) const { return bytes[(size_t)row * cols + col]; }
};

class Segmentation {
 public:
  // inference.h:57-58.  modelFilePath names the reference's network and is
  // not read (the labels are given); fov_up / fov_down in degrees; img_d, the
  // network's input depth, is kept for source compatibility.
  explicit Segmentation(const std::string modelFilePath, const float fov_up, const float fov_down,
                        const int img_w, const int img_h, const int img_d, bool do_destagger);

  Segmentation(const Segmentation&) = delete;
  Segmentation operator=(const Segmentation&) = delete;
  using Ptr = std::shared_ptr<Segmentation>;
  using ConstPtr = std::shared_ptr<const Segmentation>;

  // Labels of the points of the next run(cloud, mask) (one per point).
  void setPointLabels(std::vector<uint8_t> labels);
  // inference.cpp:379-436: projects the cloud, fills maskImg (img_h x img_w)
  // and keeps the projection for maskCloud.  Throws if no labels were set
  // for a cloud of this size.
  void run(const Cloud::Ptr cloud, MaskImage& maskImg);
  void run(const Cloud::Ptr cloud, const std::vector<uint8_t>& labels, MaskImage& maskImg);
  // inference.cpp:230-273: the points whose pixel (in the last run's
  // projection) has mask value val, in input order; with dense, an
  // img_h x img_w cloud with NaN points elsewhere, destaggered when the
  // object was made with do_destagger.  The reference's two calls
  // (sloamNode.cpp:212,215: ground with val 1, the dense tree cloud with val
  // 255) come from the GPU partition of run() when the mask is the one run()
  // made; other values, or an edited mask, are selected from the kept
  // projection.
  void maskCloud(const Cloud::Ptr cloud, const MaskImage& mask, Cloud::Ptr& outCloud,
                 unsigned char val, bool dense = false);
  // inference.cpp:448-493: times numTests runs on the cloud (with the labels
  // set) and prints the mean latency
  void speedTest(const Cloud::Ptr cloud, size_t numTests);

 private:
  void project(const Cloud& cloud, const std::vector<uint8_t>& labels, MaskImage& maskImg);

  float fov_up_deg_, fov_down_deg_;
  int img_w_, img_h_, img_d_;
  bool do_destagger_;
  std::vector<uint8_t> labels_;
  // the last run(): projection per point, its mask, the GPU partition
  std::vector<int32_t> proj_;
  std::vector<unsigned char> mask_;
  std::vector<float> ground_, tree_;
  size_t n_points_ = 0;
};

}  // namespace seg
