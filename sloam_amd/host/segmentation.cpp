// Segmentation front end (sloam/src/segmentation/inference.cpp) with given
// labels, over the GPU projection and partition.
#include "segmentation.h"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <utility>

#include "gpu_bridge.h"

namespace seg {

Segmentation::Segmentation(const std::string /*modelFilePath*/, const float fov_up,
                           const float fov_down, const int img_w, const int img_h,
                           const int img_d, bool do_destagger)
    : fov_up_deg_(fov_up),
      fov_down_deg_(fov_down),
      img_w_(img_w),
      img_h_(img_h),
      img_d_(img_d),
      do_destagger_(do_destagger) {
  if (img_w < 1 || img_h < 1) throw std::runtime_error("Segmentation: empty image size");
}

void Segmentation::setPointLabels(std::vector<uint8_t> labels) { labels_ = std::move(labels); }

void Segmentation::run(const Cloud::Ptr cloud, MaskImage& maskImg) {
  if (!cloud) throw std::runtime_error("Segmentation::run: null cloud");
  if (labels_.size() != cloud->points.size())
    throw std::runtime_error("Segmentation::run: " + std::to_string(labels_.size()) +
                             " labels set for a cloud of " +
                             std::to_string(cloud->points.size()) +
                             " points (setPointLabels first)");
  project(*cloud, labels_, maskImg);
}

void Segmentation::run(const Cloud::Ptr cloud, const std::vector<uint8_t>& labels,
                       MaskImage& maskImg) {
  if (!cloud) throw std::runtime_error("Segmentation::run: null cloud");
  if (labels.size() != cloud->points.size())
    throw std::runtime_error("Segmentation::run: labels and cloud differ in size");
  project(*cloud, labels, maskImg);
}

void Segmentation::project(const Cloud& cloud, const std::vector<uint8_t>& labels,
                           MaskImage& maskImg) {
  const size_t n = cloud.points.size();
  // the reference's destagger walks the full image (inference.cpp:200-228)
  if (do_destagger_ && n != (size_t)img_w_ * img_h_)
    throw std::runtime_error("Segmentation: destagger needs an img_h x img_w cloud");
  std::vector<float> xyzi(4 * n);
  for (size_t i = 0; i < n; ++i) {
    const Point& p = cloud.points[i];
    xyzi[4 * i] = p.x;
    xyzi[4 * i + 1] = p.y;
    xyzi[4 * i + 2] = p.z;
    xyzi[4 * i + 3] = p.intensity;
  }
  sloam_proj_cfg cfg{};
  cfg.height = img_h_;
  cfg.width = img_w_;
  cfg.fov_up_deg = fov_up_deg_;
  cfg.fov_down_deg = fov_down_deg_;
  cfg.destagger = do_destagger_ ? 1 : 0;
  auto r = sloam_gpu::projectPartition(xyzi, labels, cfg, true);
  proj_ = std::move(r.proj);
  mask_ = std::move(r.mask);
  ground_ = std::move(r.ground);
  tree_ = std::move(r.tree);
  n_points_ = n;
  maskImg.rows = img_h_;
  maskImg.cols = img_w_;
  maskImg.bytes = mask_;
}

void Segmentation::maskCloud(const Cloud::Ptr cloud, const MaskImage& mask, Cloud::Ptr& outCloud,
                             unsigned char val, bool dense) {
  if (!cloud || cloud->points.size() != n_points_ || proj_.size() != 2 * n_points_)
    throw std::runtime_error("Segmentation::maskCloud: run() this cloud first");
  if (mask.rows != img_h_ || mask.cols != img_w_ ||
      mask.bytes.size() != (size_t)img_w_ * img_h_)
    throw std::runtime_error("Segmentation::maskCloud: mask is not img_h x img_w");
  if (!outCloud) outCloud = std::make_shared<Cloud>();
  Cloud& out = *outCloud;
  const bool same_mask =
      mask.bytes.size() == mask_.size() &&
      std::memcmp(mask.bytes.data(), mask_.data(), mask_.size()) == 0;
  auto from_floats = [&](const std::vector<float>& v) {
    out.points.resize(v.size() / 4);
    for (size_t i = 0; i < out.points.size(); ++i)
      out.points[i] = Point{v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]};
  };
  if (same_mask && val == 1 && !dense) {
    from_floats(ground_);  // the GPU partition's ground list, input order
  } else if (same_mask && val == 255 && dense) {
    from_floats(tree_);  // the GPU partition's dense (destaggered) tree cloud
  } else {
    // any other selection from the kept projection; non-finite points have
    // no pixel and take mask 0
    const float qn = std::numeric_limits<float>::quiet_NaN();
    std::vector<Point> tmp;
    tmp.reserve(dense ? n_points_ : 0);
    for (size_t i = 0; i < n_points_; ++i) {
      const int col = proj_[2 * i], row = proj_[2 * i + 1];
      const unsigned char m = col < 0 ? 0 : mask.bytes[(size_t)row * img_w_ + col];
      if (m == val)
        tmp.push_back(cloud->points[i]);
      else if (dense)
        tmp.push_back(Point{qn, qn, qn, 0.f});
    }
    if (dense && do_destagger_) {
      // even rows: x, y, z move 32 columns right; the intensity stays
      std::vector<Point> shifted = tmp;
      for (int r = 0; r < img_h_; r += 2)
        for (int c = 0; c + 32 < img_w_; ++c) {
          const Point& p = tmp[(size_t)r * img_w_ + c];
          Point& q = shifted[(size_t)r * img_w_ + c + 32];
          q.x = p.x;
          q.y = p.y;
          q.z = p.z;
        }
      tmp.swap(shifted);
    }
    out.points.swap(tmp);
  }
  if (dense) {
    out.width = (uint32_t)img_w_;
    out.height = (uint32_t)img_h_;
    out.is_dense = true;
  } else {
    out.width = (uint32_t)out.points.size();
    out.height = 1;
    out.is_dense = false;
  }
}

void Segmentation::speedTest(const Cloud::Ptr cloud, size_t numTests) {
  MaskImage m;
  run(cloud, m);  // warm-up
  const auto t0 = std::chrono::steady_clock::now();
  for (size_t i = 0; i < numTests; ++i) run(cloud, m);
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::printf("Segmentation::speedTest: %zu runs, mean %.4f ms\n", numTests,
              numTests ? ms / (double)numTests : 0.0);
}

}  // namespace seg
