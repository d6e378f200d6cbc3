// Ground plane landmark (sloam/src/objects/plane.cpp), GPU-fitted.
#include "plane.h"

#include <cmath>
#include <limits>
#include <utility>

#include "geometry.h"
#include "gpu_bridge.h"

Plane::Plane(const VectorType& points, const FeatureModelParams& fmParams) {
  features = points;
  if ((int)features.size() < fmParams.numGroundFeatures || features.empty()) {
    isValid = false;
    return;
  }
  // one cell holding every point, no bottom-k reduction (retain fraction ~0):
  // the fit of Plane::computeModel over the points in their given order
  sloam_ground_cfg cfg{};
  cfg.radii_bins = 1;
  cfg.theta_bins = 1;
  cfg.min_dist = -1.0;
  cfg.max_dist = std::numeric_limits<double>::max();
  cfg.retain_thresh = 1e-300;
  cfg.num_features = fmParams.numGroundFeatures;
  const auto cells = sloam_gpu::groundPlanes(points, cfg);
  const auto& f = cells[0];
  isValid = f.valid;
  if (!isValid) return;
  for (int i = 0; i < 4; ++i) model.plane(i) = f.plane[i];
  for (int i = 0; i < 3; ++i) model.centroid(i) = f.centroid[i];
  features = f.features;  // the first numGroundFeatures points (plane.cpp:14)
}

Plane::Plane(const PlaneParameters& fitted, VectorType feats, bool valid) {
  model = fitted;
  features = std::move(feats);
  isValid = valid;
}

Scalar Plane::distance(const PlaneParameters& tgt) const {
  return sloam_geom::plane_model_distance(model.centroid, tgt.centroid);
}

Scalar Plane::distance(const PointT& p) const {
  return sloam_geom::plane_point_distance(model.plane, p);
}

void Plane::project(const SE3& tf) {
  const Matrix4 tfm = tf.matrix();
  for (auto& p : features) {
    if (!std::isfinite(p.x) || !std::isfinite(p.y) || !std::isfinite(p.z)) continue;
    // pcl::transformPoint with a double affine: computed in double, stored as float
    const double x = p.x, y = p.y, z = p.z;
    p.x = (float)(tfm(0, 0) * x + tfm(0, 1) * y + tfm(0, 2) * z + tfm(0, 3));
    p.y = (float)(tfm(1, 0) * x + tfm(1, 1) * y + tfm(1, 2) * z + tfm(1, 3));
    p.z = (float)(tfm(2, 0) * x + tfm(2, 1) * y + tfm(2, 2) * z + tfm(2, 3));
  }
  model = projectModel(tf);
}

PlaneParameters Plane::projectModel(const SE3& tf) const {
  PlaneParameters m = model;
  sloam_geom::plane_project(tf, model.plane, model.centroid, m.plane, m.centroid);
  return m;
}
