// Tree trunk landmark (sloam/src/objects/cylinder.cpp), axis fitted on the GPU.
#include "cylinder.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <utility>

#include "geometry.h"
#include "gpu_bridge.h"

sloam_line_fit_cfg Cylinder::lineFitCfg() {
  // SACSegmentation settings of Cylinder::computeModel (cylinder.cpp:116-124)
  // and the library defaults (max 50 iterations, p = 0.99, seed 12345)
  sloam_line_fit_cfg c{};
  c.threshold = 0.25;
  c.probability = 0.99;
  c.max_iterations = 50;
  c.optimize = 1;
  c.seed = 12345;
  return c;
}

bool Cylinder::withinRange(const std::vector<TreeVertex>& vertices, const Plane& gplane,
                           const FeatureModelParams& fmParams) {
  return sloam_geom::cyl_within_range(vertices[1].coords, gplane.model.centroid,
                                      fmParams.maxLidarDist);
}

bool Cylinder::needsLineFit(const std::vector<TreeVertex>& v) {
  if (v.size() < 3) return false;
  return sloam_geom::cyl_needs_line_fit((int)v.size(), v[1].coords, v[v.size() - 2].coords);
}

Cylinder::Cylinder(const std::vector<TreeVertex> vertices, const Plane& gplane,
                   const FeatureModelParams& fmParams) {
  LineFit fit;
  if (vertices.size() >= 3 && withinRange(vertices, gplane, fmParams) && needsLineFit(vertices)) {
    VectorType medians;
    for (const auto& v : vertices) medians.push_back(v.coords);
    const auto r = sloam_gpu::fitLines({medians}, lineFitCfg());
    fit.ok = r[0].ok;
    fit.inliers = r[0].inliers;
    std::copy(r[0].coeff, r[0].coeff + 6, fit.coeff);
  }
  build(vertices, gplane, fmParams, &fit, nullptr);
}

Cylinder::Cylinder(const std::vector<TreeVertex>& vertices, const Plane& gplane,
                   const FeatureModelParams& fmParams, const LineFit& fit) {
  build(vertices, gplane, fmParams, &fit, nullptr);
}

Cylinder::Cylinder(const std::vector<TreeVertex>& vertices, const Plane& gplane,
                   const FeatureModelParams& fmParams, const CylinderFit& fit) {
  build(vertices, gplane, fmParams, nullptr, &fit);
}

Cylinder::Cylinder(const CylinderParameters& m, size_t treeId, VectorType feats) {
  model = m;
  id = treeId;
  features = std::move(feats);
  isValid = true;
}

void Cylinder::build(const std::vector<TreeVertex>& vertices, const Plane& gplane,
                     const FeatureModelParams& fmParams, const LineFit* fit,
                     const CylinderFit* cfit) {
  isValid = false;
  rejectReason = kRejectNoModel;
  if (vertices.size() < 3) return;  // the reference indexes vertices [1] and [2]
  if (withinRange(vertices, gplane, fmParams)) {
    computeModel(vertices, fmParams.defaultTreeRadius, fmParams.featuresPerTree, fit, cfit);
    const bool validZ = groundBasedRoot(gplane);
    const bool validNorm = model.root.norm() > 0.01;
    const bool validRadius = model.radius > 0.0;
    const int fails = filterFails(fmParams.maxTreeRadius, fmParams.maxAxisTheta, gplane);
    const bool validTree = fails == 0;
    isValid = validZ && validRadius && validTree && validNorm;
    rejectReason = (validZ ? 0 : kRejectRootOffGround) | (validNorm ? 0 : kRejectRootAtOrigin) |
                   (validRadius ? 0 : kRejectRadiusNotPositive) |
                   ((fails & sloam_geom::kCylNoModel) ? kRejectNoModel : 0) |
                   ((fails & sloam_geom::kCylRadiusCap) ? kRejectRadiusCap : 0) |
                   ((fails & sloam_geom::kCylAxisTilt) ? kRejectAxisTilt : 0);
  } else {
    rejectReason = kRejectOutOfRange;
  }
}

PointT Cylinder::asPoint(const Scalar height) {
  const Scalar t = (height - model.root[2]) / model.ray[2];
  const Vector3 p = model.root + t * model.ray;
  PointT out;
  out.x = (float)p[0];
  out.y = (float)p[1];
  out.z = 0;
  return out;
}

bool Cylinder::groundBasedRoot(const Plane& gplane) {
  sloam_geom::CylModel m{model.root, model.ray, model.radius};
  const bool near = sloam_geom::cyl_ground_based_root(gplane.model.plane, gplane.model.centroid, m);
  model.root = m.root;
  return near;
}

int Cylinder::filterFails(const Scalar& maxTreeRadius, const Scalar& maxAxisTheta,
                          const Plane& gplane) {
  const sloam_geom::CylModel m{model.root, model.ray, model.radius};
  return sloam_geom::cyl_filter_fails(m, maxTreeRadius, maxAxisTheta, gplane.model.plane);
}

void Cylinder::computeModel(const std::vector<TreeVertex>& landmarkVtxs,
                            const Scalar defaultTreeRadius, const Scalar featuresPerTree,
                            const LineFit* fit, const CylinderFit* cfit) {
  const TreeVertex& firstVtx = landmarkVtxs[2];
  // a model that is computed to the end keeps only its first featuresPerTree
  // features (the resize below), so only those are copied; the early returns
  // keep every point, as the reference's do
  const bool toEnd = needsLineFit(landmarkVtxs) &&
                     (cfit ? (cfit->ok && cfit->radius > 0.0) : (fit->ok && fit->inliers != 0));
  const size_t keep =
      toEnd ? (size_t)featuresPerTree : std::numeric_limits<size_t>::max();
  std::vector<float> validRadii;
  for (const auto& vtx : landmarkVtxs) {
    for (auto p : vtx.points) {
      if (features.size() >= keep) break;
      p.intensity = (float)firstVtx.treeId;
      features.push_back(p);
    }
    model.radii.push_back(vtx.radius);
    if (vtx.points.size() > 3) validRadii.push_back((float)vtx.radius);
  }
  id = (size_t)firstVtx.treeId;
  const PointT bottom = landmarkVtxs[1].coords;
  model.root = Vector3(bottom.x, bottom.y, bottom.z);
  // too short: no model (cylinder.cpp:109-113)
  if (!needsLineFit(landmarkVtxs)) {
    model.radius = -1;
    return;
  }
  if (cfit) {
    // axis, root and radius of the cylinder fit; the radius floor of the
    // medians' model.  The root is the foot on the fitted axis of the bottom
    // vertex, which the medians' model takes as its root (groundBasedRoot
    // then keeps it: the axis points up, away from the ground)
    sloam_geom::CylFitRecord f;
    f.ok = cfit->ok;
    for (int i = 0; i < 3; ++i) {
      f.root[i] = cfit->root[i];
      f.ray[i] = cfit->ray[i];
    }
    f.radius = cfit->radius;
    sloam_geom::CylModel m{model.root, model.ray, model.radius};
    sloam_geom::cyl_model_from_fit(bottom, f, defaultTreeRadius, m);
    model.root = m.root;
    model.ray = m.ray;
    model.radius = m.radius;
    if (model.radius == -1) return;
    model.vertices = landmarkVtxs;
    model.lambda = 1.0;
    features.resize((size_t)featuresPerTree);
    return;
  }
  if (!fit->ok || fit->inliers == 0) {
    model.radius = -1;
    return;
  }
  model.ray = Vector3(fit->coeff[3], fit->coeff[4], fit->coeff[5]);
  if (!validRadii.empty()) {
    std::sort(validRadii.begin(), validRadii.end());
    const int d = (int)(std::find_if(validRadii.begin(), validRadii.end(),
                                     [](float x) { return x > 0; }) -
                        validRadii.begin());
    const int middle =
        std::min((int)(validRadii.size() - 1), (d + 1) + (int)((validRadii.size() - d) / 2));
    model.radius = validRadii[middle];
    if (model.radius == 0)
      model.radius = -1;
    else if (model.radius < defaultTreeRadius)
      model.radius = defaultTreeRadius;
  } else {
    model.radius = -1;
  }
  model.vertices = landmarkVtxs;
  model.lambda = 1.0;
  features.resize((size_t)featuresPerTree);  // zero-pads (cylinder.cpp:172)
}

Scalar Cylinder::distance(const CylinderParameters& tgt) const {
  return sloam_geom::cyl_model_distance(model.root, model.ray, tgt.root, tgt.ray);
}

Scalar Cylinder::distance(const PointT& point) const {
  const Vector3 p(point.x, point.y, point.z);
  const Vector3 proj = model.root + ((p - model.root).dot(model.ray) / model.ray.dot(model.ray)) * model.ray;
  return (p - proj).norm() - model.radius;
}

CylinderParameters Cylinder::projectModel(const SE3& tf) const {
  CylinderParameters m;
  m.radius = model.radius;
  m.lambda = model.lambda;
  sloam_geom::cyl_project(tf, model.root, model.ray, m.root, m.ray);
  return m;
}

void Cylinder::project(const SE3& tf) {
  const CylinderParameters m = projectModel(tf);
  model.root = m.root;
  model.ray = m.ray;
}
