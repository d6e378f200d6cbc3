// Tree trunk landmark (sloam/src/objects/cylinder.cpp), axis fitted on the GPU.
#include "cylinder.h"

#include <algorithm>
#include <cmath>
#include <limits>

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

namespace {

float squared_distance(const PointT& a, const PointT& b) {
  const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return dx * dx + dy * dy + dz * dz;
}

}  // namespace

bool Cylinder::withinRange(const std::vector<TreeVertex>& vertices, const Plane& gplane,
                           const FeatureModelParams& fmParams) {
  PointT centroid;
  centroid.x = (float)gplane.model.centroid(0);
  centroid.y = (float)gplane.model.centroid(1);
  centroid.z = (float)gplane.model.centroid(2);
  return euclideanDist2D(vertices[1].coords, centroid) < fmParams.maxLidarDist;
}

bool Cylinder::needsLineFit(const std::vector<TreeVertex>& v) {
  if (v.size() < 3) return false;
  return !(squared_distance(v[1].coords, v[v.size() - 2].coords) < 1.5f);
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

void Cylinder::build(const std::vector<TreeVertex>& vertices, const Plane& gplane,
                     const FeatureModelParams& fmParams, const LineFit* fit,
                     const CylinderFit* cfit) {
  isValid = false;
  if (vertices.size() < 3) return;  // the reference indexes vertices [1] and [2]
  if (withinRange(vertices, gplane, fmParams)) {
    computeModel(vertices, fmParams.defaultTreeRadius, fmParams.featuresPerTree, fit, cfit);
    const bool validZ = groundBasedRoot(gplane);
    const bool validNorm = model.root.norm() > 0.01;
    const bool validRadius = model.radius > 0.0;
    const bool validTree = filter(fmParams.maxTreeRadius, fmParams.maxAxisTheta, gplane);
    isValid = validZ && validRadius && validTree && validNorm;
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
  const Vector4& g = gplane.model.plane;
  const float dist = (float)(std::fabs(g[0] * model.root(0) + g[1] * model.root(1) +
                                       g[2] * model.root(2) + g[3]) /
                             g.head3().norm());
  if (dist < 2.0) {
    model.root = rayPlaneIntersection(gplane.model.centroid, g.head3(), model.root, model.ray);
    return true;
  }
  return false;
}

bool Cylinder::filter(const Scalar& maxTreeRadius, const Scalar& maxAxisTheta, const Plane& gplane) {
  if (model.radius == -1) return false;
  const Vector3 up = gplane.model.plane.head3();
  const Scalar theta = (180 / PIDEF) * std::acos(model.ray.dot(up) / (model.ray.norm() * up.norm()));
  return (model.radius < maxTreeRadius) && ((theta <= maxAxisTheta) || (theta >= 180 - maxAxisTheta));
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
    // medians' model
    if (!cfit->ok || !(cfit->radius > 0.0)) {
      model.radius = -1;
      return;
    }
    const double sgn = cfit->ray[2] < 0.0 ? -1.0 : 1.0;  // axis pointing up
    model.ray = Vector3(sgn * cfit->ray[0], sgn * cfit->ray[1], sgn * cfit->ray[2]);
    // root: the foot on the fitted axis of the bottom vertex, which the
    // medians' model takes as its root (groundBasedRoot then keeps it: the
    // axis points up, away from the ground)
    const Vector3 c(cfit->root[0], cfit->root[1], cfit->root[2]);
    model.root = c + (model.root - c).dot(model.ray) * model.ray;
    model.radius = cfit->radius < defaultTreeRadius ? defaultTreeRadius : cfit->radius;
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
  const Scalar heights[3] = {0.0, 3.0, 6.0};
  Scalar d = 0.0;
  for (const Scalar h : heights) {
    const Scalar st = (h - model.root[2]) / model.ray[2];
    const Vector3 mp = model.root + st * model.ray;
    const Scalar tt = (h - tgt.root[2]) / tgt.ray[2];
    const Vector3 tp = tgt.root + tt * tgt.ray;
    d += (mp - tp).norm();
  }
  return d / 3.0;
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
  Vector3 other = model.root + model.ray;
  m.root = tf * model.root;
  other = tf * other;
  m.ray = other - m.root;
  return m;
}

void Cylinder::project(const SE3& tf) {
  const CylinderParameters m = projectModel(tf);
  model.root = m.root;
  model.ray = m.ray;
}
