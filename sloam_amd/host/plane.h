// Ground plane landmark: the interface of sloam/include/objects/plane.h:14-32.
// The fit runs on the GPU (sloam_ground_planes); computeModels fits all
// ground cells in one batch and builds Planes from the results.
#pragma once

#include "definitions.h"
#include "semanticObject.h"
#include "utils.h"

struct PlaneParameters {
  Vector4 plane;
  Vector3 centroid;
  double dist = 0.0;
};

class Plane : public SemanticObject<PlaneParameters> {
 public:
  // Fits the plane to `points` (plane.cpp:3-17): invalid when there are fewer
  // than numGroundFeatures points; features keep the first numGroundFeatures.
  explicit Plane(const VectorType& points, const FeatureModelParams& fmParams);
  // A plane already fitted on the GPU (batched path of computeModels).
  Plane(const PlaneParameters& fitted, VectorType features, bool valid);
  Scalar distance(const PlaneParameters& tgt) const override;
  Scalar distance(const PointT& point) const override;
  void project(const SE3& tf) override;
  // the model as project() leaves it, without touching this object
  PlaneParameters projectModel(const SE3& tf) const;
};
