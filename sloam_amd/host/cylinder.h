// Tree trunk landmark: the interface of sloam/include/objects/cylinder.h:14-40.
// The axis is fitted on the GPU (sloam_fit_lines, PCL sample-consensus
// semantics); computeModels fits all trees of a scan in one batch.
#pragma once

#include <vector>

#include "definitions.h"
#include "plane.h"
#include "semanticObject.h"
#include "sloam_gpu.h"
#include "utils.h"

struct CylinderParameters {
  Vector3 root;  // a point on the axis
  Vector3 ray;   // axis direction
  std::vector<double> radii;
  std::vector<TreeVertex> vertices;
  double radius = 0.0;
  double lambda = 0.0;
};

// Result of the GPU line fit of one tree (coefficients point + direction).
struct LineFit {
  bool ok = false;
  float coeff[6] = {0, 0, 0, 0, 0, 0};
  int inliers = 0;
};

// Result of the batched GPU cylinder fit of one tree's points
// (sloam_cyl_plan_exec / sloam_cyl_learned_exec model record): a point on the
// axis, the unit axis, the radius and the inlier count.
struct CylinderFit {
  bool ok = false;
  double root[3] = {0, 0, 0};
  double ray[3] = {0, 0, 0};
  double radius = -1.0;
  int inliers = 0;
};

class Cylinder : public SemanticObject<CylinderParameters> {
 public:
  explicit Cylinder(const std::vector<TreeVertex> vertices, const Plane& gplane,
                    const FeatureModelParams& fmParams);
  // Construction from a batched GPU line fit (`fit` of vertices' medians);
  // `fit` is ignored when the tree is out of range or too short.
  Cylinder(const std::vector<TreeVertex>& vertices, const Plane& gplane,
           const FeatureModelParams& fmParams, const LineFit& fit);
  // Construction from a cylinder fit of the tree's points (`fit`): axis,
  // root and radius from the fit instead of the medians' line, the radius
  // and validity rules of the line-fit model otherwise.
  Cylinder(const std::vector<TreeVertex>& vertices, const Plane& gplane,
           const FeatureModelParams& fmParams, const CylinderFit& fit);
  // A valid landmark with a model computed elsewhere (the GPU keyframe
  // estimator): root, ray, radius, id and features as given, no vertices.
  // No counterpart in the reference.
  Cylinder(const CylinderParameters& model, size_t id, VectorType features);
  Scalar distance(const CylinderParameters& model) const override;
  Scalar distance(const PointT& point) const override;
  void project(const SE3& tf) override;
  // the model's root and ray as project() leaves them, without touching this
  // object (radius and lambda copied; radii and vertices left empty)
  CylinderParameters projectModel(const SE3& tf) const;
  PointT asPoint(const Scalar height);

  // cylinder.cpp:3-13: vertex [1] within maxLidarDist (2-D) of the plane centroid
  static bool withinRange(const std::vector<TreeVertex>& vertices, const Plane& gplane,
                          const FeatureModelParams& fmParams);
  // cylinder.cpp:101-113: the line fit is needed (tree tall enough, has vertices)
  static bool needsLineFit(const std::vector<TreeVertex>& vertices);
  // the sample-consensus line settings of computeModel (cylinder.cpp:116-124)
  static sloam_line_fit_cfg lineFitCfg();

  // Why construction left the tree invalid, as bits (0 when valid): vertex [1]
  // out of range, no model (line or cylinder fit failed, tree too short), root
  // more than 2 m from the ground plane, root at the origin, radius not
  // positive, radius at or over the cap, axis tilted past maxAxisTheta.  No
  // counterpart in the reference, which keeps only isValid.
  enum RejectReason : int {
    kRejectOutOfRange = 1,
    kRejectNoModel = 2,
    kRejectRootOffGround = 4,
    kRejectRootAtOrigin = 8,
    kRejectRadiusNotPositive = 16,
    kRejectRadiusCap = 32,
    kRejectAxisTilt = 64,
  };
  int rejectReason = 0;

 private:
  // exactly one of fit / cfit is given
  void build(const std::vector<TreeVertex>& vertices, const Plane& gplane,
             const FeatureModelParams& fmParams, const LineFit* fit, const CylinderFit* cfit);
  // the rules of filter (cylinder.cpp:57-73) the model fails
  // (sloam_geom::CylFilterFail bits; 0: it passes)
  int filterFails(const Scalar& maxTreeRadius, const Scalar& maxAxisTheta, const Plane& gplane);
  void computeModel(const std::vector<TreeVertex>& landmarkVtxs, const Scalar defaultTreeRadius,
                    const Scalar featuresPerTree, const LineFit* fit, const CylinderFit* cfit);
  bool groundBasedRoot(const Plane& gplane);
};
