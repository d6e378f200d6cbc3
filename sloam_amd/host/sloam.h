// Per-keyframe SLOAM estimator: the interface of sloam/include/core/sloam.h.
// computeModels fits every ground cell and every tree of a scan in batched GPU
// calls; the pose solves run on the GPU LM kernels.
#pragma once

#include <utility>
#include <vector>

#include "cylinder.h"
#include "definitions.h"
#include "plane.h"
#include "sloam_gpu.h"
#include "utils.h"

template <typename T>
struct ObjectMatch {
  ObjectMatch(PointT ft, T obj, Scalar d) : feature(ft.x, ft.y, ft.z), object(obj), dist(d) {}
  Vector3 feature;
  T object;
  double dist;
};

struct SloamInput {
  SloamInput() : groundCloud(new CloudT()) {}
  SE3 poseEstimate;
  Scalar distance = 0.0;
  CloudT::Ptr groundCloud;
  std::vector<Cylinder> mapModels;
  std::vector<std::vector<TreeVertex>> landmarks;
  // Optional: the ground cells already fitted on the device (one Plane per
  // radii x theta cell, r-major, isValid as fitted), e.g. by the keyframe
  // front end; computeModels then skips binning groundCloud.
  std::vector<Plane> groundCells;
  // Optional: one axis line fit per landmark, fitted on the device to the
  // landmark's vertex medians with Cylinder::lineFitCfg() (e.g. by the
  // keyframe front end); computeModels then takes the fit of every tree that
  // needs one from here instead of fitting the lines itself.
  std::vector<LineFit> treeLineFits;
  // Optional: one cylinder fit per landmark of the landmark's points (e.g. by
  // the keyframe runner's batched plain or learned fit); when given, every
  // tree takes its axis, root and radius from here (Cylinder's CylinderFit
  // constructor) and treeLineFits is not used.
  std::vector<CylinderFit> treeCylinderFits;
};

struct SloamOutput {
  SloamOutput() {}
  std::vector<int> matches;
  std::vector<Cylinder> tm;
  SE3 T_Map_Curr;
  SE3 T_Delta;
};

// map trees handed to RunSloam by reference (e.g. MapManager::getSubmap's
// view): pointers valid for the duration of the call
using MapView = std::vector<const Cylinder*>;

// ground cells [radii bin][theta bin] (the reference's boost::multi_array)
using GroundGrid = std::vector<std::vector<VectorType>>;

// wall time (ms) of the phases of the last RunSloam call: model fits,
// feature association, pose solve, model projection + map matching
struct SloamPhaseTimes {
  double models = 0, match = 0, pose = 0, finish = 0;
};

// what the last RunSloam call decided: whether its XY-yaw (tree) and
// Z-roll-pitch (ground) steps had enough matched features to run, and how
// many tree and ground feature matches it had (zero before association);
// the LM iterations of its tree (or joint) and ground solves, -1 for a solve
// that did not run
struct SloamStepInfo {
  bool treeCheck = false, groundCheck = false;
  int treeMatches = 0, planeMatches = 0;
  int treeIters = -1, groundIters = -1;
};

namespace sloam {

class sloam {
 public:
  explicit sloam();

  const FeatureModelParams& fmParams() const { return fmParams_; }
  void setFmParams(const FeatureModelParams& fmParams) {
    fmParams_ = fmParams;
    minPlanes_ = (fmParams_.groundRadiiBins * fmParams_.groundThetaBins) * 0.1;
  }
  bool RunSloam(SloamInput& in, SloamOutput& out);
  // The same step with the map trees passed apart from `in` (in.mapModels
  // is not read): a caller that keeps its submap need not copy it, with
  // every tree's vertices and their points, into the input each keyframe.
  bool RunSloam(SloamInput& in, const std::vector<Cylinder>& mapModels, SloamOutput& out);
  bool RunSloam(SloamInput& in, const MapView& mapModels, SloamOutput& out);

  // Pose optimization
  void OptimizeXYYaw(const SE3& poseEstimate, const bool optimize,
                     const std::vector<ObjectMatch<Cylinder>>& allTMatch, double* out);
  void OptimizeZRollPitch(const SE3& poseEstimate, const bool optimize,
                          const std::vector<ObjectMatch<Plane>>& allGMatch, double* out);
  bool TwoStepOptimizePose(const SE3& poseEstimate, const bool optimTrees, const bool optimGround,
                           const std::vector<ObjectMatch<Cylinder>>& allTMatch,
                           const std::vector<ObjectMatch<Plane>>& allGMatch, SE3& tf);
  bool OptimizePose(const SE3& poseEstimate, const std::vector<ObjectMatch<Cylinder>>& allTMatch,
                    const std::vector<ObjectMatch<Plane>>& allGMatch, SE3& tf);
  // Model estimation
  void projectModels(const SE3& tf, std::vector<Cylinder>& landmarks, std::vector<Plane>& planes);
  void computeModels(SloamInput& in, std::vector<Cylinder>& landmarks, std::vector<Plane>& planes);
  void binGroundPoints(const SE3 pose, const VectorType& points, GroundGrid& scgf);
  // Data association
  template <typename T>
  void matchModels(const std::vector<T>& currObjects, const std::vector<T>& mapObjects,
                   std::vector<int>& matchIndices);
  template <typename T>
  std::vector<ObjectMatch<T>> matchFeatures(const SE3 tf, const std::vector<T>& currObjects,
                                            const std::vector<T>& mapObjects,
                                            const Scalar distThresh);
  template <typename T>
  void addFeatureMatches(const VectorType& features, const T& object, const double dist,
                         std::vector<ObjectMatch<T>>& matches);

  std::vector<Plane> getPrevGroundModel();
  // Replaces the previous scan's ground models, which the next RunSloam
  // matches the scan's planes against (the reference has the getter only):
  // a second estimator then starts a keyframe from another's state, so that
  // two tree models can be compared one keyframe at a time.
  void setPrevGroundModel(std::vector<Plane> planes) { prevGPlanes_ = std::move(planes); }
  CloudT getPrevGroundFeatures();
  const SloamPhaseTimes& lastPhaseTimes() const { return phase_; }
  const SloamStepInfo& lastStep() const { return step_; }
  // The trees the last RunSloam call modelled but rejected: tree id and
  // Cylinder::RejectReason bits.  No counterpart in the reference.
  const std::vector<std::pair<int, int>>& lastRejectedTrees() const { return rejected_; }
  // Arithmetic of the residuals, Jacobians and normal equations of the pose
  // solves: 1 (default) FP64, as Ceres evaluates them; 0 FP32 with
  // compensated FP32 sums (sloam_pose_cfg.precision).  No counterpart in the
  // reference.
  void setPosePrecision(int precision);
  int posePrecision() const { return posePrecision_; }

 private:
  // ground binning / fit settings of fmParams_, binned around `pose`
  sloam_ground_cfg groundCfg(const SE3& pose) const;

  SE3 T_Map_Anchor_;
  std::vector<Plane> prevGPlanes_;
  int numMapTrees_ = 0;
  double totalDistance_ = 0.0;
  FeatureModelParams fmParams_;
  bool firstScan_ = true;
  double minPlanes_ = 0.0;
  SloamPhaseTimes phase_;
  SloamStepInfo step_;
  std::vector<std::pair<int, int>> rejected_;
  int posePrecision_ = 1;
};

}  // namespace sloam
