// Per-keyframe SLOAM estimator (sloam/src/core/sloam.cpp) over the GPU
// library: every ground cell of a scan is fitted in one sloam_ground_planes
// call, every tree axis in one sloam_fit_lines call, and the pose solves
// (joint, or the two steps together) in one sloam_pose_plan batch.
#include "sloam.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <utility>

#include "geometry.h"
#include "gpu_bridge.h"

namespace sloam {

namespace {

// Ceres settings of the reference's solves (sloam.cpp:88-97, 219-226):
// LM, HuberLoss(0.1), 50 iterations, Ceres' default tolerances; precision as
// sloam_pose_cfg.precision.
sloam_pose_cfg reference_pose_cfg(int parametrization, int precision) {
  sloam_pose_cfg c{};
  c.max_iters = 50;
  c.stop_on_convergence = 1;
  c.huber_delta = 0.1;
  c.function_tolerance = 1e-6;
  c.gradient_tolerance = 1e-10;
  c.parameter_tolerance = 1e-8;
  c.initial_radius = 1e4;
  c.min_relative_decrease = 1e-3;
  c.parametrization = parametrization;
  c.precision = precision;
  return c;
}

constexpr int32_t kXYYawFree = (1 << 0) | (1 << 1) | (1 << 5);       // x, y, w_z
constexpr int32_t kZRollPitchFree = (1 << 2) | (1 << 3) | (1 << 4);  // z, w_x, w_y

void push_point(sloam_gpu::PoseProblem& pr, const Vector3& f, int32_t id) {
  pr.points.push_back((float)f[0]);
  pr.points.push_back((float)f[1]);
  pr.points.push_back((float)f[2]);
  float bits;
  std::memcpy(&bits, &id, 4);
  pr.points.push_back(bits);
}

bool same_cylinder(const double* a, const CylinderParameters& m) {
  return a[0] == m.root[0] && a[1] == m.root[1] && a[2] == m.root[2] && a[3] == m.ray[0] &&
         a[4] == m.ray[1] && a[5] == m.ray[2] && a[6] == m.radius;
}

bool same_plane(const double* a, const PlaneParameters& m) {
  return a[0] == m.plane[0] && a[1] == m.plane[1] && a[2] == m.plane[2] && a[3] == m.plane[3];
}

// A feature matched to a map model, pointing at the map object instead of
// copying it (ObjectMatch<T> holds a full copy of T, vertices and points
// included): the form RunSloam uses between association and the pose solve.
template <typename M>
struct ModelMatch {
  Vector3 feature;
  const M* model;
  double dist;
};

const CylinderParameters& model_of(const ObjectMatch<Cylinder>& m) { return m.object.model; }
const PlaneParameters& model_of(const ObjectMatch<Plane>& m) { return m.object.model; }
template <typename M>
const M& model_of(const ModelMatch<M>& m) {
  return *m.model;
}

// matchFeatures (sloam.cpp:443-471) without the copies: the same nearest map
// model (strict <, first wins) for each current object's projected model, and
// one match per feature of the current object.
// the map objects as a list of references: a vector of objects or a view
template <typename T>
const T& map_at(const std::vector<T>& v, size_t i) {
  return v[i];
}
template <typename T>
const T& map_at(const std::vector<const T*>& v, size_t i) {
  return *v[i];
}

template <typename T, typename M, typename MapList>
std::vector<ModelMatch<M>> match_features(const SE3& tf, const std::vector<T>& currObjects,
                                          const MapList& mapObjects, const Scalar distThresh) {
  std::vector<ModelMatch<M>> matches;
  if (mapObjects.empty()) return matches;
  for (const auto& co : currObjects) {
    const M pm = co.projectModel(tf);
    Scalar bestDist = distThresh + 100;
    size_t best = 0;
    for (size_t i = 0; i < mapObjects.size(); ++i) {
      const Scalar d = map_at(mapObjects, i).distance(pm);
      if (d < bestDist) {
        bestDist = d;
        best = i;
      }
    }
    if (bestDist < distThresh)
      for (const auto& f : co.features)
        matches.push_back(
            ModelMatch<M>{Vector3(f.x, f.y, f.z), &map_at(mapObjects, best).model, bestDist});
  }
  return matches;
}

// matchModels (sloam.cpp:299-328) over either map list form
template <typename T, typename MapList>
void match_models(const std::vector<T>& currObjects, const MapList& mapObjects,
                  const FeatureModelParams& fm, std::vector<int>& matchIndices) {
  if (mapObjects.empty()) return;
  size_t counter = 0;
  for (const auto& co : currObjects) {
    Scalar bestDist = fm.treeMatchThresh + 100;
    size_t bestKey = 0;
    for (size_t key = 0; key < mapObjects.size(); ++key) {
      const Scalar d = map_at(mapObjects, key).distance(co.model);
      if (d < bestDist) {
        bestDist = d;
        bestKey = key;
      }
    }
    if (bestDist < fm.AddNewTreeThreshDist) matchIndices[counter] = (int)bestKey;
    counter++;
  }
}

// Consecutive matches against the same landmark share one table entry (the
// residuals are the same as with one entry per match).
template <typename V>
void add_tree_matches(const V& matches, sloam_gpu::PoseProblem& pr) {
  for (const auto& m : matches) {
    const CylinderParameters& o = model_of(m);
    const size_t n = pr.cyl.size() / 8;
    if (n == 0 || !same_cylinder(&pr.cyl[8 * (n - 1)], o)) {
      const double row[8] = {o.root[0], o.root[1], o.root[2], o.ray[0],
                             o.ray[1],  o.ray[2],  o.radius,  0.0};
      pr.cyl.insert(pr.cyl.end(), row, row + 8);
    }
    push_point(pr, m.feature, (int32_t)(pr.cyl.size() / 8 - 1));
  }
}

template <typename V>
void add_ground_matches(const V& matches, sloam_gpu::PoseProblem& pr) {
  for (const auto& m : matches) {
    const PlaneParameters& o = model_of(m);
    const size_t n = pr.planes.size() / 4;
    if (n == 0 || !same_plane(&pr.planes[4 * (n - 1)], o))
      for (int i = 0; i < 4; ++i) pr.planes.push_back(o.plane[i]);
    push_point(pr, m.feature, -(int32_t)(pr.planes.size() / 4));
  }
}

void set_initial_pose(const SE3& pose, sloam_gpu::PoseProblem& pr) {
  const Vector3& t = pose.translation();
  const Quat& q = pose.unit_quaternion();
  const double x[7] = {t[0], t[1], t[2], q.x, q.y, q.z, q.w};
  std::memcpy(pr.pose, x, sizeof(x));
}

// [t, angle-axis] of a pose, as the reference forms its 6-vector
// (sloam.cpp:59-64)
void pose_params(const SE3& pose, double params[6]) {
  const Vector3& t = pose.translation();
  const Quat& q = pose.unit_quaternion();
  const double quat[4] = {q.w, q.x, q.y, q.z};
  params[0] = t[0];
  params[1] = t[1];
  params[2] = t[2];
  quaternionToAngleAxis(quat, params + 3);
}

// Runs the XY-yaw (trees) and Z-roll-pitch (ground) solves of a pose as one
// batch; a step that is not run or does not converge keeps the initial
// values (sloam.cpp:99-110, 151-161).
template <typename TV, typename GV>
void two_step(const SE3& poseEstimate, bool optimTrees, bool optimGround, const TV& allTMatch,
              const GV& allGMatch, double treeOut[3], double groundOut[3], int precision,
              int* iters = nullptr) {
  double p0[6];
  pose_params(poseEstimate, p0);
  treeOut[0] = p0[0];
  treeOut[1] = p0[1];
  treeOut[2] = p0[5];
  groundOut[0] = p0[2];
  groundOut[1] = p0[3];
  groundOut[2] = p0[4];
  std::vector<sloam_gpu::PoseProblem> probs;
  int tree_k = -1, ground_k = -1;
  // a solve without residuals has nothing to move (Ceres converges at once)
  if (optimTrees && !allTMatch.empty()) {
    sloam_gpu::PoseProblem pr;
    set_initial_pose(poseEstimate, pr);
    pr.free_mask = kXYYawFree;
    add_tree_matches(allTMatch, pr);
    tree_k = (int)probs.size();
    probs.push_back(std::move(pr));
  }
  if (optimGround && !allGMatch.empty()) {
    sloam_gpu::PoseProblem pr;
    set_initial_pose(poseEstimate, pr);
    pr.free_mask = kZRollPitchFree;
    add_ground_matches(allGMatch, pr);
    ground_k = (int)probs.size();
    probs.push_back(std::move(pr));
  }
  if (iters) iters[0] = iters[1] = -1;
  if (probs.empty()) return;
  const auto res = sloam_gpu::solvePoses(probs, reference_pose_cfg(1, precision));
  if (iters) {
    if (tree_k >= 0) iters[0] = res[tree_k].iterations;
    if (ground_k >= 0) iters[1] = res[ground_k].iterations;
  }
  if (tree_k >= 0 && res[tree_k].termination == 0) {
    treeOut[0] = res[tree_k].params[0];
    treeOut[1] = res[tree_k].params[1];
    treeOut[2] = res[tree_k].params[5];
  }
  if (ground_k >= 0 && res[ground_k].termination == 0) {
    groundOut[0] = res[ground_k].params[2];
    groundOut[1] = res[ground_k].params[3];
    groundOut[2] = res[ground_k].params[4];
  }
}

// TwoStepOptimizePose (sloam.cpp:33-53) over either match form
template <typename TV, typename GV>
bool two_step_pose(const SE3& poseEstimate, bool optimTrees, bool optimGround, const TV& allTMatch,
                   const GV& allGMatch, SE3& tf, int precision, int* iters = nullptr) {
  double treeOut[3], groundOut[3];
  two_step(poseEstimate, optimTrees, optimGround, allTMatch, allGMatch, treeOut, groundOut,
           precision, iters);
  // roll, pitch from the ground step, yaw from the tree step (sloam.cpp:41-51)
  const double angleAxis[3] = {groundOut[1], groundOut[2], treeOut[2]};
  double q[4];
  angleAxisToQuaternion(angleAxis, q);
  tf.setQuaternion(Quat(q[0], q[1], q[2], q[3]));
  tf.translation()[0] = treeOut[0];
  tf.translation()[1] = treeOut[1];
  tf.translation()[2] = groundOut[0];
  return true;
}

// OptimizePose (sloam.cpp:174-255) over either match form
template <typename TV, typename GV>
bool joint_pose(const SE3& poseEstimate, const TV& allTMatch, const GV& allGMatch, SE3& tf,
                int precision, int* iters = nullptr) {
  std::vector<sloam_gpu::PoseProblem> probs(1);
  set_initial_pose(poseEstimate, probs[0]);
  add_tree_matches(allTMatch, probs[0]);
  add_ground_matches(allGMatch, probs[0]);
  bool success;
  double x[7];
  if (probs[0].points.empty()) {
    // no residuals: the solve converges at the initial pose
    success = true;
    std::memcpy(x, probs[0].pose, sizeof(x));
  } else {
    const auto res = sloam_gpu::solvePoses(probs, reference_pose_cfg(0, precision));
    success = res[0].termination == 0;
    std::memcpy(x, res[0].pose, sizeof(x));
    if (iters) iters[0] = res[0].iterations;
  }
  // the output is written only on convergence (sloam.cpp:238-244)
  if (success) {
    tf.setQuaternion(Quat(x[6], x[3], x[4], x[5]));
    tf.translation() = Vector3(x[0], x[1], x[2]);
  }
  return success;
}

}  // namespace

sloam::sloam() {
  T_Map_Anchor_ = SE3();
  firstScan_ = true;
  minPlanes_ = (fmParams_.groundRadiiBins * fmParams_.groundThetaBins) * 0.1;
}

std::vector<Plane> sloam::getPrevGroundModel() { return prevGPlanes_; }

void sloam::setPosePrecision(int precision) {
  if (precision != 0 && precision != 1)
    throw std::invalid_argument("sloam: pose precision is 0 (FP32) or 1 (FP64)");
  posePrecision_ = precision;
}

CloudT sloam::getPrevGroundFeatures() {
  CloudT pc;
  for (const auto& ground : prevGPlanes_)
    for (const auto& pt : ground.features) pc.points.push_back(pt);
  pc.height = 1;
  pc.width = (uint32_t)pc.points.size();
  return pc;
}

bool sloam::TwoStepOptimizePose(const SE3& poseEstimate, const bool optimTrees,
                                const bool optimGround,
                                const std::vector<ObjectMatch<Cylinder>>& allTMatch,
                                const std::vector<ObjectMatch<Plane>>& allGMatch, SE3& tf) {
  return two_step_pose(poseEstimate, optimTrees, optimGround, allTMatch, allGMatch, tf,
                       posePrecision_);
}

void sloam::OptimizeXYYaw(const SE3& poseEstimate, const bool optimize,
                          const std::vector<ObjectMatch<Cylinder>>& allTMatch, double* out) {
  double ground[3];
  two_step(poseEstimate, optimize, false, allTMatch, std::vector<ObjectMatch<Plane>>(), out,
           ground, posePrecision_);
}

void sloam::OptimizeZRollPitch(const SE3& poseEstimate, const bool optimize,
                               const std::vector<ObjectMatch<Plane>>& allGMatch, double* out) {
  double tree[3];
  two_step(poseEstimate, false, optimize, std::vector<ObjectMatch<Cylinder>>(), allGMatch, tree,
           out, posePrecision_);
}

bool sloam::OptimizePose(const SE3& poseEstimate,
                         const std::vector<ObjectMatch<Cylinder>>& allTMatch,
                         const std::vector<ObjectMatch<Plane>>& allGMatch, SE3& tf) {
  return joint_pose(poseEstimate, allTMatch, allGMatch, tf, posePrecision_);
}

template <typename T>
std::vector<ObjectMatch<T>> sloam::matchFeatures(const SE3 tf, const std::vector<T>& currObjects,
                                                 const std::vector<T>& mapObjects,
                                                 const Scalar distThresh) {
  std::vector<ObjectMatch<T>> matches;
  if (mapObjects.empty()) return matches;  // the reference reads mapObjects[0]
  for (const auto& co : currObjects) {
    // the model of the object projected by tf (project() without the copy)
    const auto pm = co.projectModel(tf);
    // closest map model; ties keep the first (strict <)
    Scalar bestDist = distThresh + 100;
    size_t best = 0;
    for (size_t i = 0; i < mapObjects.size(); ++i) {
      const Scalar d = mapObjects[i].distance(pm);
      if (d < bestDist) {
        bestDist = d;
        best = i;
      }
    }
    if (bestDist < distThresh) addFeatureMatches(co.features, mapObjects[best], bestDist, matches);
  }
  return matches;
}

template <typename T>
void sloam::addFeatureMatches(const VectorType& features, const T& object, const double dist,
                              std::vector<ObjectMatch<T>>& matches) {
  for (const auto& feature : features) matches.push_back(ObjectMatch<T>(feature, object, dist));
}

template <typename T>
void sloam::matchModels(const std::vector<T>& currObjects, const std::vector<T>& mapObjects,
                        std::vector<int>& matchIndices) {
  match_models(currObjects, mapObjects, fmParams_, matchIndices);
}

template std::vector<ObjectMatch<Cylinder>> sloam::matchFeatures<Cylinder>(
    const SE3, const std::vector<Cylinder>&, const std::vector<Cylinder>&, const Scalar);
template std::vector<ObjectMatch<Plane>> sloam::matchFeatures<Plane>(
    const SE3, const std::vector<Plane>&, const std::vector<Plane>&, const Scalar);
template void sloam::matchModels<Cylinder>(const std::vector<Cylinder>&,
                                           const std::vector<Cylinder>&, std::vector<int>&);

sloam_ground_cfg sloam::groundCfg(const SE3& pose) const {
  sloam_ground_cfg cfg{};
  cfg.radii_bins = fmParams_.groundRadiiBins;
  cfg.theta_bins = fmParams_.groundThetaBins;
  cfg.min_dist = fmParams_.minGroundLidarDist;
  cfg.max_dist = fmParams_.maxGroundLidarDist;
  cfg.retain_thresh = fmParams_.groundRetainThresh;
  cfg.num_features = fmParams_.numGroundFeatures;
  // the binning origin is a PointT (sloam.cpp:333-336)
  cfg.origin[0] = (float)pose.translation()[0];
  cfg.origin[1] = (float)pose.translation()[1];
  cfg.origin[2] = (float)pose.translation()[2];
  return cfg;
}

void sloam::binGroundPoints(const SE3 pose, const VectorType& points, GroundGrid& scgf) {
  const auto cells = sloam_gpu::groundPlanes(points, groundCfg(pose));
  scgf.assign(fmParams_.groundRadiiBins, std::vector<VectorType>(fmParams_.groundThetaBins));
  for (int r = 0; r < fmParams_.groundRadiiBins; ++r)
    for (int t = 0; t < fmParams_.groundThetaBins; ++t)
      for (const int32_t i : cells[(size_t)r * fmParams_.groundThetaBins + t].kept_index)
        scgf[r][t].push_back(points[i]);
}

void sloam::computeModels(SloamInput& in, std::vector<Cylinder>& landmarks,
                          std::vector<Plane>& planes) {
  rejected_.clear();
  // ground: all cells of the scan in one call (binned at the origin,
  // sloam.cpp:390-393), unless the caller brings them
  std::vector<Plane> fittedCells;
  if (in.groundCells.empty()) {
    for (const auto& cell : sloam_gpu::groundPlanes(in.groundCloud->points, groundCfg(SE3()))) {
      PlaneParameters fitted;
      for (int i = 0; i < 4; ++i) fitted.plane(i) = cell.plane[i];
      for (int i = 0; i < 3; ++i) fitted.centroid(i) = cell.centroid[i];
      fittedCells.emplace_back(fitted, cell.features, cell.valid);
    }
  }
  const std::vector<Plane>& cells = in.groundCells.empty() ? fittedCells : in.groundCells;
  const Matrix4 tfm = in.poseEstimate.inverse().matrix().transpose();
  for (const Plane& ground : cells) {
    if (!ground.isValid) continue;
    // roughly horizontal in the map frame and under the robot
    // (sloam.cpp:401-408)
    if (sloam_geom::plane_accepted(in.poseEstimate, tfm, ground.model.plane,
                                   ground.model.centroid))
      planes.push_back(ground);
  }
  if (planes.empty()) return;

  // trees: nearest plane to vertex [1] of each, then one batched line fit
  // for the trees in range and tall enough (sloam.cpp:416-435)
  // every set is fitted on its own, so a fit made earlier on the device for
  // the same landmark is the same fit
  const bool have_cyl = !in.landmarks.empty() && in.treeCylinderFits.size() == in.landmarks.size();
  const bool have_fits = have_cyl || in.treeLineFits.size() == in.landmarks.size();
  std::vector<size_t> best(in.landmarks.size(), 0);
  std::vector<VectorType> sets;
  std::vector<int> set_of(in.landmarks.size(), -1);
  for (size_t k = 0; k < in.landmarks.size(); ++k) {
    const auto& t = in.landmarks[k];
    if (t.size() < 3) continue;  // the reference indexes vertices [1] and [2]
    Scalar bestDist = 100000;
    for (size_t i = 0; i < planes.size(); ++i) {
      const Scalar d = planes[i].distance(t[1].coords);
      if (d < bestDist) {
        bestDist = d;
        best[k] = i;
      }
    }
    if (Cylinder::withinRange(t, planes[best[k]], fmParams_) && Cylinder::needsLineFit(t)) {
      VectorType medians;
      if (!have_fits)
        for (const auto& v : t) medians.push_back(v.coords);
      set_of[k] = (int)sets.size();
      sets.push_back(std::move(medians));
    }
  }
  std::vector<sloam_gpu::LineResult> fits;
  if (!have_fits) fits = sloam_gpu::fitLines(sets, Cylinder::lineFitCfg());
  for (size_t k = 0; k < in.landmarks.size(); ++k) {
    if (in.landmarks[k].size() < 3) continue;
    if (have_cyl) {
      // a tree outside the gates gets no fit (as it gets no line fit)
      const CylinderFit none;
      Cylinder c(in.landmarks[k], planes[best[k]], fmParams_,
                 set_of[k] >= 0 ? in.treeCylinderFits[k] : none);
      if (c.isValid)
        landmarks.push_back(std::move(c));
      else
        rejected_.emplace_back(in.landmarks[k][2].treeId, c.rejectReason);
      continue;
    }
    LineFit fit;
    if (set_of[k] >= 0) {
      if (have_fits) {
        fit = in.treeLineFits[k];
      } else {
        const auto& r = fits[set_of[k]];
        fit.ok = r.ok;
        fit.inliers = r.inliers;
        std::memcpy(fit.coeff, r.coeff, sizeof(fit.coeff));
      }
    }
    Cylinder c(in.landmarks[k], planes[best[k]], fmParams_, fit);
    if (c.isValid)
      landmarks.push_back(std::move(c));
    else
      rejected_.emplace_back(in.landmarks[k][2].treeId, c.rejectReason);
  }
}

void sloam::projectModels(const SE3& tf, std::vector<Cylinder>& landmarks,
                          std::vector<Plane>& planes) {
  for (auto& p : planes) p.project(tf);
  for (auto& l : landmarks) l.project(tf);
}

namespace {
using PhaseClock = std::chrono::steady_clock;
double ms_between(PhaseClock::time_point a, PhaseClock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}
}  // namespace

bool sloam::RunSloam(SloamInput& in, SloamOutput& out) { return RunSloam(in, in.mapModels, out); }

bool sloam::RunSloam(SloamInput& in, const std::vector<Cylinder>& mapModels, SloamOutput& out) {
  MapView view;
  view.reserve(mapModels.size());
  for (const Cylinder& c : mapModels) view.push_back(&c);
  return RunSloam(in, view, out);
}

bool sloam::RunSloam(SloamInput& in, const MapView& mapModels, SloamOutput& out) {
  phase_ = SloamPhaseTimes();
  step_ = SloamStepInfo();
  const auto p0 = PhaseClock::now();
  std::vector<Plane> planes;
  std::vector<Cylinder> landmarks;
  computeModels(in, landmarks, planes);
  const auto p1 = PhaseClock::now();
  phase_.models = ms_between(p0, p1);

  std::vector<int> matchIndices(landmarks.size(), -1);
  bool success = true;
  if (firstScan_) {
    // the first scan's pose is the odometry's; its models start the map
    projectModels(in.poseEstimate, landmarks, planes);
    out.T_Map_Curr = in.poseEstimate;
    out.matches = matchIndices;
    out.tm = landmarks;
    prevGPlanes_ = planes;
    firstScan_ = false;
    return success;
  }
  if (mapModels.empty()) return false;                   // empty map
  if (planes.empty() || landmarks.empty()) return false;  // no models in this scan

  // matchFeatures' associations, pointing at the map objects instead of
  // copying them (they stay alive until the pose solve is done)
  const auto treeMatches = match_features<Cylinder, CylinderParameters>(
      in.poseEstimate, landmarks, mapModels, fmParams_.treeMatchThresh);
  // the ground association threshold is fixed at 1.0 (sloam.cpp:482)
  const auto planeMatches =
      match_features<Plane, PlaneParameters>(in.poseEstimate, planes, prevGPlanes_, 1.0);
  const auto p2 = PhaseClock::now();
  phase_.match = ms_between(p1, p2);

  SE3 T_Delta = SE3();
  SE3 currPose = in.poseEstimate;
  const bool treeCheck = landmarks.size() > fmParams_.minTreeModels &&
                         treeMatches.size() > 5 * fmParams_.featuresPerTree;
  const bool groundCheck =
      planeMatches.size() > fmParams_.minGroundModels && planes.size() > minPlanes_;
  step_ = {treeCheck, groundCheck, (int)treeMatches.size(), (int)planeMatches.size()};
  int iters[2] = {-1, -1};
  if (fmParams_.twoStepOptim) {
    success = two_step_pose(in.poseEstimate, treeCheck, groundCheck, treeMatches, planeMatches,
                            currPose, posePrecision_, iters);
  } else if (treeCheck && groundCheck) {
    success = joint_pose(in.poseEstimate, treeMatches, planeMatches, T_Delta, posePrecision_,
                         iters);
    if (success) currPose = T_Delta;
  }
  step_.treeIters = iters[0];
  step_.groundIters = iters[1];
  const auto p3 = PhaseClock::now();
  phase_.pose = ms_between(p2, p3);

  projectModels(currPose, landmarks, planes);
  match_models(landmarks, mapModels, fmParams_, matchIndices);
  prevGPlanes_ = std::move(planes);
  out.matches = matchIndices;
  out.T_Map_Curr = currPose;
  out.T_Delta = T_Delta;
  out.tm = std::move(landmarks);
  phase_.finish = ms_between(p3, PhaseClock::now());
  return success;
}

}  // namespace sloam
