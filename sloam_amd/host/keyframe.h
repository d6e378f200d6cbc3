// One SLOAM keyframe on the GPU: the front end (projection + ground / tree
// partition, tree instances, ground cells; one HIP graph per scan after the
// first) and the estimator (sloam::sloam::RunSloam) fed from its results.
// The per-scan path of the reference's SLOAMNode::run
// (sloam/src/core/sloamNode.cpp:235-300) with the labels given instead of the
// segmentation network.
#pragma once

#include <chrono>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "definitions.h"
#include "sloam.h"
#include "sloam_gpu.h"

struct KeyframeTiming {
  double upload_ms = 0, frontend_ms = 0, download_ms = 0, estimator_ms = 0, total_ms = 0;
  double build_ms = 0;      // estimator input assembled from the front end's arrays
  double fit_ms = 0;        // tree cylinder fits (TreeModel::Plain / Learned)
  SloamPhaseTimes phases;   // RunSloam's phases
  int trees = 0, planes = 0;
  int fitted = 0;           // trees given a cylinder fit
  int fallback = 0;         // of those, refitted by the plain method (Learned)
};

// How each tree's model is made.
enum class TreeModel {
  // the reference's: a sample-consensus line through the vertex medians
  // (fitted in the front end) and the median vertex radius
  MedianLines = 0,
  // batched point-level cylinder RANSAC over each tree's points (1024
  // hypotheses, 5 cm threshold, LO refit): axis, root and radius
  Plain = 1,
  // the learned proposal (PointNet) verified by 16 hypotheses, with the plain
  // fit as the fallback for trees whose best verifying hypothesis keeps
  // fewer than half their points
  Learned = 2,
};

class KeyframeRunner {
 public:
  // height x width organised scans; fmParams for the estimator; use_graph
  // replays the front end as a HIP graph; pipelined gives the front end two
  // slots on streams of their own, so that submit() can run a scan's front
  // end while finish() works on the previous one
  KeyframeRunner(int height, int width, const FeatureModelParams& fmParams, bool use_graph = true,
                 float fov_up_deg = 22.5f, float fov_down_deg = -22.5f, bool pipelined = false);
  ~KeyframeRunner();
  KeyframeRunner(const KeyframeRunner&) = delete;
  KeyframeRunner& operator=(const KeyframeRunner&) = delete;

  // xyzi: height * width points of 4 floats (sensor frame, NaN = no return);
  // labels: 0 other, 1 ground, 2 tree.  mapModels: the map trees (world
  // frame).  Returns RunSloam's result.
  bool run(const float* xyzi, const uint8_t* labels, const SE3& poseEstimate,
           const std::vector<Cylinder>& mapModels, SloamOutput& out,
           KeyframeTiming* timing = nullptr);
  // the same with the map trees by reference (MapManager::getSubmap's view)
  bool run(const float* xyzi, const uint8_t* labels, const SE3& poseEstimate,
           const MapView& mapModels, SloamOutput& out, KeyframeTiming* timing = nullptr);

  // The same step in two halves, for a stream of scans: submit() enqueues a
  // scan's upload and front end and returns; finish() waits for the oldest
  // submitted scan's front end and runs the estimator on it (the front end
  // does not depend on the pose, the estimator does).  At most as many scans
  // as front-end slots (1, or 2 when pipelined) may be submitted and not
  // finished; run() needs none pending.  A finish()'s timing counts from its
  // scan's submit(): upload_ms is 0 and frontend_ms is the wait for the
  // front end.
  void submit(const float* xyzi, const uint8_t* labels);
  bool finish(const SE3& poseEstimate, const std::vector<Cylinder>& mapModels, SloamOutput& out,
              KeyframeTiming* timing = nullptr);
  bool finish(const SE3& poseEstimate, const MapView& mapModels, SloamOutput& out,
              KeyframeTiming* timing = nullptr);
  // releases the oldest submitted scan without running the estimator (a
  // keyframe the caller discards)
  void skip();
  int pending() const { return pending_; }
  int slots() const { return (int)slots_.size(); }

  sloam::sloam& estimator() { return sloam_; }

  // Runs the whole keyframe on the GPU, for the cylinder-fit tree models
  // (Plain, Learned): the scan's upload, the front end, the tree fits on
  // counted plans (sizes read on the device), the estimator
  // (sloam_kf_estimator instead of the host sloam::sloam: the same
  // decisions, pose and landmarks) and the results into pinned memory, with
  // no host step in between; with use_graph the sequence is captured once
  // and every later keyframe is one HIP graph launch.  The returned
  // landmarks carry their model, id and features (no vertices).  The device
  // keeps its own first-scan flag and previous planes: switching on starts
  // from a first scan, and estimator() is not used meanwhile.  run() only
  // (not submit / finish); at most 512 trees per scan.
  void setDeviceEstimator(bool on);
  bool deviceEstimator() const { return device_est_; }
  // what the last keyframe's estimator decided (either estimator)
  SloamStepInfo lastStep() const {
    return device_est_ ? dev_step_ : sloam_.lastStep();
  }

  // Called with each keyframe's estimator input right before RunSloam (its
  // landmarks, ground cells and tree fits as the front end made them, and
  // the map trees): for recording keyframes, e.g. to replay them elsewhere.
  // The input is valid only during the call.
  using InputRecorder =
      std::function<void(const SloamInput& in, const MapView& mapModels)>;
  void setInputRecorder(InputRecorder r) { recorder_ = std::move(r); }

  // Pinned host buffers of one scan (height * width x 4 floats, height *
  // width labels) of the front-end slot the next run() / submit() uses, owned
  // by the runner: a scan written here (e.g. by the sensor driver) and passed
  // on is uploaded by DMA straight from them, at about twice the rate of
  // pageable memory.
  float* pinnedXyzi();
  uint8_t* pinnedLabels();

  // Tree model of the following keyframes; Learned needs the proposal
  // network's parameters (float32 .npy of SLOAM_PN_PARAMS values).
  void setTreeModel(TreeModel mode, const std::string& weights_npy = std::string());
  TreeModel treeModel() const { return mode_; }
  // Key of the cylinder fits' hypothesis draws (default 0x5EED), before the
  // first keyframe: another key draws another hypothesis stream, e.g. to
  // measure how far the plain fit differs from itself.  No counterpart in
  // the reference.
  void setHypothesisSeed(uint64_t seed);
  // Arithmetic of the pose solves (sloam::sloam::setPosePrecision), for both
  // estimators; set before the device estimator is switched on.
  void setPosePrecision(int precision);

 private:
  // grow-only device buffer
  struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t n);
    ~DevBuf();
  };
  // one front-end plan with its buffers, pinned scan input and stream
  struct Slot {
    sloam_ctx* ctx = nullptr;  // the bridge's context, or one of its own
    bool own_ctx = false;
    sloam_frontend* fe = nullptr;
    sloam_frontend_out buf{};
    sloam_frontend_host host{};
    void* h_in = nullptr;   // pinned scan: xyzi then labels
    void* done = nullptr;   // hipEvent_t recorded after the front end
    std::chrono::steady_clock::time_point t_submit;
  };
  // cylinder fits of the nt trees of slot s's last front-end run: launches
  // them and the download of their models and returns (the host assembles
  // the estimator's input meanwhile); fitTreesCollect waits for them and
  // unpacks the models into fits.  launch_ms: host time of the launch.
  void fitTreesLaunch(const Slot& s, int nt);
  void fitTreesCollect(int nt, std::vector<CylinderFit>& fits, double launch_ms,
                       KeyframeTiming* timing);
  // (re)creates every slot's front-end plan from fe_cfg_
  void makeFrontend();
  // the estimator on slot s's front-end results; t0..t2: submit, upload done,
  // front end done
  // the device keyframe (setDeviceEstimator): stages the scan, the pose and
  // the map trees in pinned memory, enqueues (or replays) the sequence,
  // unpacks the results into out
  bool runDevice(const float* xyzi, const uint8_t* labels, const SE3& poseEstimate,
                 const MapView& mapModels, SloamOutput& out, KeyframeTiming* timing);
  void enqueueDevice(Slot& s);
  bool estimate(const Slot& s, const SE3& poseEstimate, const MapView& mapModels,
                SloamOutput& out, KeyframeTiming* timing, std::chrono::steady_clock::time_point t0,
                std::chrono::steady_clock::time_point t1,
                std::chrono::steady_clock::time_point t2);

  int H_, W_;
  sloam_frontend_cfg fe_cfg_{};
  std::vector<Slot> slots_;
  int next_ = 0;     // slot of the next run() / submit()
  int pending_ = 0;  // submitted and not finished
  FeatureModelParams fm_;
  sloam::sloam sloam_;
  TreeModel mode_ = TreeModel::MedianLines;
  uint64_t fit_seed_ = 0x5EED;
  InputRecorder recorder_;
  sloam_cyl_plan* plain_ = nullptr;
  sloam_cyl_plan* verify_ = nullptr;
  int32_t* h_off_ = nullptr;    // pinned: tree point offsets of the last fit
  size_t h_off_cap_ = 0;
  DevBuf d_params_, d_xyz_, d_off_, d_model_, d_prop_, d_flags_;
  double* h_model_ = nullptr;   // pinned: models of the last fit + fallback count
  size_t h_model_cap_ = 0;      // trees
  // device estimator: pose, map count and map trees staged in pinned memory
  // and uploaded in one copy
  bool device_est_ = false;
  sloam_kf_estimator* kf_est_ = nullptr;
  int kf_cap_trees_ = 0, kf_cap_map_ = 0;
  int64_t kf_cap_points_ = 0;
  DevBuf d_kf_in_;
  void* h_kf_in_ = nullptr;
  SloamStepInfo dev_step_;
  // the device keyframe's tree batch (offsets, xyz, count), counted fit
  // plans, proposals, fallback flags and models, and its captured graph
  sloam_cyl_plan* plain_c_ = nullptr;
  sloam_cyl_plan* verify_c_ = nullptr;
  DevBuf d_kt_off_, d_kt_xyz_, d_kt_n_, d_kt_prop_, d_kt_flags_, d_kt_model_;
  sloam_graph* kf_graph_ = nullptr;
};
