// One SLOAM keyframe on the GPU (front-end plan + estimator).
#include "keyframe.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>

#include "gpu_bridge.h"

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

void check(int st, const char* what) {
  if (st != SLOAM_OK) {
    std::string msg = std::string(what) + ": " + sloam_strerror(st);
    const char* d = sloam_ctx_last_error(sloam_gpu::context());
    if (d && *d) msg += std::string(" (") + d + ")";
    throw std::runtime_error(msg);
  }
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// n little-endian float32 values of a version 1 / 2 .npy file holding
// exactly that 1-D array
std::vector<float> load_npy_f32(const std::string& path, size_t n) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot open " + path);
  std::vector<float> out;
  std::string err;
  unsigned char pre[10];
  if (std::fread(pre, 1, 8, f) != 8 || std::memcmp(pre, "\x93NUMPY", 6) != 0) {
    err = "not an .npy file";
  } else {
    const int major = pre[6];
    uint32_t hlen = 0;
    if (major == 1) {
      unsigned char b[2];
      if (std::fread(b, 1, 2, f) == 2) hlen = b[0] | (b[1] << 8);
    } else {
      unsigned char b[4];
      if (std::fread(b, 1, 4, f) == 4) hlen = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24);
    }
    std::string hdr(hlen, '\0');
    if (hlen == 0 || std::fread(&hdr[0], 1, hlen, f) != hlen) {
      err = "short header";
    } else if (hdr.find("'<f4'") == std::string::npos ||
               hdr.find("'fortran_order': False") == std::string::npos ||
               hdr.find("(" + std::to_string(n) + ",)") == std::string::npos) {
      err = "expected float32 [" + std::to_string(n) + "]";
    } else {
      out.resize(n);
      if (std::fread(out.data(), 4, n, f) != n) err = "short data";
    }
  }
  std::fclose(f);
  if (!err.empty()) throw std::runtime_error(path + ": " + err);
  return out;
}

}  // namespace

void KeyframeRunner::DevBuf::ensure(size_t n) {
  if (n <= bytes) return;
  if (p) hip_check(hipFree(p), "hipFree");
  p = nullptr;
  bytes = 0;
  const size_t sz = n + n / 2;
  hip_check(hipMalloc(&p, sz), "hipMalloc");
  bytes = sz;
}

KeyframeRunner::DevBuf::~DevBuf() {
  if (p) (void)hipFree(p);
}

float* KeyframeRunner::pinnedXyzi() {
  Slot& sl = slots_[next_];
  if (!sl.h_in) {
    const size_t N = (size_t)H_ * W_;
    hip_check(hipHostMalloc(&sl.h_in, 17 * N), "hipHostMalloc");
  }
  return static_cast<float*>(sl.h_in);
}

uint8_t* KeyframeRunner::pinnedLabels() {
  return reinterpret_cast<uint8_t*>(pinnedXyzi() + 4 * (size_t)H_ * W_);
}

void KeyframeRunner::setTreeModel(TreeModel mode, const std::string& weights_npy) {
  if (device_est_)
    throw std::runtime_error("keyframe: set the tree model before switching to the device keyframe");
  if (mode == TreeModel::Learned) {
    if (weights_npy.empty()) throw std::runtime_error("keyframe: the learned model needs weights");
    const std::vector<float> w = load_npy_f32(weights_npy, SLOAM_PN_PARAMS);
    d_params_.ensure(sizeof(float) * w.size());
    hip_check(hipMemcpy(d_params_.p, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice),
              "hipMemcpy(weights)");
  }
  // the front end fits the trees' median lines only for the model that
  // uses them
  const int want_lines = mode == TreeModel::MedianLines ? 1 : 0;
  if (want_lines != fe_cfg_.fit_lines) {
    fe_cfg_.fit_lines = want_lines;
    makeFrontend();
  }
  mode_ = mode;
}

void KeyframeRunner::fitTreesLaunch(const Slot& sl, int nt) {
  sloam_ctx* ctx = sloam_gpu::context();
  const int32_t *to = sl.host.tree_offsets, *po = sl.host.point_offsets;
  // tree t's points are those of its vertices, contiguous in the front end's
  // point list
  const int32_t base = po[to[0]];
  if (h_off_cap_ < (size_t)nt + 1) {
    if (h_off_) hip_check(hipHostFree(h_off_), "hipHostFree");
    h_off_ = nullptr;
    h_off_cap_ = 0;
    const size_t cap = (size_t)nt + 1 + nt / 2;
    hip_check(hipHostMalloc((void**)&h_off_, sizeof(int32_t) * cap), "hipHostMalloc");
    h_off_cap_ = cap;
  }
  for (int t = 0; t <= nt; ++t) h_off_[t] = po[to[t]] - base;
  const int64_t np = h_off_[nt];
  // plans: built on the first keyframe, laid out again on every other
  sloam_cyl_fit_cfg cfg{};
  cfg.n_hyp = 1024;
  cfg.thr = 0.05f;
  cfg.max_radius = 1.0f;
  cfg.refit_iters = 2;
  cfg.refit_rounds = 3;
  cfg.seed = fit_seed_;
  cfg.stream_tag = 0;
  cfg.hyp_chunk = 128;
  cfg.tile_points = 1024;
  if (!plain_) {
    check(sloam_cyl_plan_create(ctx, h_off_, nt, &cfg, &plain_), "sloam_cyl_plan_create");
  } else {
    check(sloam_cyl_plan_reset(plain_, h_off_, nt), "sloam_cyl_plan_reset");
  }
  const bool learned = mode_ == TreeModel::Learned;
  if (learned) {
    sloam_cyl_fit_cfg vc = cfg;
    vc.n_hyp = 16;
    vc.stream_tag = 1;
    vc.hyp_chunk = 16;
    if (!verify_) {
      check(sloam_cyl_plan_create(ctx, h_off_, nt, &vc, &verify_), "sloam_cyl_plan_create");
    } else {
      check(sloam_cyl_plan_reset(verify_, h_off_, nt), "sloam_cyl_plan_reset");
    }
  }
  d_xyz_.ensure(sizeof(float) * 3 * (size_t)std::max<int64_t>(np, 1));
  d_off_.ensure(sizeof(int32_t) * ((size_t)nt + 1));
  // the models, then the fallback count (one download for both)
  d_model_.ensure(sizeof(double) * (SLOAM_CYL_MODEL_STRIDE * (size_t)nt + 1));
  d_prop_.ensure(sizeof(float) * SLOAM_PN_PROPOSAL_STRIDE * (size_t)nt);
  d_flags_.ensure(sizeof(int32_t) * (size_t)nt);
  if (h_model_cap_ < (size_t)nt) {
    if (h_model_) hip_check(hipHostFree(h_model_), "hipHostFree");
    h_model_ = nullptr;
    h_model_cap_ = 0;
    const size_t cap = (size_t)nt + nt / 2;
    hip_check(hipHostMalloc((void**)&h_model_, sizeof(double) * (SLOAM_CYL_MODEL_STRIDE * cap + 1)),
              "hipHostMalloc");
    h_model_cap_ = cap;
  }
  float* xyz = static_cast<float*>(d_xyz_.p);
  int32_t* off = static_cast<int32_t*>(d_off_.p);
  double* model = static_cast<double*>(d_model_.p);
  int32_t* flags = static_cast<int32_t*>(d_flags_.p);
  int32_t* n_fallback = reinterpret_cast<int32_t*>(model + SLOAM_CYL_MODEL_STRIDE * (size_t)nt);
  check(sloam_memcpy_async(ctx, off, h_off_, sizeof(int32_t) * ((size_t)nt + 1)),
        "upload tree offsets");
  check(sloam_xyz_from_xyzi(ctx, sl.buf.points + 4 * (size_t)base, np, xyz),
        "sloam_xyz_from_xyzi");
  if (learned) {
    float* prop = static_cast<float*>(d_prop_.p);
    check(sloam_pn_infer(ctx, static_cast<const float*>(d_params_.p), xyz, off, nt, prop),
          "sloam_pn_infer");
    check(sloam_cyl_learned_exec(verify_, plain_, xyz, off, prop, 0.5f, nullptr, nullptr, flags,
                                 n_fallback, nullptr, model),
          "sloam_cyl_learned_exec");
  } else {
    check(sloam_cyl_plan_exec(plain_, xyz, off, nullptr, nullptr, nullptr, nullptr, model),
          "sloam_cyl_plan_exec");
  }
  check(sloam_memcpy_async(ctx, h_model_, model,
                           sizeof(double) * (SLOAM_CYL_MODEL_STRIDE * (size_t)nt + (learned ? 1 : 0))),
        "download models");
}

void KeyframeRunner::setHypothesisSeed(uint64_t seed) {
  if (plain_ || verify_ || plain_c_ || verify_c_)
    throw std::runtime_error("keyframe: the hypothesis seed is set before the first keyframe");
  fit_seed_ = seed;
}

void KeyframeRunner::setPosePrecision(int precision) {
  if (kf_est_ && precision != sloam_.posePrecision())
    throw std::runtime_error("keyframe: the pose precision is set before the device estimator");
  sloam_.setPosePrecision(precision);
}

void KeyframeRunner::setDeviceEstimator(bool on) {
  if (on && mode_ == TreeModel::MedianLines)
    throw std::runtime_error("keyframe: the device keyframe serves the cylinder-fit tree models");
  if (on && slots_.size() != 1)
    throw std::runtime_error("keyframe: the device keyframe runs one scan at a time");
  sloam_ctx* ctx = sloam_gpu::context();
  if (on && !kf_est_) {
    const sloam_frontend_out& fo = slots_[0].buf;
    kf_cap_trees_ = std::min(fo.cap_trees, 512);
    kf_cap_points_ = fo.cap_points;
    kf_cap_map_ = 256;
    sloam_kf_params p{};
    p.max_lidar_dist = fm_.maxLidarDist;
    p.default_tree_radius = fm_.defaultTreeRadius;
    p.max_tree_radius = fm_.maxTreeRadius;
    p.max_axis_theta = fm_.maxAxisTheta;
    p.tree_match_thresh = fm_.treeMatchThresh;
    p.add_new_tree_thresh_dist = fm_.AddNewTreeThreshDist;
    p.min_tree_models = fm_.minTreeModels;
    p.min_ground_models = fm_.minGroundModels;
    p.min_planes = (fm_.groundRadiiBins * fm_.groundThetaBins) * 0.1;  // sloam.h's minPlanes_
    p.features_per_tree = fm_.featuresPerTree;
    p.num_ground_features = fm_.numGroundFeatures;
    p.two_step = fm_.twoStepOptim ? 1 : 0;
    p.pose_precision = sloam_.posePrecision();
    check(sloam_kf_estimator_create(ctx, &p, kf_cap_trees_, std::max(fo.n_cells, 1), kf_cap_map_,
                                    &kf_est_),
          "sloam_kf_estimator_create");
    const size_t in_bytes = sizeof(double) * (8 + 8 * (size_t)kf_cap_map_);
    d_kf_in_.ensure(in_bytes);
    hip_check(hipHostMalloc(&h_kf_in_, in_bytes), "hipHostMalloc");
    std::memset(h_kf_in_, 0, in_bytes);
    d_kt_off_.ensure(sizeof(int32_t) * ((size_t)kf_cap_trees_ + 1));
    d_kt_xyz_.ensure(sizeof(float) * 3 * (size_t)kf_cap_points_);
    d_kt_n_.ensure(sizeof(int32_t) * 4);
    d_kt_prop_.ensure(sizeof(float) * SLOAM_PN_PROPOSAL_STRIDE * (size_t)kf_cap_trees_);
    d_kt_flags_.ensure(sizeof(int32_t) * ((size_t)kf_cap_trees_ + 1));
    d_kt_model_.ensure(sizeof(double) * SLOAM_CYL_MODEL_STRIDE * (size_t)kf_cap_trees_);
  }
  if (on && (!plain_c_ || (mode_ == TreeModel::Learned && !verify_c_))) {
    // the settings of fitTreesLaunch's plans, on counted plans
    sloam_cyl_fit_cfg cfg{};
    cfg.n_hyp = 1024;
    cfg.thr = 0.05f;
    cfg.max_radius = 1.0f;
    cfg.refit_iters = 2;
    cfg.refit_rounds = 3;
    cfg.seed = fit_seed_;
    cfg.stream_tag = 0;
    cfg.hyp_chunk = 128;
    cfg.tile_points = 1024;
    const int32_t* d_n = static_cast<const int32_t*>(d_kt_n_.p);
    if (!plain_c_)
      check(sloam_cyl_plan_create_counted(ctx, kf_cap_trees_, kf_cap_points_, d_n, &cfg, &plain_c_),
            "sloam_cyl_plan_create_counted");
    if (mode_ == TreeModel::Learned && !verify_c_) {
      sloam_cyl_fit_cfg vc = cfg;
      vc.n_hyp = 16;
      vc.stream_tag = 1;
      vc.hyp_chunk = 16;
      check(sloam_cyl_plan_create_counted(ctx, kf_cap_trees_, kf_cap_points_, d_n, &vc, &verify_c_),
            "sloam_cyl_plan_create_counted");
    }
  }
  if (on && !device_est_) check(sloam_kf_estimator_reset(kf_est_), "sloam_kf_estimator_reset");
  // a captured keyframe holds the tree model it was captured with
  if (kf_graph_) {
    sloam_graph_destroy(kf_graph_);
    kf_graph_ = nullptr;
  }
  device_est_ = on;
}

void KeyframeRunner::enqueueDevice(Slot& sl) {
  sloam_ctx* ctx = sloam_gpu::context();
  const size_t N = (size_t)H_ * W_;
  check(sloam_memcpy_async(ctx, sl.buf.xyzi, pinnedXyzi(), 16 * N), "upload xyzi");
  check(sloam_memcpy_async(ctx, sl.buf.label, pinnedLabels(), N), "upload labels");
  check(sloam_memcpy_async(ctx, d_kf_in_.p, h_kf_in_, sizeof(double) * (8 + 8 * (size_t)kf_cap_map_)),
        "upload pose and map");
  check(sloam_frontend_enqueue(sl.fe), "sloam_frontend_enqueue");
  int32_t* off = static_cast<int32_t*>(d_kt_off_.p);
  float* xyz = static_cast<float*>(d_kt_xyz_.p);
  int32_t* d_n = static_cast<int32_t*>(d_kt_n_.p);
  double* model = static_cast<double*>(d_kt_model_.p);
  check(sloam_kf_tree_points(ctx, &sl.buf, kf_cap_trees_, kf_cap_points_, off, xyz, d_n),
        "sloam_kf_tree_points");
  if (mode_ == TreeModel::Learned) {
    float* prop = static_cast<float*>(d_kt_prop_.p);
    int32_t* flags = static_cast<int32_t*>(d_kt_flags_.p);
    check(sloam_pn_infer_counted(ctx, static_cast<const float*>(d_params_.p), xyz, off,
                                 kf_cap_trees_, d_n, prop),
          "sloam_pn_infer_counted");
    check(sloam_cyl_learned_exec(verify_c_, plain_c_, xyz, off, prop, 0.5f, nullptr, nullptr, flags,
                                 flags + kf_cap_trees_, nullptr, model),
          "sloam_cyl_learned_exec");
  } else {
    check(sloam_cyl_plan_exec(plain_c_, xyz, off, nullptr, nullptr, nullptr, nullptr, model),
          "sloam_cyl_plan_exec");
  }
  const double* d = static_cast<const double*>(d_kf_in_.p);
  sloam_kf_inputs in{};
  in.trellis_counts = sl.buf.trellis_counts;
  in.tree_offsets = sl.buf.tree_offsets;
  in.vertex_int = sl.buf.vertex_int;
  in.vertex_f64 = sl.buf.vertex_f64;
  in.point_offsets = sl.buf.point_offsets;
  in.points = sl.buf.points;
  in.n_cells = sl.buf.n_cells;
  in.num_features = fm_.numGroundFeatures;
  in.cell_info = sl.buf.cell_info;
  in.cell_model = sl.buf.cell_model;
  in.cell_features = sl.buf.cell_features;
  in.tree_fits = model;
  in.pose = d;
  in.n_map = reinterpret_cast<const int32_t*>(d + 7);
  in.map = d + 8;
  check(sloam_kf_estimator_exec(kf_est_, &in), "sloam_kf_estimator_exec");
}

bool KeyframeRunner::runDevice(const float* xyzi, const uint8_t* labels, const SE3& poseEstimate,
                               const MapView& mapModels, SloamOutput& out,
                               KeyframeTiming* timing) {
  sloam_ctx* ctx = sloam_gpu::context();
  Slot& sl = slots_[0];
  const size_t N = (size_t)H_ * W_;
  const auto t0 = Clock::now();
  // the scan, the pose (t, q x y z w), the map count (first word of slot 7)
  // and the map trees, staged in pinned memory the sequence uploads from
  if (xyzi != pinnedXyzi()) std::memcpy(pinnedXyzi(), xyzi, 16 * N);
  if (labels != pinnedLabels()) std::memcpy(pinnedLabels(), labels, N);
  const int K = (int)mapModels.size();
  if (K > kf_cap_map_) throw std::runtime_error("keyframe: more map trees than the estimator holds");
  double* h = static_cast<double*>(h_kf_in_);
  {
    const Vector3& t = poseEstimate.translation();
    const Quat& q = poseEstimate.unit_quaternion();
    const double x[7] = {t[0], t[1], t[2], q.x, q.y, q.z, q.w};
    std::memcpy(h, x, sizeof(x));
    const int32_t k32 = K;
    h[7] = 0.0;
    std::memcpy(h + 7, &k32, sizeof(k32));
    for (int i = 0; i < K; ++i) {
      const CylinderParameters& m = mapModels[i]->model;
      double* row = h + 8 + 8 * (size_t)i;
      for (int c = 0; c < 3; ++c) {
        row[c] = m.root[c];
        row[3 + c] = m.ray[c];
      }
      row[6] = m.radius;
      row[7] = 0.0;
    }
  }
  const auto t1 = Clock::now();
  if (kf_graph_) {
    check(sloam_graph_launch(kf_graph_), "sloam_graph_launch");
  } else {
    // the first keyframe runs eagerly (sizing every scratch buffer), then
    // the sequence is captured for the next ones
    enqueueDevice(sl);
    if (fe_cfg_.use_graph) {
      check(sloam_sync(ctx), "sloam_sync");
      check(sloam_graph_begin(ctx), "sloam_graph_begin");
      enqueueDevice(sl);
      sloam_graph* g = nullptr;
      check(sloam_graph_end(ctx, &g), "sloam_graph_end");
      kf_graph_ = g;
      // the captured copy of the sequence has not run: the eager one has
    }
  }
  check(sloam_sync(ctx), "sloam_sync");
  const auto t2 = Clock::now();
  sloam_kf_result_view r{};
  check(sloam_kf_estimator_result(kf_est_, &r), "sloam_kf_estimator_result");
  const int32_t* hd = r.head;
  // header words: see sloam_kf_result_view
  if (hd[10]) throw std::runtime_error("keyframe: a device keyframe capacity was exceeded");
  dev_step_ = SloamStepInfo{hd[3] != 0, hd[4] != 0, hd[5], hd[6], hd[18], hd[19]};
  bool ok = false;
  if (hd[17]) {
    out.T_Map_Curr = SE3::fromUnitQuaternion(Quat(r.pose[6], r.pose[3], r.pose[4], r.pose[5]),
                                             Vector3(r.pose[0], r.pose[1], r.pose[2]));
    out.T_Delta = SE3::fromUnitQuaternion(Quat(r.delta[6], r.delta[3], r.delta[4], r.delta[5]),
                                          Vector3(r.delta[0], r.delta[1], r.delta[2]));
    const int n_lm = hd[2];
    out.matches.assign(r.lm_match, r.lm_match + n_lm);
    out.tm.clear();
    out.tm.reserve(n_lm);
    const int32_t *to = sl.host.tree_offsets, *po = sl.host.point_offsets;
    const float* pts = sl.host.points;
    const int fpt = fm_.featuresPerTree;
    for (int l = 0; l < n_lm; ++l) {
      CylinderParameters m;
      const double* q = r.lm_model + 8 * (size_t)l;
      m.root = Vector3(q[0], q[1], q[2]);
      m.ray = Vector3(q[3], q[4], q[5]);
      m.radius = q[6];
      m.lambda = 1.0;
      // the tree's first featuresPerTree points (zero points past its last),
      // intensity = tree id (cylinder.cpp:86-96, 172)
      const int t = r.lm_tree[l];
      const int pb = po[to[t]], pe = po[to[t + 1]];
      VectorType f((size_t)fpt);
      for (int j = 0; j < fpt && pb + j < pe; ++j) {
        const float* p = pts + 4 * (size_t)(pb + j);
        f[j] = PointT{p[0], p[1], p[2], (float)r.lm_id[l]};
      }
      out.tm.emplace_back(m, (size_t)r.lm_id[l], std::move(f));
    }
    ok = hd[16] != 0;
  }
  if (timing) {
    *timing = KeyframeTiming();
    timing->upload_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    timing->frontend_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    timing->estimator_ms = ms_since(t2);
    timing->total_ms = ms_since(t0);
    timing->trees = (int)out.tm.size();
    timing->planes = hd[1];
    timing->fitted = hd[11];
  }
  return ok;
}

void KeyframeRunner::fitTreesCollect(int nt, std::vector<CylinderFit>& fits, double launch_ms,
                                     KeyframeTiming* timing) {
  const auto t0 = Clock::now();
  const bool learned = mode_ == TreeModel::Learned;
  int32_t fb = 0;
  check(sloam_sync(sloam_gpu::context()), "sloam_sync");
  if (learned) std::memcpy(&fb, h_model_ + SLOAM_CYL_MODEL_STRIDE * (size_t)nt, sizeof(int32_t));
  fits.assign(nt, CylinderFit());
  for (int t = 0; t < nt; ++t) {
    const double* m = h_model_ + SLOAM_CYL_MODEL_STRIDE * (size_t)t;
    CylinderFit& c = fits[t];
    for (int i = 0; i < 3; ++i) {
      c.root[i] = m[i];
      c.ray[i] = m[3 + i];
    }
    c.radius = m[6];
    c.inliers = (int)m[7];
    c.ok = m[6] > 0.0 && m[7] >= 3.0;
  }
  if (timing) {
    timing->fit_ms = launch_ms + ms_since(t0);
    timing->fitted = nt;
    timing->fallback = fb;
  }
}

KeyframeRunner::KeyframeRunner(int height, int width, const FeatureModelParams& fmParams,
                               bool use_graph, float fov_up_deg, float fov_down_deg,
                               bool pipelined)
    : H_(height), W_(width), fm_(fmParams) {
  sloam_frontend_cfg cfg{};
  cfg.proj.height = height;
  cfg.proj.width = width;
  cfg.proj.fov_up_deg = fov_up_deg;
  cfg.proj.fov_down_deg = fov_down_deg;
  cfg.proj.destagger = 0;
  // Instance parameters (trellis.cpp; sloamNode.cpp:60-70)
  cfg.trellis.cluster_dist = 1.0f;
  cfg.trellis.max_dist_to_centroid = 0.2f;
  cfg.trellis.min_label_points = 80;
  cfg.trellis.min_row_points = 3;
  cfg.trellis.min_vertices = 16;
  cfg.trellis.max_vertices = 56;
  cfg.trellis.radius_scale = 1.0;
  // ground cells of computeModels (binned at the sensor origin)
  cfg.ground.radii_bins = fmParams.groundRadiiBins;
  cfg.ground.theta_bins = fmParams.groundThetaBins;
  cfg.ground.min_dist = fmParams.minGroundLidarDist;
  cfg.ground.max_dist = fmParams.maxGroundLidarDist;
  cfg.ground.retain_thresh = fmParams.groundRetainThresh;
  cfg.ground.num_features = fmParams.numGroundFeatures;
  cfg.use_graph = use_graph ? 1 : 0;
  // results land in pinned host memory at the end of every front-end run
  cfg.host_results = 1;
  // every tree's axis line, fitted inside the front end (the median-line
  // tree model; the cylinder-fit models rebuild the plan without them)
  cfg.line = Cylinder::lineFitCfg();
  cfg.fit_lines = 1;
  fe_cfg_ = cfg;
  // one slot on the estimator's context, or two with streams of their own
  slots_.resize(pipelined ? 2 : 1);
  for (Slot& sl : slots_) {
    if (pipelined) {
      check(sloam_ctx_create(sloam_gpu::deviceIndex(), &sl.ctx), "sloam_ctx_create");
      sl.own_ctx = true;
    } else {
      sl.ctx = sloam_gpu::context();
    }
    hipEvent_t ev = nullptr;
    hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
    sl.done = ev;
  }
  makeFrontend();
  sloam_.setFmParams(fmParams);
}

void KeyframeRunner::makeFrontend() {
  if (pending_) throw std::runtime_error("keyframe: front-end slots are busy");
  for (Slot& sl : slots_) {
    if (sl.fe) {
      check(sloam_sync(sl.ctx), "sloam_sync");
      sloam_frontend_destroy(sl.fe);
      sl.fe = nullptr;
    }
    check(sloam_frontend_create(sl.ctx, &fe_cfg_, &sl.fe), "sloam_frontend_create");
    check(sloam_frontend_buffers(sl.fe, &sl.buf), "sloam_frontend_buffers");
    check(sloam_frontend_host_view(sl.fe, &sl.host), "sloam_frontend_host_view");
  }
}

KeyframeRunner::~KeyframeRunner() {
  sloam_graph_destroy(kf_graph_);
  sloam_cyl_plan_destroy(plain_c_);
  sloam_cyl_plan_destroy(verify_c_);
  sloam_kf_estimator_destroy(kf_est_);
  if (h_kf_in_) (void)hipHostFree(h_kf_in_);
  sloam_cyl_plan_destroy(plain_);
  sloam_cyl_plan_destroy(verify_);
  if (h_model_) (void)hipHostFree(h_model_);

  if (h_off_) (void)hipHostFree(h_off_);
  for (Slot& sl : slots_) {
    if (sl.ctx) (void)sloam_sync(sl.ctx);
    sloam_frontend_destroy(sl.fe);
    if (sl.done) (void)hipEventDestroy((hipEvent_t)sl.done);
    if (sl.h_in) (void)hipHostFree(sl.h_in);
    if (sl.own_ctx) sloam_ctx_destroy(sl.ctx);
  }
}

namespace {
MapView view_of(const std::vector<Cylinder>& trees) {
  MapView v;
  v.reserve(trees.size());
  for (const Cylinder& c : trees) v.push_back(&c);
  return v;
}
}  // namespace

bool KeyframeRunner::run(const float* xyzi, const uint8_t* labels, const SE3& poseEstimate,
                         const std::vector<Cylinder>& mapModels, SloamOutput& out,
                         KeyframeTiming* timing) {
  return run(xyzi, labels, poseEstimate, view_of(mapModels), out, timing);
}

bool KeyframeRunner::finish(const SE3& poseEstimate, const std::vector<Cylinder>& mapModels,
                            SloamOutput& out, KeyframeTiming* timing) {
  return finish(poseEstimate, view_of(mapModels), out, timing);
}

bool KeyframeRunner::run(const float* xyzi, const uint8_t* labels, const SE3& poseEstimate,
                         const MapView& mapModels, SloamOutput& out, KeyframeTiming* timing) {
  if (pending_) throw std::runtime_error("keyframe: run() with submitted scans pending");
  if (device_est_) return runDevice(xyzi, labels, poseEstimate, mapModels, out, timing);
  Slot& sl = slots_[next_];
  next_ = (next_ + 1) % (int)slots_.size();
  const size_t N = (size_t)H_ * W_;
  const auto t0 = Clock::now();
  check(sloam_memcpy_async(sl.ctx, sl.buf.xyzi, xyzi, 16 * N), "upload xyzi");
  check(sloam_memcpy_async(sl.ctx, sl.buf.label, labels, N), "upload labels");
  check(sloam_sync(sl.ctx), "sloam_sync");
  const auto t1 = Clock::now();
  check(sloam_frontend_exec(sl.fe), "sloam_frontend_exec");
  check(sloam_sync(sl.ctx), "sloam_sync");
  return estimate(sl, poseEstimate, mapModels, out, timing, t0, t1, Clock::now());
}

void KeyframeRunner::submit(const float* xyzi, const uint8_t* labels) {
  if (device_est_) throw std::runtime_error("keyframe: the device keyframe runs through run()");
  if (pending_ == (int)slots_.size())
    throw std::runtime_error("keyframe: every front-end slot is busy (finish() first)");
  Slot& sl = slots_[next_];
  next_ = (next_ + 1) % (int)slots_.size();
  const size_t N = (size_t)H_ * W_;
  sl.t_submit = Clock::now();
  check(sloam_memcpy_async(sl.ctx, sl.buf.xyzi, xyzi, 16 * N), "upload xyzi");
  check(sloam_memcpy_async(sl.ctx, sl.buf.label, labels, N), "upload labels");
  check(sloam_frontend_exec(sl.fe), "sloam_frontend_exec");
  hip_check(hipEventRecord((hipEvent_t)sl.done, (hipStream_t)sloam_ctx_stream(sl.ctx)),
            "hipEventRecord");
  ++pending_;
}

bool KeyframeRunner::finish(const SE3& poseEstimate, const MapView& mapModels, SloamOutput& out,
                            KeyframeTiming* timing) {
  if (!pending_) throw std::runtime_error("keyframe: finish() with no submitted scan");
  const int n = (int)slots_.size();
  Slot& sl = slots_[(next_ - pending_ + n) % n];
  hip_check(hipEventSynchronize((hipEvent_t)sl.done), "hipEventSynchronize");
  --pending_;
  return estimate(sl, poseEstimate, mapModels, out, timing, sl.t_submit, sl.t_submit,
                  Clock::now());
}

void KeyframeRunner::skip() {
  if (!pending_) throw std::runtime_error("keyframe: skip() with no submitted scan");
  const int n = (int)slots_.size();
  Slot& sl = slots_[(next_ - pending_ + n) % n];
  hip_check(hipEventSynchronize((hipEvent_t)sl.done), "hipEventSynchronize");
  --pending_;
}

bool KeyframeRunner::estimate(const Slot& sl, const SE3& poseEstimate,
                              const MapView& mapModels, SloamOutput& out,
                              KeyframeTiming* timing, Clock::time_point t0,
                              Clock::time_point t1, Clock::time_point t2) {
  if (timing) {
    timing->fit_ms = 0;
    timing->fitted = 0;
    timing->fallback = 0;
  }
  // results: already in pinned host memory (written by the front end's last
  // kernel)
  const sloam_frontend_host& hv = sl.host;
  const int32_t* cnt = hv.trellis_counts;
  if (cnt[4]) throw std::runtime_error("keyframe: trellis output capacity exceeded");
  const int nt = cnt[0], nc = sl.buf.n_cells;
  const int nf = fm_.numGroundFeatures;
  const int32_t *to = hv.tree_offsets, *vi = hv.vertex_int, *po = hv.point_offsets,
                *info = hv.cell_info;
  const double *vf = hv.vertex_f64, *model = hv.cell_model;
  const float *pts = hv.points, *feat = hv.cell_features;
  const auto t3 = Clock::now();
  // the tree fits run on the GPU while the host assembles the input below
  const bool fit_trees = mode_ != TreeModel::MedianLines && nt > 0;
  double fit_launch_ms = 0.0;
  if (fit_trees) {
    fitTreesLaunch(sl, nt);
    fit_launch_ms = ms_since(t3);
  }

  SloamInput in;  // the map trees go to RunSloam apart, not copied into it
  in.poseEstimate = poseEstimate;
  in.landmarks.reserve(nt);
  for (int t = 0; t < nt; ++t) {
    std::vector<TreeVertex> tree;
    tree.reserve(to[t + 1] - to[t]);
    for (int v = to[t]; v < to[t + 1]; ++v) {
      TreeVertex x;
      x.treeId = vi[4 * (size_t)v];
      x.isValid = vi[4 * (size_t)v + 3] != 0;
      x.radius = vf[4 * (size_t)v];
      x.coords.x = (float)vf[4 * (size_t)v + 1];
      x.coords.y = (float)vf[4 * (size_t)v + 2];
      x.coords.z = (float)vf[4 * (size_t)v + 3];
      x.points.reserve(po[v + 1] - po[v]);
      for (int p = po[v]; p < po[v + 1]; ++p)
        x.points.push_back(PointT{pts[4 * (size_t)p], pts[4 * (size_t)p + 1],
                                  pts[4 * (size_t)p + 2], pts[4 * (size_t)p + 3]});
      tree.push_back(std::move(x));
    }
    in.landmarks.push_back(std::move(tree));
    if (fe_cfg_.fit_lines) {
      LineFit fit;
      fit.ok = hv.line_stats[4 * (size_t)t] != 0;
      fit.inliers = hv.line_stats[4 * (size_t)t + 1];
      for (int i = 0; i < 6; ++i) fit.coeff[i] = hv.line_coeff[6 * (size_t)t + i];
      in.treeLineFits.push_back(fit);
    }
  }
  for (int c = 0; c < nc; ++c) {
    PlaneParameters pp;
    for (int i = 0; i < 4; ++i) pp.plane(i) = model[8 * (size_t)c + i];
    for (int i = 0; i < 3; ++i) pp.centroid(i) = model[8 * (size_t)c + 4 + i];
    const bool valid = info[4 * (size_t)c] != 0;
    VectorType f;
    if (valid)
      for (int j = 0; j < nf; ++j) {
        const float* q = &feat[4 * ((size_t)c * nf + j)];
        f.push_back(PointT{q[0], q[1], q[2], q[3]});
      }
    in.groundCells.emplace_back(pp, std::move(f), valid);
  }
  if (fit_trees) {
    // the tree models from cylinder fits of the trees' points
    fitTreesCollect(nt, in.treeCylinderFits, fit_launch_ms, timing);
    in.treeLineFits.clear();
  }
  const auto t3b = Clock::now();
  if (recorder_) recorder_(in, mapModels);
  const bool ok = sloam_.RunSloam(in, mapModels, out);
  const auto t4 = Clock::now();
  if (timing) {
    timing->upload_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    timing->frontend_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    timing->download_ms = std::chrono::duration<double, std::milli>(t3 - t2).count();
    timing->estimator_ms = std::chrono::duration<double, std::milli>(t4 - t3).count();
    timing->build_ms =
        std::chrono::duration<double, std::milli>(t3b - t3).count() - timing->fit_ms;
    timing->phases = sloam_.lastPhaseTimes();
    timing->total_ms = ms_since(t0);
    timing->trees = (int)out.tm.size();
    timing->planes = (int)sloam_.getPrevGroundModel().size();
  }
  return ok;
}
