// One SLOAM keyframe on the GPU (front-end plan + estimator).
#include "keyframe.h"

#include <chrono>
#include <stdexcept>
#include <string>

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

}  // namespace

KeyframeRunner::KeyframeRunner(int height, int width, const FeatureModelParams& fmParams,
                               bool use_graph, float fov_up_deg, float fov_down_deg)
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
  // every tree's axis line, fitted inside the front end
  cfg.line = Cylinder::lineFitCfg();
  cfg.fit_lines = 1;
  check(sloam_frontend_create(sloam_gpu::context(), &cfg, &fe_), "sloam_frontend_create");
  check(sloam_frontend_buffers(fe_, &buf_), "sloam_frontend_buffers");
  check(sloam_frontend_host_view(fe_, &host_), "sloam_frontend_host_view");
  sloam_.setFmParams(fmParams);
}

KeyframeRunner::~KeyframeRunner() { sloam_frontend_destroy(fe_); }

bool KeyframeRunner::run(const float* xyzi, const uint8_t* labels, const SE3& poseEstimate,
                         const std::vector<Cylinder>& mapModels, SloamOutput& out,
                         KeyframeTiming* timing) {
  sloam_ctx* ctx = sloam_gpu::context();
  const size_t N = (size_t)H_ * W_;
  const auto t0 = Clock::now();
  check(sloam_memcpy_async(ctx, buf_.xyzi, xyzi, 16 * N), "upload xyzi");
  check(sloam_memcpy_async(ctx, buf_.label, labels, N), "upload labels");
  check(sloam_sync(ctx), "sloam_sync");
  const auto t1 = Clock::now();
  check(sloam_frontend_exec(fe_), "sloam_frontend_exec");
  check(sloam_sync(ctx), "sloam_sync");
  const auto t2 = Clock::now();
  // results: already in pinned host memory (written by the front end's last
  // kernel)
  const sloam_frontend_host& hv = host_;
  const int32_t* cnt = hv.trellis_counts;
  if (cnt[4]) throw std::runtime_error("keyframe: trellis output capacity exceeded");
  const int nt = cnt[0], nc = buf_.n_cells;
  const int nf = fm_.numGroundFeatures;
  const int32_t *to = hv.tree_offsets, *vi = hv.vertex_int, *po = hv.point_offsets,
                *info = hv.cell_info;
  const double *vf = hv.vertex_f64, *model = hv.cell_model;
  const float *pts = hv.points, *feat = hv.cell_features;
  const auto t3 = Clock::now();

  SloamInput in;
  in.poseEstimate = poseEstimate;
  in.mapModels = mapModels;
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
    LineFit fit;
    fit.ok = hv.line_stats[4 * (size_t)t] != 0;
    fit.inliers = hv.line_stats[4 * (size_t)t + 1];
    for (int i = 0; i < 6; ++i) fit.coeff[i] = hv.line_coeff[6 * (size_t)t + i];
    in.treeLineFits.push_back(fit);
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
    in.groundCells.emplace_back(pp, f, valid);
  }
  const auto t3b = Clock::now();
  const bool ok = sloam_.RunSloam(in, out);
  const auto t4 = Clock::now();
  if (timing) {
    timing->upload_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    timing->frontend_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    timing->download_ms = std::chrono::duration<double, std::milli>(t3 - t2).count();
    timing->estimator_ms = std::chrono::duration<double, std::milli>(t4 - t3).count();
    timing->build_ms = std::chrono::duration<double, std::milli>(t3b - t3).count();
    timing->phases = sloam_.lastPhaseTimes();
    timing->total_ms = ms_since(t0);
    timing->trees = (int)out.tm.size();
    timing->planes = (int)sloam_.getPrevGroundModel().size();
  }
  return ok;
}
