// Keyframe-stream benchmark of the GPU SLOAM path (KeyframeRunner): runs a
// recorded sequence of organised scans through front end + estimator and
// prints one JSON object with the per-keyframe latency percentiles, stage
// medians and the pose drift against the sequence's true poses.
//
//   keyframe_bench <sequence file> [passes] [use_graph 0/1]
//
// Sequence file (little endian): int32 K, H, W; then per keyframe:
// float64 true pose [7] (t, q x y z w; map <- sensor), float64 odometry
// increment [7] (noisy relative motion from the previous keyframe), float32
// xyzi [H * W * 4], uint8 labels [H * W].  The pose estimate of keyframe k is
// the previous estimate composed with the increment; the map is the previous
// keyframe's tree models.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "keyframe.h"

namespace {

struct Frame {
  double truth[7], odom[7];
  std::vector<float> xyzi;
  std::vector<uint8_t> label;
};

SE3 to_se3(const double* p) { return SE3(Quat(p[6], p[3], p[4], p[5]), Vector3(p[0], p[1], p[2])); }

double pct(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t i = (size_t)pos;
  const double f = pos - (double)i;
  return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <sequence> [passes] [use_graph]\n", argv[0]);
    return 2;
  }
  const int passes = argc > 2 ? std::atoi(argv[2]) : 3;
  const bool use_graph = argc > 3 ? std::atoi(argv[3]) != 0 : true;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) {
    std::fprintf(stderr, "cannot open %s\n", argv[1]);
    return 2;
  }
  int32_t hdr[3];
  if (std::fread(hdr, 4, 3, f) != 3) return 2;
  const int K = hdr[0], H = hdr[1], W = hdr[2];
  std::vector<Frame> seq(K);
  for (auto& fr : seq) {
    fr.xyzi.resize((size_t)H * W * 4);
    fr.label.resize((size_t)H * W);
    if (std::fread(fr.truth, 8, 7, f) != 7 || std::fread(fr.odom, 8, 7, f) != 7 ||
        std::fread(fr.xyzi.data(), 4, fr.xyzi.size(), f) != fr.xyzi.size() ||
        std::fread(fr.label.data(), 1, fr.label.size(), f) != fr.label.size()) {
      std::fprintf(stderr, "short sequence file\n");
      return 2;
    }
  }
  std::fclose(f);

  FeatureModelParams fm;  // sloam.yaml
  std::vector<double> total, fe, up, down, est, build, ph_models, ph_match, ph_pose, ph_finish;
  // per sample: total, keyframe, upload, front end, estimator, models, pose
  std::vector<std::array<double, 7>> slow;
  std::vector<double> first;  // the warm-up pass
  double trees = 0, planes = 0;
  int samples = 0, ok_count = 0;
  double final_drift = 0, max_drift = 0, max_rot = 0;
  try {
    for (int pass = 0; pass < passes; ++pass) {
      KeyframeRunner kf(H, W, fm, use_graph);
      SE3 estimate = to_se3(seq[0].truth);
      std::vector<Cylinder> map;
      for (int k = 0; k < K; ++k) {
        if (k > 0) estimate = estimate * to_se3(seq[k].odom);
        SloamOutput out;
        KeyframeTiming t;
        const bool ok = kf.run(seq[k].xyzi.data(), seq[k].label.data(), estimate, map, out, &t);
        if (ok) estimate = out.T_Map_Curr;
        // the map for the next keyframe: this keyframe's trees (world frame)
        if (!out.tm.empty()) map = out.tm;
        if (pass == 0 && k > 0) {
          const SE3 truth = to_se3(seq[k].truth);
          const double d = (estimate.translation() - truth.translation()).norm();
          const Quat dq = estimate.unit_quaternion().conjugate() * truth.unit_quaternion();
          const double ang = 2.0 * std::acos(std::min(1.0, std::fabs(dq.w))) * 180.0 / M_PI;
          final_drift = d;
          max_drift = std::max(max_drift, d);
          max_rot = std::max(max_rot, ang);
        }
        if (k == 0) continue;  // first scan: map initialisation (and graph capture)
        if (pass == 0 && passes > 1) {
          // warm-up pass (first use of every buffer, plan and kernel of the
          // process): reported apart
          first.push_back(t.total_ms);
          continue;
        }
        total.push_back(t.total_ms);
        slow.push_back({t.total_ms, (double)k, t.upload_ms, t.frontend_ms, t.estimator_ms,
                        t.phases.models, t.phases.pose});
        fe.push_back(t.frontend_ms);
        up.push_back(t.upload_ms);
        down.push_back(t.download_ms);
        est.push_back(t.estimator_ms);
        build.push_back(t.build_ms);
        ph_models.push_back(t.phases.models);
        ph_match.push_back(t.phases.match);
        ph_pose.push_back(t.phases.pose);
        ph_finish.push_back(t.phases.finish);
        trees += t.trees;
        planes += t.planes;
        ++samples;
        ok_count += ok ? 1 : 0;
      }
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "keyframe_bench: %s\n", e.what());
    return 1;
  }
  std::printf(
      "{\"keyframes\": %d, \"passes\": %d, \"height\": %d, \"width\": %d, \"graph\": %s, "
      "\"latency_ms\": {\"p50\": %.4f, \"p99\": %.4f, \"mean\": %.4f, \"max\": %.4f}, "
      "\"stage_p50_ms\": {\"upload\": %.4f, \"frontend\": %.4f, \"download\": %.4f, "
      "\"estimator\": %.4f}, \"estimator_p50_ms\": {\"build\": %.4f, \"models\": %.4f, "
      "\"match\": %.4f, \"pose\": %.4f, \"finish\": %.4f}, "
      "\"frontend_p99_ms\": %.4f, \"trees_mean\": %.2f, "
      "\"planes_mean\": %.2f, \"success_rate\": %.4f, \"drift\": {\"final_m\": %.4f, "
      "\"max_m\": %.4f, \"max_deg\": %.4f}, \"warmup_pass_latency_ms\": {\"p50\": %.4f, "
      "\"max\": %.4f}}\n",
      K, passes, H, W, use_graph ? "true" : "false", pct(total, 0.5), pct(total, 0.99),
      samples ? [&] { double s = 0; for (double x : total) s += x; return s / samples; }() : 0.0,
      total.empty() ? 0.0 : *std::max_element(total.begin(), total.end()), pct(up, 0.5),
      pct(fe, 0.5), pct(down, 0.5), pct(est, 0.5), pct(build, 0.5), pct(ph_models, 0.5),
      pct(ph_match, 0.5), pct(ph_pose, 0.5), pct(ph_finish, 0.5), pct(fe, 0.99),
      samples ? trees / samples : 0.0, samples ? planes / samples : 0.0,
      samples ? (double)ok_count / samples : 0.0, final_drift, max_drift, max_rot,
      pct(first, 0.5), first.empty() ? 0.0 : *std::max_element(first.begin(), first.end()));
  // the five slowest keyframes, on stderr
  std::sort(slow.begin(), slow.end(), [](const auto& a, const auto& b) { return a[0] > b[0]; });
  for (size_t i = 0; i < slow.size() && i < 5; ++i)
    std::fprintf(stderr,
                 "slow keyframe %d: total %.3f ms (upload %.3f, front end %.3f, estimator %.3f: "
                 "models %.3f, pose %.3f)\n",
                 (int)slow[i][1], slow[i][0], slow[i][2], slow[i][3], slow[i][4], slow[i][5],
                 slow[i][6]);
  return 0;
}
