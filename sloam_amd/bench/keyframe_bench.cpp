// Keyframe-stream benchmark of the GPU SLOAM path (KeyframeRunner): runs a
// recorded sequence of organised scans through front end + estimator and
// prints one JSON object with the per-keyframe latency percentiles, stage
// medians and the pose drift against the sequence's true poses.
//
//   keyframe_bench <sequence file> [passes] [use_graph 0/1]
//                  [--tree-model lines|plain|learned] [--weights <npy>]
//                  [--traj-out <file>] [--log] [--pinned] [--pipelined]
//                  [--shadow lines|plain|learned] [--device-estimator]
//                  [--compare-device]
//
// --tree-model picks how each tree's model is made (KeyframeRunner's
// TreeModel; default lines, the reference's); --traj-out writes the first
// pass's pose estimates, K x 7 float64 (t, q x y z w), to compare the
// trajectories of two tree models; --log prints the first pass's
// keyframes (trees, matches, pose error before and after the solve) to
// stderr; --pinned writes each scan into the runner's pinned input buffers
// (as a sensor driver would) before its timed call, instead of passing the
// sequence's pageable arrays; --pipelined submits each scan's front end
// before the previous keyframe's estimator runs (KeyframeRunner::submit /
// finish on two front-end slots).  Throughput is keyframes per second of
// each timed pass (the copies into the pinned buffers, which a sensor driver
// would make, are not counted).  --shadow runs, in the first pass only, a
// second runner with the named tree model on every keyframe, from the same
// pose estimate, submap and previous ground models as the first, and
// reports how far its pose lands from the first runner's (the one-keyframe
// difference of the two tree models, without the compounding of two maps
// built apart); it is not allowed with --pipelined.  --dump <file> writes,
// for every keyframe of the first pass, the estimator's input (as the
// runner hands it to RunSloam), its state before the call and its output
// (format: write_dump_* below), so that another implementation of RunSloam
// can be run on the same inputs.  --max-tree-radius and --tree-match-thresh
// replace those two estimator parameters (sloam.yaml: 0.3 and 0.5).
// --device-estimator runs the estimator on the GPU (sloam_kf_estimator)
// instead of the host sloam::sloam; --compare-device runs, in the first pass,
// a shadow with the same tree model and the device estimator from its own
// state, and counts the keyframes where its pose, landmarks and step
// decisions equal the host estimator's bit for bit.  --pose-precision 0
// runs the pose solves of both runners with FP32 residuals and compensated
// FP32 sums instead of FP64 (KeyframeRunner::setPosePrecision).  The per-keyframe
// latencies never include the shadow's work, but with a single pass that
// pass's throughput does: give two passes or more to time with --shadow.
//
// Sequence file (little endian): int32 K, H, W; then per keyframe:
// float64 true pose [7] (t, q x y z w; map <- sensor), float64 odometry
// increment [7] (noisy relative motion from the previous keyframe), float32
// xyzi [H * W * 4], uint8 labels [H * W].  The pose estimate of keyframe k is
// the previous estimate composed with the increment.  The map is kept as
// SLOAMNode::run keeps it (sloam/src/core/sloamNode.cpp:186-245): a
// MapManager submap around the estimate goes in, the keyframe's trees and
// their matches update the map; a keyframe after the first with an empty
// submap is discarded (its estimate stays the odometry's).
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "keyframe.h"
#include "mapManager.h"

namespace {

struct Frame {
  double truth[7], odom[7];
  std::vector<float> xyzi;
  std::vector<uint8_t> label;
};

SE3 to_se3(const double* p) { return SE3(Quat(p[6], p[3], p[4], p[5]), Vector3(p[0], p[1], p[2])); }

bool parse_tree_model(const std::string& name, TreeModel* m) {
  if (name == "lines") {
    *m = TreeModel::MedianLines;
  } else if (name == "plain") {
    *m = TreeModel::Plain;
  } else if (name == "learned") {
    *m = TreeModel::Learned;
  } else {
    return false;
  }
  return true;
}

bool same_se3(const SE3& a, const SE3& b) {
  const Quat &qa = a.unit_quaternion(), &qb = b.unit_quaternion();
  return std::memcmp(&qa, &qb, sizeof(Quat)) == 0 &&
         std::memcmp(&a.translation(), &b.translation(), sizeof(Vector3)) == 0;
}

bool same_model(const CylinderParameters& a, const CylinderParameters& b) {
  return std::memcmp(&a.root, &b.root, sizeof(Vector3)) == 0 &&
         std::memcmp(&a.ray, &b.ray, sizeof(Vector3)) == 0 &&
         std::memcmp(&a.radius, &b.radius, sizeof(double)) == 0;
}

// translation (m) and rotation (deg) between two poses
void pose_diff(const SE3& a, const SE3& b, double* m, double* deg) {
  const Quat dq = a.unit_quaternion().conjugate() * b.unit_quaternion();
  *deg = 2.0 * std::acos(std::min(1.0, std::fabs(dq.w))) * 180.0 / M_PI;
  *m = (a.translation() - b.translation()).norm();
}

double pct(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t i = (size_t)pos;
  const double f = pos - (double)i;
  return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

// --dump records (little endian), one per RunSloam call of the first pass:
//   int32 keyframe, int32 first call (1 / 0)
//   float64 pose estimate [7] (t, q x y z w)
//   int32 cells, int32 features per cell; per cell: float64 plane[4],
//     centroid[3], int32 valid, float32 features [nf][4] (zeros if invalid)
//   int32 trees; per tree: int32 vertices; per vertex: int32 tree id,
//     int32 valid, float64 radius, float64 x, y, z (the float coordinates),
//     int32 points, float32 points [n][4]
//   int32 line fits; per fit: int32 ok, int32 inliers, float32 coeff[6]
//   int32 cylinder fits; per fit: int32 ok, float64 root[3], ray[3], radius
//   int32 map trees; per tree: float64 root[3], ray[3], radius
//   int32 previous planes; per plane: float64 plane[4], centroid[3]
// then the output of the call:
//   int32 returned, int32 tree step, int32 ground step, int32 tree feature
//   matches, int32 plane feature matches; float64 T_Map_Curr[7],
//   T_Delta[7]; int32 trees kept; per tree: int32 id, float64 root[3],
//   ray[3], radius, int32 match; int32 planes after; per plane: float64
//   plane[4], centroid[3]
struct DumpWriter {
  FILE* f = nullptr;
  template <typename T>
  void put(const T& v) {
    std::fwrite(&v, sizeof(T), 1, f);
  }
  void pose(const SE3& p) {
    const Vector3 t = p.translation();
    const Quat q = p.unit_quaternion();
    for (double v : {t[0], t[1], t[2], q.x, q.y, q.z, q.w}) put(v);
  }
  void plane_model(const PlaneParameters& m) {
    for (int i = 0; i < 4; ++i) put((double)m.plane(i));
    for (int i = 0; i < 3; ++i) put((double)m.centroid(i));
  }
  void cyl_model(const CylinderParameters& m) {
    for (int i = 0; i < 3; ++i) put((double)m.root[i]);
    for (int i = 0; i < 3; ++i) put((double)m.ray[i]);
    put((double)m.radius);
  }
  void input(int k, bool first, const SloamInput& in, const MapView& map,
             const std::vector<Plane>& prev, int nf) {
    put((int32_t)k);
    put((int32_t)(first ? 1 : 0));
    pose(in.poseEstimate);
    put((int32_t)in.groundCells.size());
    put((int32_t)nf);
    for (const Plane& c : in.groundCells) {
      plane_model(c.model);
      put((int32_t)(c.isValid ? 1 : 0));
      for (int j = 0; j < nf; ++j) {
        PointT p{0, 0, 0, 0};
        if (c.isValid && j < (int)c.features.size()) p = c.features[j];
        for (float v : {p.x, p.y, p.z, p.intensity}) put(v);
      }
    }
    put((int32_t)in.landmarks.size());
    for (const auto& tree : in.landmarks) {
      put((int32_t)tree.size());
      for (const TreeVertex& v : tree) {
        put((int32_t)v.treeId);
        put((int32_t)(v.isValid ? 1 : 0));
        put((double)v.radius);
        for (float c : {v.coords.x, v.coords.y, v.coords.z}) put((double)c);
        put((int32_t)v.points.size());
        for (const PointT& p : v.points)
          for (float c : {p.x, p.y, p.z, p.intensity}) put(c);
      }
    }
    put((int32_t)in.treeLineFits.size());
    for (const LineFit& l : in.treeLineFits) {
      put((int32_t)(l.ok ? 1 : 0));
      put((int32_t)l.inliers);
      for (int i = 0; i < 6; ++i) put(l.coeff[i]);
    }
    put((int32_t)in.treeCylinderFits.size());
    for (const CylinderFit& c : in.treeCylinderFits) {
      put((int32_t)(c.ok ? 1 : 0));
      for (int i = 0; i < 3; ++i) put(c.root[i]);
      for (int i = 0; i < 3; ++i) put(c.ray[i]);
      put(c.radius);
    }
    put((int32_t)map.size());
    for (const Cylinder* c : map) cyl_model(c->model);
    put((int32_t)prev.size());
    for (const Plane& p : prev) plane_model(p.model);
  }
  void output(bool ok, const SloamStepInfo& st, const SloamOutput& out,
              const std::vector<Plane>& after) {
    put((int32_t)(ok ? 1 : 0));
    put((int32_t)(st.treeCheck ? 1 : 0));
    put((int32_t)(st.groundCheck ? 1 : 0));
    put((int32_t)st.treeMatches);
    put((int32_t)st.planeMatches);
    put((int32_t)st.treeIters);
    put((int32_t)st.groundIters);
    pose(out.T_Map_Curr);
    pose(out.T_Delta);
    put((int32_t)out.tm.size());
    for (size_t i = 0; i < out.tm.size(); ++i) {
      put((int32_t)out.tm[i].id);
      cyl_model(out.tm[i].model);
      put((int32_t)(i < out.matches.size() ? out.matches[i] : -1));
    }
    put((int32_t)after.size());
    for (const Plane& p : after) plane_model(p.model);
  }
};

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::string> pos;
  std::string model_name = "lines", weights, traj_out, shadow_name, dump_path;
  double max_tree_radius = -1.0, tree_match_thresh = -1.0;
  // --shadow-seed: the shadow's hypothesis key (another stream of the same
  // tree model measures that model's spread against itself)
  unsigned long long shadow_seed = 0;
  int pose_precision = 1;
  bool log = false, pinned = false, pipelined = false, device_est = false, compare_device = false;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "--log") {
      log = true;
    } else if (a == "--device-estimator") {
      device_est = true;
    } else if (a == "--compare-device") {
      compare_device = true;
    } else if (a == "--pinned") {
      pinned = true;
    } else if (a == "--pipelined") {
      pipelined = true;
    } else if ((a == "--max-tree-radius" || a == "--tree-match-thresh") && i + 1 < argc) {
      (a == "--max-tree-radius" ? max_tree_radius : tree_match_thresh) = std::atof(argv[++i]);
    } else if (a == "--dump" && i + 1 < argc) {
      dump_path = argv[++i];
    } else if (a == "--pose-precision" && i + 1 < argc) {
      pose_precision = std::atoi(argv[++i]);
    } else if (a == "--shadow-seed" && i + 1 < argc) {
      shadow_seed = std::strtoull(argv[++i], nullptr, 0);
    } else if ((a == "--tree-model" || a == "--weights" || a == "--traj-out" ||
                a == "--shadow") &&
               i + 1 < argc) {
      (a == "--tree-model" ? model_name
       : a == "--weights"  ? weights
       : a == "--shadow"   ? shadow_name
                           : traj_out) = argv[++i];
    } else if (a.rfind("--", 0) == 0) {
      std::fprintf(stderr, "unknown option %s\n", a.c_str());
      return 2;
    } else {
      pos.push_back(a);
    }
  }
  if (pos.empty()) {
    std::fprintf(stderr,
                 "usage: %s <sequence> [passes] [use_graph] [--tree-model lines|plain|learned] "
                 "[--weights npy] [--traj-out file] [--log] [--pinned] [--pipelined] "
                 "[--shadow lines|plain|learned] [--dump file] [--max-tree-radius m] "
                 "[--tree-match-thresh m] [--pose-precision 0|1]\n",
                 argv[0]);
    return 2;
  }
  TreeModel tree_model = TreeModel::MedianLines, shadow_model = TreeModel::MedianLines;
  if (!parse_tree_model(model_name, &tree_model)) {
    std::fprintf(stderr, "unknown tree model %s\n", model_name.c_str());
    return 2;
  }
  if (compare_device) {
    // the shadow: the same tree model with the estimator on the device, from
    // its own state (which evolves as the host estimator's does)
    if (!shadow_name.empty() || device_est) {
      std::fprintf(stderr, "--compare-device runs its own shadow on the host estimator's run\n");
      return 2;
    }
    shadow_name = model_name;
  }
  const bool use_shadow = !shadow_name.empty();
  if (use_shadow && !parse_tree_model(shadow_name, &shadow_model)) {
    std::fprintf(stderr, "unknown shadow tree model %s\n", shadow_name.c_str());
    return 2;
  }
  if (use_shadow && pipelined) {
    std::fprintf(stderr, "--shadow runs keyframes one at a time: not with --pipelined\n");
    return 2;
  }
  const int passes = pos.size() > 1 ? std::atoi(pos[1].c_str()) : 3;
  const bool use_graph = pos.size() > 2 ? std::atoi(pos[2].c_str()) != 0 : true;
  const char* seq_path = pos[0].c_str();
  FILE* f = std::fopen(seq_path, "rb");
  if (!f) {
    std::fprintf(stderr, "cannot open %s\n", seq_path);
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
  if (max_tree_radius > 0) fm.maxTreeRadius = max_tree_radius;
  if (tree_match_thresh > 0) fm.treeMatchThresh = tree_match_thresh;
  DumpWriter dump;
  if (!dump_path.empty()) {
    dump.f = std::fopen(dump_path.c_str(), "wb");
    if (!dump.f) {
      std::fprintf(stderr, "cannot open %s\n", dump_path.c_str());
      return 2;
    }
  }
  std::vector<double> total, fe, up, down, est, build, ph_models, ph_match, ph_pose, ph_finish;
  // per sample: total, keyframe, upload, front end, estimator, models, pose
  std::vector<std::array<double, 7>> slow;
  std::vector<double> first;  // the warm-up pass
  double trees = 0, planes = 0;
  int samples = 0, ok_count = 0;
  double final_drift = 0, max_drift = 0, max_rot = 0;
  std::vector<double> fit, traj;
  // per keyframe of the timed passes: the map's submap query and update
  // (MapManager::getSubmap + updateMap, as SLOAMNode::run makes them around
  // each keyframe; counted in the throughput, not in the latency)
  std::vector<double> map_ms;
  // the keyframe's latency with its map query and update added
  std::vector<double> with_map;
  long long fitted = 0, fallback = 0;
  int discarded = 0;
  std::vector<double> pass_rate;  // keyframes per second of each timed pass
  // --shadow: per keyframe, the second runner's pose against the first's
  std::vector<double> sh_m, sh_deg;
  int sh_ok_differs = 0;
  // --compare-device: keyframes compared, where the two estimators' success
  // differs, and where the pose (bits), the landmarks (models, ids, matches)
  // and the step decisions agree; the first keyframe that differs
  int dc_calls = 0, dc_ok_differs = 0, dc_pose_eq = 0, dc_models_eq = 0, dc_step_eq = 0,
      dc_first_diff = -1;
  double dc_max_m = 0, dc_max_deg = 0;
  int dc_printed = 0;  // differing keyframes printed to stderr
  // per compared keyframe: m, deg, k, then each runner's trees, map
  // matches, tree step run, tree feature matches; largest radius difference
  // of the landmarks both kept in the same order
  std::vector<std::array<double, 12>> sh_rows;
  // radius differences of the trees both runners kept (paired by tree id)
  std::vector<double> sh_dr;
  // trees kept by one runner and rejected by the other, per rejecting rule
  // (Cylinder::RejectReason bits 0..6; [7]: not among the other's rejects)
  std::array<int, 8> flips_first_rejects{}, flips_shadow_rejects{};
  // first pass: RunSloam calls after the first, and how many of them ran the
  // XY-yaw (tree) and Z-roll-pitch (ground) steps
  int step_calls = 0, tree_steps = 0, ground_steps = 0;
  // odometry alone (the increments composed from the first true pose), against
  // the true poses: final and largest translation error, largest rotation
  double odom_final = 0, odom_max = 0, odom_max_deg = 0;
  {
    SE3 o = to_se3(seq[0].truth);
    for (int k = 1; k < K; ++k) {
      o = o * to_se3(seq[k].odom);
      double m, deg;
      pose_diff(o, to_se3(seq[k].truth), &m, &deg);
      odom_final = m;
      odom_max = std::max(odom_max, m);
      odom_max_deg = std::max(odom_max_deg, deg);
    }
  }
  try {
    std::unique_ptr<KeyframeRunner> shadow;
    if (use_shadow) {
      shadow.reset(new KeyframeRunner(H, W, fm, use_graph, 22.5f, -22.5f, false));
      shadow->setTreeModel(shadow_model, weights);
      if (shadow_seed) shadow->setHypothesisSeed(shadow_seed);
      shadow->setPosePrecision(pose_precision);
      if (compare_device) shadow->setDeviceEstimator(true);
    }
    for (int pass = 0; pass < passes; ++pass) {
      KeyframeRunner kf(H, W, fm, use_graph, 22.5f, -22.5f, pipelined);
      kf.setTreeModel(tree_model, weights);
      kf.setPosePrecision(pose_precision);
      if (device_est) kf.setDeviceEstimator(true);
      int dump_k = 0, dump_calls = 0;
      if (dump.f && pass == 0)
        kf.setInputRecorder([&](const SloamInput& in, const MapView& m) {
          dump.input(dump_k, dump_calls == 0, in, m, kf.estimator().getPrevGroundModel(),
                     fm.numGroundFeatures);
          ++dump_calls;
        });
      double staging_ms = 0.0;  // copies into the pinned buffers (not counted)
      // hands scan j to the runner: into its pinned buffers (--pinned), then
      // submitted (--pipelined) or returned for run()
      auto stage = [&](int j, const float** xp, const uint8_t** lp) {
        *xp = seq[j].xyzi.data();
        *lp = seq[j].label.data();
        if (pinned) {
          const auto c0 = std::chrono::steady_clock::now();
          std::memcpy(kf.pinnedXyzi(), *xp, sizeof(float) * seq[j].xyzi.size());
          std::memcpy(kf.pinnedLabels(), *lp, seq[j].label.size());
          *xp = kf.pinnedXyzi();
          *lp = kf.pinnedLabels();
          staging_ms += std::chrono::duration<double, std::milli>(
                            std::chrono::steady_clock::now() - c0)
                            .count();
        }
        if (pipelined) kf.submit(*xp, *lp);
      };
      const float* xyzi = nullptr;
      const uint8_t* label = nullptr;
      // every keyframe of the pass, the first (map start) included
      const auto tpass = std::chrono::steady_clock::now();
      if (pipelined) stage(0, &xyzi, &label);
      SE3 estimate = to_se3(seq[0].truth);
      MapManager semanticMap;
      bool nodeFirst = true;  // SLOAMNode's firstScan_
      for (int k = 0; k < K; ++k) {
        if (k > 0) estimate = estimate * to_se3(seq[k].odom);
        SloamOutput out;
        KeyframeTiming t;
        const SE3 before = estimate;
        MapView map;  // the submap by reference (valid until updateMap)
        const auto m0 = std::chrono::steady_clock::now();
        semanticMap.getSubmap(estimate, map);
        double kf_map_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - m0)
                .count();
        const bool discard = !nodeFirst && map.empty();
        if (pipelined) {
          // the next scan's front end runs while this keyframe's estimator does
          if (k + 1 < K) stage(k + 1, &xyzi, &label);
        } else {
          stage(k, &xyzi, &label);
        }
        bool ok = false;
        if (discard) {
          if (pass == 0) ++discarded;
          if (pipelined) kf.skip();
        } else {
          // the shadow starts from the first runner's state: its estimate,
          // submap and previous ground models
          std::vector<Plane> prevPlanes;
          if (shadow && pass == 0) prevPlanes = kf.estimator().getPrevGroundModel();
          dump_k = k;
          ok = pipelined ? kf.finish(estimate, map, out, &t)
                         : kf.run(xyzi, label, estimate, map, out, &t);
          if (dump.f && pass == 0)
            dump.output(ok, kf.lastStep(), out, kf.estimator().getPrevGroundModel());
          if (pass == 0 && !nodeFirst) {
            const SloamStepInfo st = kf.lastStep();
            ++step_calls;
            tree_steps += st.treeCheck ? 1 : 0;
            ground_steps += st.groundCheck ? 1 : 0;
          }
          if (shadow && pass == 0) {
            if (!compare_device) shadow->estimator().setPrevGroundModel(std::move(prevPlanes));
            SloamOutput so;
            const bool sok =
                shadow->run(seq[k].xyzi.data(), seq[k].label.data(), before, map, so, nullptr);
            if (compare_device) {
              // bit-level comparison of the two estimators' outputs
              ++dc_calls;
              if (ok != sok) ++dc_ok_differs;
              if (ok && sok) {
                const bool pose_eq = same_se3(out.T_Map_Curr, so.T_Map_Curr);
                bool models_eq = out.tm.size() == so.tm.size() && out.matches == so.matches;
                for (size_t i = 0; models_eq && i < out.tm.size(); ++i)
                  models_eq = same_model(out.tm[i].getModel(), so.tm[i].getModel()) &&
                              out.tm[i].id == so.tm[i].id;
                const SloamStepInfo a = kf.lastStep(), b = shadow->lastStep();
                const bool step_eq = a.treeCheck == b.treeCheck && a.groundCheck == b.groundCheck &&
                                     a.treeMatches == b.treeMatches &&
                                     a.planeMatches == b.planeMatches &&
                                     a.treeIters == b.treeIters &&
                                     a.groundIters == b.groundIters;
                dc_pose_eq += pose_eq ? 1 : 0;
                dc_models_eq += models_eq ? 1 : 0;
                dc_step_eq += step_eq ? 1 : 0;
                double m, deg;
                pose_diff(out.T_Map_Curr, so.T_Map_Curr, &m, &deg);
                dc_max_m = std::max(dc_max_m, m);
                dc_max_deg = std::max(dc_max_deg, deg);
                if ((!pose_eq || !models_eq || !step_eq) && dc_first_diff < 0) dc_first_diff = k;
                if ((!pose_eq || !models_eq || !step_eq) && dc_printed < 10) {
                  ++dc_printed;
                  std::fprintf(stderr,
                               "device differs at keyframe %d: pose %d (%.3e m %.3e deg), "
                               "landmarks %d (%zu / %zu), steps %d (tree %d/%d ground %d/%d "
                               "matches %d/%d %d/%d iterations %d/%d %d/%d)\n",
                               k, pose_eq ? 1 : 0, m, deg, models_eq ? 1 : 0, out.tm.size(),
                               so.tm.size(), step_eq ? 1 : 0, a.treeCheck, b.treeCheck,
                               a.groundCheck, b.groundCheck, a.treeMatches, b.treeMatches,
                               a.planeMatches, b.planeMatches, a.treeIters, b.treeIters,
                               a.groundIters, b.groundIters);
                }
              } else if (ok != sok && dc_printed < 10) {
                ++dc_printed;
                std::fprintf(stderr, "device differs at keyframe %d: success %d / %d\n", k,
                             ok ? 1 : 0, sok ? 1 : 0);
              }
            }
            if (compare_device) {
            } else if (k > 0 && ok && sok) {
              double m, deg;
              pose_diff(out.T_Map_Curr, so.T_Map_Curr, &m, &deg);
              sh_m.push_back(m);
              sh_deg.push_back(deg);
              auto matched = [](const SloamOutput& o) {
                int c = 0;
                for (int x : o.matches) c += x >= 0 ? 1 : 0;
                return (double)c;
              };
              const SloamStepInfo a = kf.lastStep();
              const SloamStepInfo b = shadow->lastStep();
              // trees one runner keeps and the other rejects, by the rules
              // that reject them
              auto tally = [](const SloamOutput& keeper, const SloamOutput& other,
                              const std::vector<std::pair<int, int>>& other_rejected,
                              std::array<int, 8>& counts) {
                for (const Cylinder& c : keeper.tm) {
                  bool kept = false;
                  for (const Cylinder& o : other.tm) kept = kept || o.id == c.id;
                  if (kept) continue;
                  int reason = -1;
                  for (const auto& r : other_rejected)
                    if ((size_t)r.first == c.id) reason = r.second;
                  if (reason < 0) {
                    ++counts[7];
                    continue;
                  }
                  for (int bit = 0; bit < 7; ++bit) counts[bit] += (reason >> bit) & 1;
                }
              };
              tally(out, so, shadow->estimator().lastRejectedTrees(), flips_shadow_rejects);
              tally(so, out, kf.estimator().lastRejectedTrees(), flips_first_rejects);
              for (const Cylinder& a_c : out.tm)
                for (const Cylinder& b_c : so.tm)
                  if (a_c.id == b_c.id) {
                    sh_dr.push_back(std::fabs(a_c.getModel().radius - b_c.getModel().radius));
                    break;
                  }
              double dr = -1.0;
              if (out.tm.size() == so.tm.size()) {
                dr = 0.0;
                for (size_t i = 0; i < out.tm.size(); ++i)
                  dr = std::max(dr, std::fabs(out.tm[i].getModel().radius -
                                              so.tm[i].getModel().radius));
              }
              sh_rows.push_back({m, deg, (double)k, (double)out.tm.size(), matched(out),
                                 a.treeCheck ? 1.0 : 0.0, (double)a.treeMatches,
                                 (double)so.tm.size(), matched(so), b.treeCheck ? 1.0 : 0.0,
                                 (double)b.treeMatches, dr});
            } else if (k > 0 && ok != sok) {
              ++sh_ok_differs;
            }
          }
          const auto m1 = std::chrono::steady_clock::now();
          semanticMap.updateMap(out.tm, out.matches);
          kf_map_ms += std::chrono::duration<double, std::milli>(
                           std::chrono::steady_clock::now() - m1)
                           .count();
          if (ok && nodeFirst && !out.tm.empty()) nodeFirst = false;
        }
        if (log && pass == 0) {
          const SE3 truth = to_se3(seq[k].truth);
          auto err = [&](const SE3& e, double* deg) {
            const Quat dq = e.unit_quaternion().conjugate() * truth.unit_quaternion();
            *deg = 2.0 * std::acos(std::min(1.0, std::fabs(dq.w))) * 180.0 / M_PI;
            return (e.translation() - truth.translation()).norm();
          };
          double d0, d1;
          const double e0 = err(before, &d0);
          const double e1 = err(ok ? out.T_Map_Curr : before, &d1);
          int matched = 0;
          for (int m : out.matches) matched += m >= 0 ? 1 : 0;
          std::fprintf(stderr,
                       "keyframe %d: ok %d, trees %d (fitted %d, fallback %d), map %zu, "
                       "matched %d, planes %d, error before %.4f m %.3f deg, after %.4f m "
                       "%.3f deg\n",
                       k, ok ? 1 : 0, t.trees, t.fitted, t.fallback, map.size(), matched,
                       t.planes, e0, d0, e1, d1);
        }
        if (ok) estimate = out.T_Map_Curr;
        if (pass == 0) {
          const Vector3 tr = estimate.translation();
          const Quat q = estimate.unit_quaternion();
          const double row[7] = {tr[0], tr[1], tr[2], q.x, q.y, q.z, q.w};
          traj.insert(traj.end(), row, row + 7);
        }
        if (pass == 0 && k > 0) {
          const SE3 truth = to_se3(seq[k].truth);
          const double d = (estimate.translation() - truth.translation()).norm();
          const Quat dq = estimate.unit_quaternion().conjugate() * truth.unit_quaternion();
          const double ang = 2.0 * std::acos(std::min(1.0, std::fabs(dq.w))) * 180.0 / M_PI;
          final_drift = d;
          max_drift = std::max(max_drift, d);
          max_rot = std::max(max_rot, ang);
        }
        // first scan: map initialisation (and graph capture); discarded
        // keyframes ran nothing
        if (k == 0 || discard) continue;
        if (pass == 0 && passes > 1) {
          // warm-up pass (first use of every buffer, plan and kernel of the
          // process): reported apart
          first.push_back(t.total_ms);
          continue;
        }
        total.push_back(t.total_ms);
        map_ms.push_back(kf_map_ms);
        with_map.push_back(t.total_ms + kf_map_ms);
        slow.push_back({t.total_ms, (double)k, t.upload_ms, t.frontend_ms, t.estimator_ms,
                        t.phases.models, t.phases.pose});
        fe.push_back(t.frontend_ms);
        up.push_back(t.upload_ms);
        down.push_back(t.download_ms);
        est.push_back(t.estimator_ms);
        build.push_back(t.build_ms);
        fit.push_back(t.fit_ms);
        fitted += t.fitted;
        fallback += t.fallback;
        ph_models.push_back(t.phases.models);
        ph_match.push_back(t.phases.match);
        ph_pose.push_back(t.phases.pose);
        ph_finish.push_back(t.phases.finish);
        trees += t.trees;
        planes += t.planes;
        ++samples;
        ok_count += ok ? 1 : 0;
      }
      if (pass > 0 || passes == 1) {
        const double ms = std::chrono::duration<double, std::milli>(
                              std::chrono::steady_clock::now() - tpass)
                              .count() -
                          staging_ms;
        if (ms > 0) pass_rate.push_back(K / (ms / 1e3));
      }
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "keyframe_bench: %s\n", e.what());
    return 1;
  }
  if (dump.f && std::fclose(dump.f) != 0) {
    std::fprintf(stderr, "cannot write %s\n", dump_path.c_str());
    return 1;
  }
  if (!traj_out.empty()) {
    FILE* o = std::fopen(traj_out.c_str(), "wb");
    if (!o || std::fwrite(traj.data(), 8, traj.size(), o) != traj.size()) {
      std::fprintf(stderr, "cannot write %s\n", traj_out.c_str());
      if (o) std::fclose(o);
      return 1;
    }
    std::fclose(o);
  }
  std::printf(
      "{\"tree_model\": \"%s\", \"pinned_input\": %s, \"tree_fit_p50_ms\": %.4f, "
      "\"trees_fitted\": %lld, \"trees_fallback\": %lld, \"discarded\": %d, "
      "\"pipelined\": %s, \"throughput_kf_per_s\": %.1f, \"device_estimator\": %s, ",
      model_name.c_str(), pinned ? "true" : "false", pct(fit, 0.5), fitted, fallback, discarded,
      pipelined ? "true" : "false", pct(pass_rate, 0.5), device_est ? "true" : "false");
  if (compare_device)
    std::printf(
        "\"device_compare\": {\"calls\": %d, \"success_differs\": %d, \"pose_bit_equal\": %d, "
        "\"landmarks_equal\": %d, \"steps_equal\": %d, \"first_difference\": %d, "
        "\"max_m\": %.3e, \"max_deg\": %.3e}, ",
        dc_calls, dc_ok_differs, dc_pose_eq, dc_models_eq, dc_step_eq, dc_first_diff, dc_max_m,
        dc_max_deg);
  std::printf(
      "\"params\": {\"max_tree_radius\": %.4f, \"tree_match_thresh\": %.4f, "
      "\"max_axis_theta\": %.2f, \"two_step\": %s}, \"steps\": {\"calls\": %d, "
      "\"tree_step_frac\": %.4f, \"ground_step_frac\": %.4f}, \"odometry_drift\": "
      "{\"final_m\": %.4f, \"max_m\": %.4f, \"max_deg\": %.4f}, ",
      (double)fm.maxTreeRadius, (double)fm.treeMatchThresh, (double)fm.maxAxisTheta,
      fm.twoStepOptim ? "true" : "false", step_calls,
      step_calls ? (double)tree_steps / step_calls : 0.0,
      step_calls ? (double)ground_steps / step_calls : 0.0, odom_final, odom_max, odom_max_deg);
  std::printf(
      "\"keyframes\": %d, \"passes\": %d, \"height\": %d, \"width\": %d, \"graph\": %s, "
      "\"latency_ms\": {\"p50\": %.4f, \"p99\": %.4f, \"mean\": %.4f, \"max\": %.4f}, "
      "\"stage_p50_ms\": {\"upload\": %.4f, \"frontend\": %.4f, \"download\": %.4f, "
      "\"estimator\": %.4f}, \"estimator_p50_ms\": {\"build\": %.4f, \"models\": %.4f, "
      "\"match\": %.4f, \"pose\": %.4f, \"finish\": %.4f}, "
      "\"frontend_p99_ms\": %.4f, \"map_p50_ms\": %.4f, \"latency_with_map_ms\": "
      "{\"p50\": %.4f, \"p99\": %.4f}, \"trees_mean\": %.2f, "
      "\"planes_mean\": %.2f, \"success_rate\": %.4f, \"drift\": {\"final_m\": %.4f, "
      "\"max_m\": %.4f, \"max_deg\": %.4f}, \"warmup_pass_latency_ms\": {\"p50\": %.4f, "
      "\"max\": %.4f}, ",
      K, passes, H, W, use_graph ? "true" : "false", pct(total, 0.5), pct(total, 0.99),
      samples ? [&] { double s = 0; for (double x : total) s += x; return s / samples; }() : 0.0,
      total.empty() ? 0.0 : *std::max_element(total.begin(), total.end()), pct(up, 0.5),
      pct(fe, 0.5), pct(down, 0.5), pct(est, 0.5), pct(build, 0.5), pct(ph_models, 0.5),
      pct(ph_match, 0.5), pct(ph_pose, 0.5), pct(ph_finish, 0.5), pct(fe, 0.99), pct(map_ms, 0.5),
      pct(with_map, 0.5), pct(with_map, 0.99),
      samples ? trees / samples : 0.0, samples ? planes / samples : 0.0,
      samples ? (double)ok_count / samples : 0.0, final_drift, max_drift, max_rot,
      pct(first, 0.5), first.empty() ? 0.0 : *std::max_element(first.begin(), first.end()));
  if (use_shadow && !compare_device) {
    // all compared keyframes, then those where both runners kept the same
    // number of trees and took the same XY-yaw decision
    std::vector<double> same_m, same_deg;
    for (const auto& r : sh_rows)
      if (r[3] == r[7] && r[5] == r[9]) {
        same_m.push_back(r[0]);
        same_deg.push_back(r[1]);
      }
    auto stats = [&](const std::vector<double>& m, const std::vector<double>& deg) {
      int within = 0;
      for (size_t i = 0; i < m.size(); ++i) within += m[i] <= 0.01 && deg[i] <= 0.1;
      std::printf(
          "\"compared\": %zu, \"m\": {\"p50\": %.6f, \"p99\": %.6f, \"max\": %.6f}, "
          "\"deg\": {\"p50\": %.6f, \"p99\": %.6f, \"max\": %.6f}, "
          "\"within_1cm_0p1deg\": %.4f",
          m.size(), pct(m, 0.5), pct(m, 0.99), m.empty() ? 0.0 : *std::max_element(m.begin(), m.end()),
          pct(deg, 0.5), pct(deg, 0.99),
          deg.empty() ? 0.0 : *std::max_element(deg.begin(), deg.end()),
          m.empty() ? 0.0 : (double)within / m.size());
    };
    std::printf("\"shadow\": {\"tree_model\": \"%s\", \"success_differs\": %d, ",
                shadow_name.c_str(), sh_ok_differs);
    stats(sh_m, sh_deg);
    std::printf(", \"same_decisions\": {");
    stats(same_m, same_deg);
    int over = 0;
    for (double x : sh_dr) over += x > 0.01 ? 1 : 0;
    static const char* kRules[8] = {"out_of_range",     "no_model",   "root_off_ground",
                                    "root_at_origin",   "radius_not_positive",
                                    "radius_cap",       "axis_tilt",  "unexplained"};
    auto flips = [&](const std::array<int, 8>& c) {
      std::printf("{");
      for (int i = 0; i < 8; ++i) std::printf("%s\"%s\": %d", i ? ", " : "", kRules[i], c[i]);
      std::printf("}");
    };
    std::printf("}, \"validity_flips\": {\"rejected_by_first\": ");
    flips(flips_first_rejects);
    std::printf(", \"rejected_by_shadow\": ");
    flips(flips_shadow_rejects);
    std::printf(
        "}, \"tree_radius_diff_m\": {\"trees\": %zu, \"p50\": %.6f, \"p99\": %.6f, "
        "\"max\": %.6f, \"frac_over_1cm\": %.4f}}}\n",
        sh_dr.size(), pct(sh_dr, 0.5), pct(sh_dr, 0.99),
        sh_dr.empty() ? 0.0 : *std::max_element(sh_dr.begin(), sh_dr.end()),
        sh_dr.empty() ? 0.0 : (double)over / sh_dr.size());
  } else {
    std::printf("\"shadow\": null}\n");
  }
  // --shadow: the five keyframes whose poses differ most, on stderr
  std::sort(sh_rows.begin(), sh_rows.end(),
            [](const auto& a, const auto& b) { return a[0] > b[0]; });
  for (size_t i = 0; i < sh_rows.size() && i < 5; ++i) {
    const auto& r = sh_rows[i];
    std::fprintf(stderr,
                 "shadow keyframe %d: %.4f m %.3f deg; %s trees %d matched %d tree step %d "
                 "(%d feature matches); %s trees %d matched %d tree step %d (%d); largest "
                 "radius difference %.4f m\n",
                 (int)r[2], r[0], r[1], model_name.c_str(), (int)r[3], (int)r[4], (int)r[5],
                 (int)r[6], shadow_name.c_str(), (int)r[7], (int)r[8], (int)r[9], (int)r[10],
                 r[11]);
  }
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
