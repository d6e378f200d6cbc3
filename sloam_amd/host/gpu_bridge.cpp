// Host-side bridge from the drop-in classes to the C ABI.
#include "gpu_bridge.h"

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

namespace sloam_gpu {
namespace {

void fail(const char* what, int st, sloam_ctx* ctx) {
  std::string msg = std::string(what) + ": " + sloam_strerror(st);
  if (ctx) {
    const char* d = sloam_ctx_last_error(ctx);
    if (d && *d) msg += std::string(" (") + d + ")";
  }
  throw std::runtime_error(msg);
}

void check(int st, const char* what) {
  if (st != SLOAM_OK) fail(what, st, context());
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

struct CtxHolder {
  sloam_ctx* ctx = nullptr;
  ~CtxHolder() {
    if (ctx) sloam_ctx_destroy(ctx);
  }
};

// Per-thread grow-only device buffers.  All work of the bridge runs on the
// context's one stream and every call ends with a stream synchronisation, so
// a call reuses the buffers of the previous one: the k-th buffer a call asks
// for is slot k.  A slot that must grow is freed after the stream drains.
struct Arena {
  std::vector<std::pair<void*, size_t>> slots;
  size_t next = 0;
  ~Arena() {
    for (auto& s : slots)
      if (s.first) (void)hipFree(s.first);
  }
  void* take(size_t bytes) {
    if (next == slots.size()) slots.emplace_back(nullptr, 0);
    auto& s = slots[next++];
    if (s.second < bytes) {
      hip_check(hipStreamSynchronize((hipStream_t)sloam_ctx_stream(context())),
                "hipStreamSynchronize");
      if (s.first) (void)hipFree(s.first);
      s.first = nullptr;
      s.second = 0;
      const size_t sz = bytes + bytes / 2;
      hip_check(hipMalloc(&s.first, sz), "hipMalloc");
      s.second = sz;
    }
    return s.first;
  }
};

Arena& arena() {
  thread_local Arena a;
  return a;
}

// Per-thread grow-only pinned host buffer that stages a call's uploads and
// downloads (one copy each way); every call synchronises before it returns,
// so the next call may rewrite it.
struct Pinned {
  void* p = nullptr;
  size_t bytes = 0;
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
  char* take(size_t n) {
    if (n > bytes) {
      if (p) hip_check(hipHostFree(p), "hipHostFree");
      p = nullptr;
      bytes = 0;
      const size_t sz = n + n / 2;
      hip_check(hipHostMalloc(&p, sz), "hipHostMalloc");
      bytes = sz;
    }
    return static_cast<char*>(p);
  }
};

Pinned& pinned() {
  thread_local Pinned s;
  return s;
}

// byte offsets of 256-byte aligned sections of one staging block
struct Sections {
  size_t size = 0;
  size_t add(size_t bytes) {
    const size_t o = size;
    size += (bytes + 255) / 256 * 256;
    return o;
  }
};

template <typename T>
void put(char* base, size_t off, const std::vector<T>& v) {
  if (!v.empty()) std::memcpy(base + off, v.data(), sizeof(T) * v.size());
}

// opened at the start of every bridge call: slots are numbered from 0 again
struct CallScope {
  CallScope() { arena().next = 0; }
};

// a device buffer of at least `bytes` for the rest of the call
struct Dev {
  void* p = nullptr;
  explicit Dev(size_t bytes) { p = arena().take(bytes ? bytes : 16); }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

template <typename T>
void up(Dev& d, const std::vector<T>& v) {
  if (!v.empty())
    hip_check(hipMemcpyAsync(d.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice,
                             (hipStream_t)sloam_ctx_stream(context())),
              "hipMemcpyAsync(H2D)");
}

template <typename T>
void down(std::vector<T>& v, const Dev& d) {
  if (!v.empty())
    hip_check(hipMemcpyAsync(v.data(), d.p, sizeof(T) * v.size(), hipMemcpyDeviceToHost,
                             (hipStream_t)sloam_ctx_stream(context())),
              "hipMemcpyAsync(D2H)");
}

}  // namespace

int deviceIndex() {
  int dev = 0;
  if (const char* e = std::getenv("SLOAM_DEVICE")) dev = std::atoi(e);
  return dev;
}

sloam_ctx* context() {
  thread_local CtxHolder holder;
  if (!holder.ctx) {
    const int dev = deviceIndex();
    const int st = sloam_ctx_create(dev, &holder.ctx);
    if (st != SLOAM_OK) fail("sloam_ctx_create", st, nullptr);
  }
  return holder.ctx;
}

std::vector<CellFit> groundPlanes(const VectorType& points, const sloam_ground_cfg& cfg) {
  CallScope scope;
  sloam_ctx* ctx = context();
  const int n_cells = cfg.radii_bins * cfg.theta_bins;
  const int nf = cfg.num_features;
  std::vector<float> pts(4 * points.size());
  for (size_t i = 0; i < points.size(); ++i) {
    pts[4 * i] = points[i].x;
    pts[4 * i + 1] = points[i].y;
    pts[4 * i + 2] = points[i].z;
    pts[4 * i + 3] = points[i].intensity;
  }
  Dev d_pts(sizeof(float) * pts.size());
  Dev d_info(sizeof(int32_t) * 4 * n_cells);
  Dev d_model(sizeof(double) * 8 * n_cells);
  Dev d_feat(sizeof(float) * 4 * (size_t)n_cells * (size_t)(nf > 0 ? nf : 1));
  up(d_pts, pts);
  Dev d_ko(sizeof(int32_t) * (n_cells + 1));
  Dev d_ki(sizeof(int32_t) * (points.empty() ? 1 : points.size()));
  check(sloam_ground_planes(ctx, d_pts.as<float>(), (int32_t)points.size(), nullptr, &cfg,
                            d_info.as<int32_t>(), d_model.as<double>(), d_feat.as<float>(),
                            d_ko.as<int32_t>(), d_ki.as<int32_t>()),
        "sloam_ground_planes");
  std::vector<int32_t> ko(n_cells + 1), ki(points.size());
  down(ko, d_ko);
  down(ki, d_ki);
  std::vector<int32_t> info(4 * n_cells);
  std::vector<double> model(8 * n_cells);
  std::vector<float> feat(4 * (size_t)n_cells * (size_t)(nf > 0 ? nf : 1));
  down(info, d_info);
  down(model, d_model);
  down(feat, d_feat);
  check(sloam_sync(ctx), "sloam_sync");
  std::vector<CellFit> out(n_cells);
  for (int c = 0; c < n_cells; ++c) {
    CellFit& f = out[c];
    f.valid = info[4 * c] != 0;
    f.kept = info[4 * c + 1];
    for (int i = 0; i < 4; ++i) f.plane[i] = model[8 * c + i];
    for (int i = 0; i < 3; ++i) f.centroid[i] = model[8 * c + 4 + i];
    if (f.valid)
      for (int j = 0; j < nf; ++j) {
        const float* q = &feat[4 * ((size_t)c * nf + j)];
        f.features.push_back(PointT{q[0], q[1], q[2], q[3]});
      }
    if (!points.empty())
      for (int j = ko[c]; j < ko[c + 1] && j - ko[c] < f.kept; ++j) f.kept_index.push_back(ki[j]);
  }
  return out;
}

std::vector<LineResult> fitLines(const std::vector<VectorType>& sets,
                                 const sloam_line_fit_cfg& cfg) {
  CallScope scope;
  std::vector<LineResult> out(sets.size());
  if (sets.empty()) return out;
  sloam_ctx* ctx = context();
  std::vector<float> xyz;
  std::vector<int32_t> off(1, 0);
  for (const auto& s : sets) {
    for (const auto& p : s) {
      xyz.push_back(p.x);
      xyz.push_back(p.y);
      xyz.push_back(p.z);
    }
    off.push_back((int32_t)(xyz.size() / 3));
  }
  const int S = (int)sets.size();
  Dev d_xyz(sizeof(float) * (xyz.empty() ? 3 : xyz.size()));
  Dev d_off(sizeof(int32_t) * off.size());
  Dev d_coeff(sizeof(float) * 6 * S);
  Dev d_stats(sizeof(int32_t) * 4 * S);
  up(d_xyz, xyz);
  up(d_off, off);
  check(sloam_fit_lines(ctx, d_xyz.as<float>(), d_off.as<int32_t>(), S, &cfg, d_coeff.as<float>(),
                        d_stats.as<int32_t>(), nullptr),
        "sloam_fit_lines");
  std::vector<float> coeff(6 * S);
  std::vector<int32_t> stats(4 * S);
  down(coeff, d_coeff);
  down(stats, d_stats);
  check(sloam_sync(ctx), "sloam_sync");
  for (int s = 0; s < S; ++s) {
    out[s].ok = stats[4 * s] != 0;
    out[s].inliers = stats[4 * s + 1];
    for (int i = 0; i < 6; ++i) out[s].coeff[i] = coeff[6 * s + i];
  }
  return out;
}

std::vector<PoseResult> solvePoses(const std::vector<PoseProblem>& problems,
                                   const sloam_pose_cfg& cfg) {
  CallScope scope;
  std::vector<PoseResult> out(problems.size());
  if (problems.empty()) return out;
  sloam_ctx* ctx = context();
  std::vector<float> pts;
  std::vector<double> cyl, pl, pose;
  std::vector<int32_t> res_off(1, 0), cyl_off(1, 0), pl_off(1, 0), masks;
  for (const auto& pr : problems) {
    // landmark ids are global: shift this problem's local ids
    const int32_t cb = cyl_off.back(), pb = pl_off.back();
    for (size_t i = 0; i < pr.points.size(); i += 4) {
      pts.push_back(pr.points[i]);
      pts.push_back(pr.points[i + 1]);
      pts.push_back(pr.points[i + 2]);
      int32_t id;
      std::memcpy(&id, &pr.points[i + 3], 4);
      id = id >= 0 ? id + cb : id - pb;
      float f;
      std::memcpy(&f, &id, 4);
      pts.push_back(f);
    }
    cyl.insert(cyl.end(), pr.cyl.begin(), pr.cyl.end());
    pl.insert(pl.end(), pr.planes.begin(), pr.planes.end());
    res_off.push_back((int32_t)(pts.size() / 4));
    cyl_off.push_back((int32_t)(cyl.size() / 8));
    pl_off.push_back((int32_t)(pl.size() / 4));
    pose.insert(pose.end(), pr.pose, pr.pose + 7);
    masks.push_back(pr.free_mask);
  }
  const int P = (int)problems.size();
  // one plan per thread, re-targeted at each problem set (grow-only buffers)
  thread_local std::unique_ptr<sloam_pose_plan, void (*)(sloam_pose_plan*)> cached(
      nullptr, sloam_pose_plan_destroy);
  if (!cached) {
    sloam_pose_plan* p = nullptr;
    check(sloam_pose_plan_create(ctx, res_off.data(), P, cyl_off.back(), pl_off.back(), &cfg, &p),
          "sloam_pose_plan_create");
    cached.reset(p);
  } else {
    check(sloam_pose_plan_reset(cached.get(), res_off.data(), P, cyl_off.back(), pl_off.back(),
                                &cfg),
          "sloam_pose_plan_reset");
  }
  sloam_pose_plan* plan = cached.get();
  // inputs in one upload, outputs in one download, through pinned staging
  Sections in, res;
  const size_t o_pts = in.add(sizeof(float) * (pts.empty() ? 4 : pts.size()));
  const size_t o_cyl = in.add(sizeof(double) * (cyl.empty() ? 8 : cyl.size()));
  const size_t o_pl = in.add(sizeof(double) * (pl.empty() ? 4 : pl.size()));
  const size_t o_co = in.add(sizeof(int32_t) * cyl_off.size());
  const size_t o_po = in.add(sizeof(int32_t) * pl_off.size());
  const size_t o_pose = in.add(sizeof(double) * pose.size());
  const size_t o_mask = in.add(sizeof(int32_t) * masks.size());
  const size_t o_out = res.add(sizeof(double) * 7 * P);
  const size_t o_par = res.add(sizeof(double) * 6 * P);
  const size_t o_st = res.add(sizeof(int32_t) * 2 * P);
  const size_t o_cost = res.add(sizeof(double) * 2 * P);
  Dev d_in(in.size);
  Dev d_res(res.size);
  char* h = pinned().take(in.size + res.size);
  put(h, o_pts, pts);
  put(h, o_cyl, cyl);
  put(h, o_pl, pl);
  put(h, o_co, cyl_off);
  put(h, o_po, pl_off);
  put(h, o_pose, pose);
  put(h, o_mask, masks);
  const hipStream_t stream = (hipStream_t)sloam_ctx_stream(ctx);
  hip_check(hipMemcpyAsync(d_in.p, h, in.size, hipMemcpyHostToDevice, stream),
            "hipMemcpyAsync(H2D)");
  char* di = d_in.as<char>();
  char* dr = d_res.as<char>();
  check(sloam_pose_plan_exec(plan, (const float*)(di + o_pts), (const double*)(di + o_cyl),
                             (const int32_t*)(di + o_co), (const double*)(di + o_pl),
                             (const int32_t*)(di + o_po), (const double*)(di + o_pose),
                             (const int32_t*)(di + o_mask), (double*)(dr + o_out),
                             (double*)(dr + o_par), (int32_t*)(dr + o_st),
                             (double*)(dr + o_cost)),
        "sloam_pose_plan_exec");
  char* hr = h + in.size;
  hip_check(hipMemcpyAsync(hr, d_res.p, res.size, hipMemcpyDeviceToHost, stream),
            "hipMemcpyAsync(D2H)");
  check(sloam_sync(ctx), "sloam_sync");
  const double* o = (const double*)(hr + o_out);
  const double* par = (const double*)(hr + o_par);
  const int32_t* st = (const int32_t*)(hr + o_st);
  const double* cost = (const double*)(hr + o_cost);
  for (int k = 0; k < P; ++k) {
    std::memcpy(out[k].pose, &o[7 * k], 7 * sizeof(double));
    std::memcpy(out[k].params, &par[6 * k], 6 * sizeof(double));
    out[k].termination = st[2 * k];
    out[k].iterations = st[2 * k + 1];
    out[k].cost[0] = cost[2 * k];
    out[k].cost[1] = cost[2 * k + 1];
  }
  return out;
}

TrellisResult trellis(const VectorType& cloud, int height, int width,
                      const sloam_trellis_cfg& cfg, bool want_labels) {
  CallScope scope;
  TrellisResult out;
  sloam_ctx* ctx = context();
  if (height < 1 || width < 1 || (long long)height * width > 0x7fffffffll)
    throw std::runtime_error("trellis: bad organisation " + std::to_string(height) + " x " +
                             std::to_string(width));
  const int N = height * width;
  if (cloud.size() != (size_t)N)
    throw std::runtime_error("trellis: cloud holds " + std::to_string(cloud.size()) +
                             " points, its organisation " + std::to_string(height) + " x " +
                             std::to_string(width) + " needs " + std::to_string(N));
  const int cap_t = N / (cfg.min_label_points + 1) + 1;
  const int cap_v = cap_t * cfg.max_vertices;
  const int cap_p = N > 0 ? N : 1;
  std::vector<float> pts(4 * (size_t)N);
  for (int i = 0; i < N; ++i) {
    pts[4 * (size_t)i] = cloud[i].x;
    pts[4 * (size_t)i + 1] = cloud[i].y;
    pts[4 * (size_t)i + 2] = cloud[i].z;
    pts[4 * (size_t)i + 3] = cloud[i].intensity;
  }
  Dev d_cloud(sizeof(float) * (pts.empty() ? 4 : pts.size()));
  Dev d_labels(sizeof(int32_t) * (want_labels && N > 0 ? N : 1));
  Dev d_to(sizeof(int32_t) * (cap_t + 1));
  Dev d_vi(sizeof(int32_t) * 4 * (size_t)cap_v);
  Dev d_vf(sizeof(double) * 4 * (size_t)cap_v);
  Dev d_po(sizeof(int32_t) * (cap_v + 1));
  Dev d_pts(sizeof(float) * 4 * (size_t)cap_p);
  Dev d_cnt(sizeof(int32_t) * 8);
  up(d_cloud, pts);
  check(sloam_trellis(ctx, d_cloud.as<float>(), height, width, &cfg,
                      want_labels ? d_labels.as<int32_t>() : nullptr, cap_t, cap_v, cap_p,
                      d_to.as<int32_t>(), d_vi.as<int32_t>(), d_vf.as<double>(),
                      d_po.as<int32_t>(), d_pts.as<float>(), d_cnt.as<int32_t>()),
        "sloam_trellis");
  std::vector<int32_t> cnt(8);
  down(cnt, d_cnt);
  check(sloam_sync(ctx), "sloam_sync");
  if (cnt[4]) throw std::runtime_error("sloam_trellis: output capacity exceeded");
  const int nt = cnt[0], nv = cnt[1], np = cnt[2];
  out.n_labels = cnt[3];
  out.tree_offsets.resize(nt + 1);
  out.vertex_int.resize(4 * (size_t)nv);
  out.vertex_f64.resize(4 * (size_t)nv);
  out.point_offsets.resize(nv + 1);
  out.points.resize(4 * (size_t)np);
  if (want_labels) out.labels.resize(N);
  down(out.tree_offsets, d_to);
  down(out.vertex_int, d_vi);
  down(out.vertex_f64, d_vf);
  down(out.point_offsets, d_po);
  down(out.points, d_pts);
  if (want_labels) down(out.labels, d_labels);
  check(sloam_sync(ctx), "sloam_sync");
  return out;
}

ProjectionResult projectPartition(const std::vector<float>& xyzi,
                                  const std::vector<uint8_t>& labels, const sloam_proj_cfg& cfg,
                                  bool want_tree) {
  CallScope scope;
  sloam_ctx* ctx = context();
  const size_t n = xyzi.size() / 4;
  if (labels.size() != n)
    throw std::runtime_error("projectPartition: " + std::to_string(labels.size()) +
                             " labels for " + std::to_string(n) + " points");
  const size_t pix = (size_t)cfg.height * (size_t)cfg.width;
  Dev d_xyzi(sizeof(float) * 4 * (n ? n : 1));
  Dev d_label(n ? n : 1);
  Dev d_proj(sizeof(int32_t) * 2 * (n ? n : 1));
  Dev d_mask(pix ? pix : 1);
  Dev d_ground(sizeof(float) * 4 * (n ? n : 1));
  Dev d_ng(sizeof(int32_t));
  Dev d_tree(want_tree ? sizeof(float) * 4 * (n ? n : 1) : 16);
  up(d_xyzi, xyzi);
  up(d_label, labels);
  check(sloam_project_partition(ctx, d_xyzi.as<float>(), d_label.as<uint8_t>(), (int64_t)n, &cfg,
                                d_proj.as<int32_t>(), nullptr, nullptr, d_mask.as<uint8_t>(),
                                d_ground.as<float>(), d_ng.as<int32_t>(),
                                want_tree ? d_tree.as<float>() : nullptr),
        "sloam_project_partition");
  ProjectionResult r;
  r.proj.resize(2 * n);
  r.mask.resize(pix);
  std::vector<int32_t> ng(1);
  down(r.proj, d_proj);
  down(r.mask, d_mask);
  down(ng, d_ng);
  if (want_tree) {
    r.tree.resize(4 * n);
    down(r.tree, d_tree);
  }
  check(sloam_sync(ctx), "sloam_sync");
  r.ground.resize(4 * (size_t)ng[0]);
  down(r.ground, d_ground);
  check(sloam_sync(ctx), "sloam_sync");
  return r;
}

}  // namespace sloam_gpu
