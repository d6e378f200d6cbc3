// Keyframe front end on one stream: range projection + ground / tree
// partition, tree instances, ground-plane cells (sloam_project_partition,
// sloam_trellis, sloam_ground_planes) over buffers the plan owns.  With
// use_graph the sequence is captured into a HIP graph at the first exec
// (after one eager run that sizes the scratch buffers) and every later exec
// is a single graph launch.  Every kernel of the sequence reads its sizes
// from the device (ground count, segment counts), and no memset, copy or
// allocation happens inside it, so one graph serves every scan.
#include "common.hpp"

struct sloam_frontend {
  sloam_ctx* ctx = nullptr;
  sloam_frontend_cfg cfg{};
  int N = 0;
  sloam_frontend_out out{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int runs = 0;
};

namespace {

int enqueue(sloam_frontend* f) {
  sloam_ctx* ctx = f->ctx;
  const sloam_frontend_out& o = f->out;
  int st = sloam_project_partition(ctx, o.xyzi, o.label, f->N, &f->cfg.proj, nullptr, nullptr,
                                   nullptr, o.mask, o.ground, o.n_ground, o.tree);
  if (st) return st;
  st = sloam_trellis(ctx, o.tree, f->cfg.proj.height, f->cfg.proj.width, &f->cfg.trellis,
                     nullptr, o.cap_trees, o.cap_vertices, o.cap_points, o.tree_offsets,
                     o.vertex_int, o.vertex_f64, o.point_offsets, o.points, o.trellis_counts);
  if (st) return st;
  return sloam_ground_planes(ctx, o.ground, f->N, o.n_ground, &f->cfg.ground, o.cell_info,
                             o.cell_model, o.cell_features, o.kept_offsets, o.kept_index);
}

}  // namespace

extern "C" {

int sloam_frontend_create(sloam_ctx* ctx, const sloam_frontend_cfg* cfg, sloam_frontend** out) {
  if (!ctx || !cfg || !out) return SLOAM_ERR_INVALID_ARG;
  *out = nullptr;
  const sloam_proj_cfg& pc = cfg->proj;
  if (pc.height < 1 || pc.width < 1) return SLOAM_ERR_INVALID_ARG;
  const long long N = (long long)pc.height * pc.width;
  if (N >= (1ll << 20)) return SLOAM_ERR_TOO_LARGE;
  const int n_cells = cfg->ground.radii_bins * cfg->ground.theta_bins;
  if (n_cells < 1 || n_cells > 65535 || cfg->ground.num_features < 1 ||
      cfg->trellis.max_vertices < 1)
    return SLOAM_ERR_INVALID_ARG;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  sloam_frontend* f = new (std::nothrow) sloam_frontend();
  if (!f) return SLOAM_ERR_ALLOC;
  f->ctx = ctx;
  f->cfg = *cfg;
  f->N = (int)N;
  sloam_frontend_out& o = f->out;
  o.cap_trees = (int)(N / (cfg->trellis.min_label_points + 1)) + 1;
  o.cap_vertices = o.cap_trees * cfg->trellis.max_vertices;
  o.cap_points = (int)N;
  o.n_cells = n_cells;
  struct Buf {
    void** p;
    size_t bytes;
  };
  const size_t n = (size_t)N, nf = (size_t)cfg->ground.num_features;
  Buf bufs[] = {
      {(void**)&o.xyzi, 16 * n},
      {(void**)&o.label, n},
      {(void**)&o.mask, n},
      {(void**)&o.ground, 16 * n},
      {(void**)&o.n_ground, 4},
      {(void**)&o.tree, 16 * n},
      {(void**)&o.tree_offsets, 4 * (size_t)(o.cap_trees + 1)},
      {(void**)&o.vertex_int, 16 * (size_t)o.cap_vertices},
      {(void**)&o.vertex_f64, 32 * (size_t)o.cap_vertices},
      {(void**)&o.point_offsets, 4 * (size_t)(o.cap_vertices + 1)},
      {(void**)&o.points, 16 * n},
      {(void**)&o.trellis_counts, 32},
      {(void**)&o.cell_info, 16 * (size_t)n_cells},
      {(void**)&o.cell_model, 64 * (size_t)n_cells},
      {(void**)&o.cell_features, 16 * (size_t)n_cells * nf},
      {(void**)&o.kept_offsets, 4 * (size_t)(n_cells + 1)},
      {(void**)&o.kept_index, 4 * n},
  };
  for (auto& b : bufs) {
    const int st = sloam_dev_alloc(ctx, b.p, b.bytes);
    if (st) {
      sloam_frontend_destroy(f);
      return st;
    }
  }
  *out = f;
  return SLOAM_OK;
}

void sloam_frontend_destroy(sloam_frontend* f) {
  if (!f) return;
  if (f->ctx) (void)hipStreamSynchronize(f->ctx->stream);
  if (f->exec) (void)hipGraphExecDestroy(f->exec);
  if (f->graph) (void)hipGraphDestroy(f->graph);
  sloam_frontend_out& o = f->out;
  void* ptrs[] = {o.xyzi,          o.label,      o.mask,          o.ground,      o.n_ground,
                  o.tree,          o.tree_offsets, o.vertex_int,  o.vertex_f64,  o.point_offsets,
                  o.points,        o.trellis_counts, o.cell_info, o.cell_model,  o.cell_features,
                  o.kept_offsets,  o.kept_index};
  for (void* p : ptrs) sloam_dev_free(p);
  delete f;
}

int sloam_frontend_buffers(const sloam_frontend* f, sloam_frontend_out* out) {
  if (!f || !out) return SLOAM_ERR_INVALID_ARG;
  *out = f->out;
  return SLOAM_OK;
}

int sloam_frontend_exec(sloam_frontend* f) {
  if (!f) return SLOAM_ERR_INVALID_ARG;
  sloam_ctx* ctx = f->ctx;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  if (f->exec) {
    SloamKernelScope ks(ctx, "frontend_graph");
    SLOAM_HIP(ctx, hipGraphLaunch(f->exec, ctx->stream));
    ++f->runs;
    return SLOAM_OK;
  }
  int st = enqueue(f);
  if (st) return st;
  ++f->runs;
  if (f->cfg.use_graph) {
    // the eager run has sized every scratch buffer: capture the same sequence
    SLOAM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const bool timing = ctx->timing;
    ctx->timing = false;
    SLOAM_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    st = enqueue(f);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    ctx->timing = timing;
    if (st) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    if (e != hipSuccess) return sloam_record_hip_error(ctx, e, "hipStreamEndCapture");
    f->graph = g;
    SLOAM_HIP(ctx, hipGraphInstantiate(&f->exec, g, nullptr, nullptr, 0));
  }
  return SLOAM_OK;
}

}  // extern "C"
