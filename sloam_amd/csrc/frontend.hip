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
  uint32_t slots = 0;  // context scratch slots held by the graph
  int runs = 0;
  // host_results: pinned host buffer (host and device address) and the word
  // offsets of its regions
  void* host = nullptr;
  void* host_dev = nullptr;
  size_t region[11] = {};
  float* medians = nullptr;  // fit_lines: vertex medians as float, 3 per vertex
};

namespace {

// Regions of the pinned result buffer, in 32-bit words: counts, tree offsets,
// vertex ints, vertex doubles, point offsets, points, cell info, cell models,
// cell features, line coefficients, line stats; each starts on a 64-word
// boundary.
enum { kRCounts, kRTreeOff, kRVtxInt, kRVtxF64, kRPtOff, kRPoints, kRCellInfo, kRCellModel,
       kRCellFeat, kRLineCoeff, kRLineStats, kRegions };

struct PackArgs {
  const int32_t* counts;
  const uint32_t* src[kRegions];
  uint32_t* dst[kRegions];
  int cap_trees, cap_vertices, cap_points, n_cells, nf, lines;
};

// Copies the used part of every output into the pinned host buffer (one
// grid-stride loop per region; lengths from the device-side counts, clamped
// to the capacities).
__global__ __launch_bounds__(256) void fe_pack_kernel(PackArgs a) {
  const int nt = min(max(a.counts[0], 0), a.cap_trees);
  const int nv = min(max(a.counts[1], 0), a.cap_vertices);
  const int np = min(max(a.counts[2], 0), a.cap_points);
  const size_t len[kRegions] = {8,
                                (size_t)nt + 1,
                                4 * (size_t)nv,
                                8 * (size_t)nv,
                                (size_t)nv + 1,
                                4 * (size_t)np,
                                4 * (size_t)a.n_cells,
                                16 * (size_t)a.n_cells,
                                4 * (size_t)a.n_cells * (size_t)a.nf,
                                a.lines ? 6 * (size_t)nt : 0,
                                a.lines ? 4 * (size_t)nt : 0};
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
  for (int r = 0; r < kRegions; ++r)
    for (size_t i = t0; i < len[r]; i += stride) a.dst[r][i] = a.src[r][i];
}

// Vertex medians (vertex_f64 x, y, z) as float, the points the line fit of a
// tree takes.
__global__ __launch_bounds__(256) void fe_medians_kernel(const double* __restrict__ vf,
                                                         const int32_t* __restrict__ counts,
                                                         int cap_vertices,
                                                         float* __restrict__ med) {
  const int nv = min(max(counts[1], 0), cap_vertices);
  for (int v = blockIdx.x * 256 + threadIdx.x; v < nv; v += gridDim.x * 256) {
    med[3 * (size_t)v] = (float)vf[4 * (size_t)v + 1];
    med[3 * (size_t)v + 1] = (float)vf[4 * (size_t)v + 2];
    med[3 * (size_t)v + 2] = (float)vf[4 * (size_t)v + 3];
  }
}

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
  if (f->cfg.fit_lines) {
    hipLaunchKernelGGL(fe_medians_kernel, dim3(64), dim3(256), 0, ctx->stream, o.vertex_f64,
                       o.trellis_counts, o.cap_vertices, f->medians);
    SLOAM_HIP(ctx, hipGetLastError());
    // one set per tree (up to the device-side tree count)
    st = sloam_fit_lines_dev(ctx, f->medians, o.tree_offsets, o.cap_trees, o.trellis_counts,
                             &f->cfg.line, o.line_coeff, o.line_stats, nullptr);
    if (st) return st;
  }
  st = sloam_ground_planes(ctx, o.ground, f->N, o.n_ground, &f->cfg.ground, o.cell_info,
                           o.cell_model, o.cell_features, o.kept_offsets, o.kept_index);
  if (st || !f->host) return st;
  PackArgs a;
  a.counts = o.trellis_counts;
  const void* src[kRegions] = {o.trellis_counts, o.tree_offsets,  o.vertex_int, o.vertex_f64,
                               o.point_offsets,  o.points,        o.cell_info,  o.cell_model,
                               o.cell_features,  o.line_coeff,    o.line_stats};
  for (int r = 0; r < kRegions; ++r) {
    a.src[r] = static_cast<const uint32_t*>(src[r]);
    a.dst[r] = static_cast<uint32_t*>(f->host_dev) + f->region[r];
  }
  a.cap_trees = o.cap_trees;
  a.cap_vertices = o.cap_vertices;
  a.cap_points = o.cap_points;
  a.n_cells = o.n_cells;
  a.nf = f->cfg.ground.num_features;
  a.lines = f->cfg.fit_lines ? 1 : 0;
  SloamKernelScope ks(ctx, "frontend_pack");
  hipLaunchKernelGGL(fe_pack_kernel, dim3(64), dim3(256), 0, ctx->stream, a);
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
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
  if (cfg->fit_lines &&
      (!(cfg->line.threshold > 0.0) || cfg->line.max_iterations < 0 ||
       !(cfg->line.probability > 0.0) || !(cfg->line.probability < 1.0)))
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
      {(void**)&o.line_coeff, cfg->fit_lines ? 24 * (size_t)o.cap_trees : 0},
      {(void**)&o.line_stats, cfg->fit_lines ? 16 * (size_t)o.cap_trees : 0},
      {(void**)&f->medians, cfg->fit_lines ? 12 * (size_t)o.cap_vertices : 0},
  };
  for (auto& b : bufs) {
    const int st = sloam_dev_alloc(ctx, b.p, b.bytes);
    if (st) {
      sloam_frontend_destroy(f);
      return st;
    }
  }
  if (cfg->host_results) {
    const size_t words[kRegions] = {8,
                                    (size_t)o.cap_trees + 1,
                                    4 * (size_t)o.cap_vertices,
                                    8 * (size_t)o.cap_vertices,
                                    (size_t)o.cap_vertices + 1,
                                    4 * (size_t)o.cap_points,
                                    4 * (size_t)n_cells,
                                    16 * (size_t)n_cells,
                                    4 * (size_t)n_cells * nf,
                                    cfg->fit_lines ? 6 * (size_t)o.cap_trees : 0,
                                    cfg->fit_lines ? 4 * (size_t)o.cap_trees : 0};
    size_t total = 0;
    for (int r = 0; r < kRegions; ++r) {
      f->region[r] = total;
      total += (words[r] + 63) / 64 * 64;
    }
    hipError_t e = hipHostMalloc(&f->host, 4 * total, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostGetDevicePointer(&f->host_dev, f->host, 0);
    if (e != hipSuccess) {
      sloam_record_hip_error(ctx, e, "hipHostMalloc(frontend results)");
      sloam_frontend_destroy(f);
      return SLOAM_ERR_ALLOC;
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
  sloam_ctx_pin_slots(f->ctx, f->slots, -1);
  sloam_frontend_out& o = f->out;
  void* ptrs[] = {o.xyzi,          o.label,      o.mask,          o.ground,      o.n_ground,
                  o.tree,          o.tree_offsets, o.vertex_int,  o.vertex_f64,  o.point_offsets,
                  o.points,        o.trellis_counts, o.cell_info, o.cell_model,  o.cell_features,
                  o.kept_offsets,  o.kept_index, o.line_coeff,  o.line_stats,  f->medians};
  for (void* p : ptrs) sloam_dev_free(p);
  if (f->host) (void)hipHostFree(f->host);
  delete f;
}

int sloam_frontend_buffers(const sloam_frontend* f, sloam_frontend_out* out) {
  if (!f || !out) return SLOAM_ERR_INVALID_ARG;
  *out = f->out;
  return SLOAM_OK;
}

int sloam_frontend_host_view(const sloam_frontend* f, sloam_frontend_host* out) {
  if (!f || !out || !f->host) return SLOAM_ERR_INVALID_ARG;
  const uint32_t* h = static_cast<const uint32_t*>(f->host);
  out->trellis_counts = reinterpret_cast<const int32_t*>(h + f->region[kRCounts]);
  out->tree_offsets = reinterpret_cast<const int32_t*>(h + f->region[kRTreeOff]);
  out->vertex_int = reinterpret_cast<const int32_t*>(h + f->region[kRVtxInt]);
  out->vertex_f64 = reinterpret_cast<const double*>(h + f->region[kRVtxF64]);
  out->point_offsets = reinterpret_cast<const int32_t*>(h + f->region[kRPtOff]);
  out->points = reinterpret_cast<const float*>(h + f->region[kRPoints]);
  out->cell_info = reinterpret_cast<const int32_t*>(h + f->region[kRCellInfo]);
  out->cell_model = reinterpret_cast<const double*>(h + f->region[kRCellModel]);
  out->cell_features = reinterpret_cast<const float*>(h + f->region[kRCellFeat]);
  out->line_coeff =
      f->cfg.fit_lines ? reinterpret_cast<const float*>(h + f->region[kRLineCoeff]) : nullptr;
  out->line_stats =
      f->cfg.fit_lines ? reinterpret_cast<const int32_t*>(h + f->region[kRLineStats]) : nullptr;
  return SLOAM_OK;
}

int sloam_frontend_enqueue(sloam_frontend* f) {
  if (!f) return SLOAM_ERR_INVALID_ARG;
  SLOAM_HIP(f->ctx, hipSetDevice(f->ctx->device));
  const int st = enqueue(f);
  if (!st) ++f->runs;
  return st;
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
    if (ctx->capturing) return SLOAM_ERR_INVALID_ARG;
    const bool timing = ctx->timing;
    ctx->timing = false;
    SLOAM_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    ctx->capture_slots = 0;
    st = enqueue(f);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    ctx->capturing = false;
    ctx->timing = timing;
    if (st) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    if (e != hipSuccess) return sloam_record_hip_error(ctx, e, "hipStreamEndCapture");
    f->graph = g;
    SLOAM_HIP(ctx, hipGraphInstantiate(&f->exec, g, nullptr, nullptr, 0));
    f->slots = ctx->capture_slots;
    sloam_ctx_pin_slots(ctx, f->slots, +1);
  }
  return SLOAM_OK;
}

}  // extern "C"
