/*
 * sloam_gpu.h -- C ABI of the MI355X (gfx950) implementation of SLOAM's
 * geometric hot path.
 *
 * The reference (KumarRobotics/sloam) exposes no FFI for this path: its
 * boundary is the C++ class API (sloam::sloam::RunSloam, Cylinder, Plane,
 * seg::Segmentation).  The C++ drop-in classes in sloam_amd/host/ keep those
 * names and call the entry points below; any other host language binds them
 * directly (plain C types and pointers: ctypes, cgo, JNI, N-API).
 *
 * Conventions
 *   - every function returns int status: 0 = OK, < 0 = error (sloam_strerror);
 *     no C++ exception crosses this boundary;
 *   - "d_" pointers are device (HBM) pointers owned by the caller, "h_" pointers
 *     are host pointers owned by the caller; scratch memory is owned by the
 *     context (or by a plan created from it);
 *   - all work is enqueued on the context's HIP stream and is asynchronous until
 *     sloam_sync(); a context is not thread-safe (one context per host thread);
 *   - points are float32 xyz, 12 bytes each, array-of-structs.
 */
#ifndef SLOAM_GPU_H_
#define SLOAM_GPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define SLOAM_OK 0
#define SLOAM_ERR_INVALID_ARG (-1)
#define SLOAM_ERR_HIP (-2)
#define SLOAM_ERR_NO_DEVICE (-3)
#define SLOAM_ERR_TOO_LARGE (-4)
#define SLOAM_ERR_ALLOC (-5)
#define SLOAM_ERR_INTERNAL (-6)

typedef struct sloam_ctx sloam_ctx;

/* ---- context ------------------------------------------------------------
 * A context owns one HIP stream (default priority) on one device, the
 * scratch buffers of the entry points below and optional per-kernel timers.
 * Replaces: the per-process state of sloam::sloam (sloam/include/core/sloam.h:98-106)
 * as far as device resources go; there is no device state in the reference. */
int sloam_ctx_create(int device, sloam_ctx** out);
void sloam_ctx_destroy(sloam_ctx* ctx);
int sloam_sync(sloam_ctx* ctx);
const char* sloam_strerror(int status);
/* Last HIP error string recorded by the context (empty if none). */
const char* sloam_ctx_last_error(const sloam_ctx* ctx);
/* The context's hipStream_t, as an opaque pointer (for event timing). */
void* sloam_ctx_stream(sloam_ctx* ctx);
/* Per-kernel timing with HIP events on the context stream.  When enabled,
 * every launch of a named kernel is bracketed by events; after sloam_sync()
 * sloam_ctx_kernel_time() returns the summed milliseconds and launch count. */
int sloam_ctx_enable_timing(sloam_ctx* ctx, int enable);
/* Copy of `bytes` bytes between host and device memory (either direction,
 * unified addressing) on the context stream; host memory must stay valid
 * until sloam_sync.                                                         */
int sloam_memcpy_async(sloam_ctx* ctx, void* dst, const void* src, size_t bytes);
/* d_xyz[3 i + k] = d_xyzi[4 i + k], k < 3, for n points (device buffers,
 * d_xyzi 16-byte aligned): the x y z intensity points of the front end
 * (sloam_frontend_out.points, the reference's PointT) as the packed xyz the
 * cylinder fit reads.  No counterpart in the reference (its fits read PCL
 * clouds in place).                                                        */
int sloam_xyz_from_xyzi(sloam_ctx* ctx, const float* d_xyzi, int64_t n, float* d_xyz);
int sloam_ctx_kernel_time(sloam_ctx* ctx, const char* kernel, double* ms_total, int64_t* launches);
int sloam_ctx_reset_timing(sloam_ctx* ctx);

/* ---- HIP graphs of library calls ---------------------------------------
 * No counterpart in the reference.  Between sloam_graph_begin and
 * sloam_graph_end, the entry points called on ctx are recorded into a HIP
 * graph instead of run (with their fork / join over the context's auxiliary
 * streams); sloam_graph_launch replays the whole sequence as one launch on
 * the context stream, over the same buffers and arguments.  Run the sequence
 * once before capturing it (it sizes the context's scratch buffers; a
 * buffer that would have to grow during capture makes the call fail); entry
 * points that synchronise, copy from the host or capture graphs themselves
 * (sloam_frontend_exec, plan creation) do not belong in a capture.  Timing
 * is off during capture; a replay is timed as "graph".
 * A captured graph (this one, or a front end's) holds the context scratch
 * buffers its kernels use until it is destroyed: while it lives, a call on
 * the same context that would need a larger scratch buffer than the one the
 * graph holds fails with SLOAM_ERR_INVALID_ARG (sloam_ctx_last_error names
 * the buffer) instead of moving memory the graph still writes.  Run the
 * largest inputs before capturing, or give them a context of their own.
 * Destroying the context first releases the graph's HIP objects: the graph
 * is then orphaned (sloam_graph_launch returns SLOAM_ERR_INVALID_ARG) and
 * sloam_graph_destroy only frees its handle.  Plans, front ends and
 * estimators created from a context must be destroyed before it.          */
typedef struct sloam_graph sloam_graph;
int sloam_graph_begin(sloam_ctx* ctx);
int sloam_graph_end(sloam_ctx* ctx, sloam_graph** out);
int sloam_graph_launch(sloam_graph* graph);
void sloam_graph_destroy(sloam_graph* graph);

/* ---- batched cylinder RANSAC fitting (point-level, fixed hypotheses) ------
 * For each cluster c (points d_xyz[offsets[c] .. offsets[c+1])):
 *   1. axis seed: largest-eigenvector of the cluster's point covariance
 *      (integer-exact sums, FP64 3x3 Jacobi eigen-solve);
 *   2. n_hyp hypotheses; hypothesis h draws 3 distinct point indices from
 *      Philox4x32-10(counter = {h, c, stream_tag, 0}, key = seed), projects them
 *      on the plane orthogonal to the axis and takes their circumcircle
 *      (centre on the axis, radius r); hypotheses with a degenerate triangle or
 *      r > max_radius score 0;
 *   3. inlier predicate | ||(p - c) x a|| - r | < thr, evaluated for every
 *      (point, hypothesis) pair in FP32 as L < u*A + v*B + s < U (squared
 *      form, see sloam_amd/csrc/cylfit.hip); counts per hypothesis;
 *   4. best hypothesis = max count, lowest index on ties; inlier bit mask;
 *   5. FP64 LO-RANSAC refit: refit_rounds x (refit_iters Gauss-Newton steps
 *      of the 5-parameter geometric cylinder fit over the current inliers,
 *      then re-selection |d| < thr under the new model); the first round
 *      starts from the best hypothesis and its FP32 inliers.
 * The per-point predicate is the reference's point-to-cylinder distance
 * (Cylinder::distance(const PointT&), sloam/src/objects/cylinder.cpp:196-203)
 * with a unit ray; the hypothesis generator has no counterpart in the
 * reference, whose Cylinder::computeModel (cylinder.cpp:75-173) fits a line to
 * trellis vertex medians instead.                                            */
typedef struct sloam_cyl_fit_cfg {
  int32_t n_hyp;          /* hypotheses per cluster (e.g. 200, 1024)          */
  float thr;              /* inlier threshold, metres (e.g. 0.05)             */
  float max_radius;       /* hypotheses with r > max_radius are rejected      */
  int32_t refit_iters;    /* Gauss-Newton steps per LO refit round            */
  uint64_t seed;          /* Philox key                                       */
  int32_t stream_tag;     /* Philox counter word 2 (separates sample streams) */
  int32_t hyp_chunk;      /* hypotheses per workgroup (0 = autotuned)         */
  int32_t tile_points;    /* points per LDS tile (0 = autotuned)              */
  int32_t refit_rounds;   /* LO refit rounds (Gauss-Newton + re-selection)    */
} sloam_cyl_fit_cfg;

/* Output model record, one per cluster (8 doubles):
 *   [0..2] root (axis point at the height of the final inlier centroid)
 *   [3..5] ray  (unit axis; largest |component| positive)
 *   [6]    radius (-1 if the fit is invalid: fewer than 3 inliers)
 *   [7]    number of inliers of the refitted model (FP64 |d| < thr)       */
#define SLOAM_CYL_MODEL_STRIDE 8

typedef struct sloam_cyl_plan sloam_cyl_plan;

/* Plan a batch: uploads the work list (clusters ordered largest first and
 * split into hypothesis chunks) once; h_offsets has n_clusters+1 entries.  */
int sloam_cyl_plan_create(sloam_ctx* ctx, const int32_t* h_offsets, int32_t n_clusters,
                          const sloam_cyl_fit_cfg* cfg, sloam_cyl_plan** out);
void sloam_cyl_plan_destroy(sloam_cyl_plan* plan);
/* Re-plan for another batch with the same settings (no counterpart in the
 * reference, whose fits allocate per call): waits for the context stream,
 * then lays out h_offsets (n_clusters + 1 entries) as sloam_cyl_plan_create
 * does, reusing the plan's device buffers (they only grow).  Not allowed
 * during graph capture.  On error the plan holds an empty batch.          */
int sloam_cyl_plan_reset(sloam_cyl_plan* plan, const int32_t* h_offsets, int32_t n_clusters);
/* Counted plan (no counterpart in the reference): the batch's size is read
 * from device memory at every exec (*d_n_clusters clusters, their offsets in
 * the exec's d_offsets), so that an exec needs no host-side sizes and can be
 * captured in a HIP graph whose replays fit other batches, e.g. a
 * keyframe's trees straight from its front end.  The layout (work list,
 * refit chunks) is made on the device by the exec's first kernel; clusters
 * run in index order, and every result equals a host-planned plan's.  At
 * most cap_clusters clusters and cap_points points, and at most 131072
 * points per cluster (the host-planned plans' bound, SLOAM_ERR_TOO_LARGE
 * there): a larger batch sets the overflow word of sloam_cyl_plan_counts and
 * fits nothing.  Exec without
 * d_mask; not with sloam_cyl_plan_reset.                                    */
int sloam_cyl_plan_create_counted(sloam_ctx* ctx, int32_t cap_clusters, int64_t cap_points,
                                  const int32_t* d_n_clusters, const sloam_cyl_fit_cfg* cfg,
                                  sloam_cyl_plan** out);
/* Device pointer to 4 int32 of a counted plan, written by each exec:
 * clusters, work items, refit chunks, overflow (1: the batch exceeded the
 * capacities).                                                              */
int sloam_cyl_plan_counts(const sloam_cyl_plan* plan, const int32_t** d_counts);
/* Run a planned batch.  d_offsets must hold the same offsets as h_offsets.
 * Optional outputs may be NULL:
 *   d_counts  [n_clusters * n_hyp] int32 inlier count per hypothesis
 *   d_samples [n_clusters * n_hyp * 3] int32 sampled point indices (cluster-local)
 *   d_mask    [(n_points + 31) / 32] uint32 inlier bits of the best hypothesis,
 *             bit i of word w = global point 32*w + i
 *   d_best    [n_clusters] int32 best hypothesis index (-1 if n < 3)
 * Required: d_model [n_clusters * SLOAM_CYL_MODEL_STRIDE] double.            */
int sloam_cyl_plan_exec(sloam_cyl_plan* plan, const float* d_xyz, const int32_t* d_offsets,
                        int32_t* d_counts, int32_t* d_samples, uint32_t* d_mask,
                        int32_t* d_best, double* d_model);
/* The same with two optional inputs:
 *   d_proposal [n_clusters * SLOAM_PN_PROPOSAL_STRIDE] floats (sloam_pn_infer):
 *     for a cluster with a valid proposal, hypothesis 0 is the proposed
 *     circle (centre taken into the cluster's PCA seed frame, proposed
 *     radius; its samples are -1);
 *   d_active [n_clusters] int32: clusters with 0 are left untouched (no
 *     seed, scoring or refit; their model records keep their contents).
 * d_mask must be NULL when d_active is given.                               */
int sloam_cyl_plan_exec_ex(sloam_cyl_plan* plan, const float* d_xyz, const int32_t* d_offsets,
                           const float* d_proposal, const int32_t* d_active, int32_t* d_counts,
                           int32_t* d_samples, uint32_t* d_mask, int32_t* d_best,
                           double* d_model);
/* Learned fit in one pass over the batch (no counterpart in the reference):
 * one axis seed per cluster; the `verify` plan's hypotheses (hypothesis 0
 * the proposal of d_proposal, as sloam_cyl_plan_exec_ex) scored for every
 * cluster; a cluster whose best verifying hypothesis holds fewer than
 * min_inlier_ratio * n inliers is flagged in d_fallback (count in
 * d_n_fallback, nullable) and scored with the `plain` plan's hypotheses
 * (no proposal); one LO refit of every cluster from its best hypothesis of
 * the set it was scored with, with that plan's refit_rounds / refit_iters.
 * Then a rescue: a cluster the verify stage kept whose refitted model is
 * invalid or keeps fewer than min_inlier_ratio * n inliers is flagged too
 * (d_fallback, d_n_fallback), scored with the plain hypotheses and refitted
 * again.  A flagged cluster's model equals the plain plan's own fit of it, an
 * unflagged one the verify plan's.  Both plans must be built on the same
 * offsets with the same thr, max_radius and seed.  d_counts_verify /
 * d_counts_plain (nullable: plan scratch) hold n_clusters x n_hyp counts of
 * their plan.                                                               */
int sloam_cyl_learned_exec(sloam_cyl_plan* verify, sloam_cyl_plan* plain, const float* d_xyz,
                           const int32_t* d_offsets, const float* d_proposal,
                           float min_inlier_ratio, int32_t* d_counts_verify,
                           int32_t* d_counts_plain, int32_t* d_fallback, int32_t* d_n_fallback,
                           int32_t* d_best, double* d_model);
/* Launch-shape autotuner: times plans with hyp_chunk in {64, 128, 256, 512}
 * and tile_points in {1024, 2048, 4096} on this batch (d_xyz / d_offsets on
 * the device, h_offsets on the host) and returns cfg with the fastest pair in
 * *out.  Every shape gives identical results; the choice is cached per
 * (device, clusters, points, hypotheses).                                   */
int sloam_cyl_autotune(sloam_ctx* ctx, const int32_t* h_offsets, int32_t n_clusters,
                       const sloam_cyl_fit_cfg* cfg, const float* d_xyz,
                       const int32_t* d_offsets, sloam_cyl_fit_cfg* out);
/* d_active[c] = 1 for clusters (>= 3 points) whose model is invalid or keeps
 * fewer than min_inlier_ratio x n points as inliers (model[7]); d_n_active
 * (nullable, device) gets their number.  With exec_ex this runs a fallback
 * fit of exactly those clusters.                                            */
int sloam_cyl_fallback(sloam_ctx* ctx, const int32_t* d_offsets, int32_t n_clusters,
                       const double* d_model, float min_inlier_ratio, int32_t* d_active,
                       int32_t* d_n_active);

/* ---- batched pose estimation (Levenberg-Marquardt, Huber loss) ------------
 * Replaces: sloam::sloam::OptimizePose (sloam/src/core/sloam.cpp:174-255) and
 * its residual functors CylinderCost (sloam/include/objects/cylinder.h:90-132)
 * and PlaneCost (sloam/include/objects/plane.h:76-116), solved there by Ceres
 * (LM, DENSE_QR, HuberLoss(0.1), max 50 iterations).
 * Each problem k owns residuals [res_offsets[k], res_offsets[k+1]) of
 * d_points, cylinders [cyl_offsets[k], ...) and planes [plane_offsets[k], ...).
 * d_points: 4 floats per residual: x, y, z (sensor frame) and the landmark
 *   reference as int32 bits: id >= 0 -> cylinder id, id < 0 -> plane -(id+1)
 *   (global indices into d_cyl / d_planes).
 * d_cyl: 8 doubles per cylinder (map frame): root[3], ray[3], radius, unused.
 * d_planes: 4 doubles per plane (map frame): a, b, c, d of a x + b y + c z + d.
 * d_pose_in / d_pose_out: 7 doubles per problem: t[3], q[4] (x, y, z, w),
 *   the map <- sensor transform (lp = q * p + t).
 * d_status (nullable): 2 int32 per problem: termination (0 CONVERGENCE,
 *   1 NO_CONVERGENCE) and iterations; d_cost (nullable): 2 doubles per problem:
 *   initial and final cost 1/2 sum rho(r^2).
 * precision 0: residuals and Jacobians are FP32 (relative to the initial
 *   translation), the normal equations are summed with error-free FP32
 *   transformations and solved in FP64 (the throughput form);
 * precision 1: residuals, Jacobians and sums are FP64, as Ceres evaluates them
 *   (the keyframe solves use it: their termination tests then see Ceres'
 *   costs).
 * parametrization 0 (OptimizePose, sloam.cpp:174-255): t + unit quaternion,
 *   rotation updated by left multiplication (Ceres' quaternion local
 *   parameterization); all six parameters free.
 * parametrization 1 (OptimizeXYYaw / OptimizeZRollPitch, sloam.cpp:55-172):
 *   the 6-vector [t, angle-axis w] (ceres::QuaternionToAngleAxis of the
 *   initial rotation) updated additively; d_free_mask (nullable: all free)
 *   holds per problem a 6-bit mask of the optimised entries, as Ceres'
 *   SubsetParameterization fixes the others ({2,3,4} fixed for XYYaw, {0,1,5}
 *   for ZRollPitch).
 * d_params_out (nullable): 6 doubles per problem, [t, angle-axis w] of the
 *   result.                                                                  */
typedef struct sloam_pose_cfg {
  int32_t max_iters;             /* LM iterations (reference: 50)             */
  int32_t stop_on_convergence;   /* 1: stop at convergence; 0: run max_iters  */
  double huber_delta;            /* Huber loss scale (reference: 0.1)         */
  double function_tolerance;     /* |dF| / F (Ceres default 1e-6)             */
  double gradient_tolerance;     /* ||g||_inf (Ceres default 1e-10)           */
  double parameter_tolerance;    /* ||dx|| / (||x|| + tol) (1e-8)             */
  double initial_radius;         /* trust region (Ceres default 1e4)          */
  double min_relative_decrease;  /* step acceptance (Ceres default 1e-3)      */
  int32_t parametrization;       /* 0 quaternion / left, 1 angle-axis subset  */
  int32_t precision;             /* 0 FP32 compensated, 1 FP64 evaluation     */
} sloam_pose_cfg;

typedef struct sloam_pose_plan sloam_pose_plan;

int sloam_pose_plan_create(sloam_ctx* ctx, const int32_t* h_res_offsets, int32_t n_problems,
                           int32_t n_cylinders, int32_t n_planes, const sloam_pose_cfg* cfg,
                           sloam_pose_plan** out);
void sloam_pose_plan_destroy(sloam_pose_plan* plan);
/* Re-targets an existing plan at another problem set (same argument meaning
 * as sloam_pose_plan_create; cfg NULL keeps the plan's configuration).  Its
 * device buffers only grow, so a caller solving one small problem set per
 * keyframe allocates nothing after the first.  Synchronises the context's
 * stream first (the previous exec may still read the plan). */
int sloam_pose_plan_reset(sloam_pose_plan* plan, const int32_t* h_res_offsets,
                          int32_t n_problems, int32_t n_cylinders, int32_t n_planes,
                          const sloam_pose_cfg* cfg);
/* Execution of sloam_pose_plan_exec:
 *   SLOAM_POSE_MODE_BATCHED: per LM iteration one residual/Jacobian kernel
 *     over 2048-residual chunks of every problem and one LM-step kernel
 *     (2 * (max_iters + 1) launches; for many or large problems);
 *   SLOAM_POSE_MODE_FUSED: one workgroup per problem runs every iteration in
 *     one launch (for few small problems, e.g. one keyframe's solve);
 *   SLOAM_POSE_MODE_AUTO (default): FUSED when no problem has more than
 *     8 chunks (16384 residuals), else BATCHED.
 * Both modes evaluate the same chunks in the same order and give
 * bit-identical results. */
#define SLOAM_POSE_MODE_AUTO 0
#define SLOAM_POSE_MODE_BATCHED 1
#define SLOAM_POSE_MODE_FUSED 2
int sloam_pose_plan_set_mode(sloam_pose_plan* plan, int32_t mode);
/* 1 if the next exec runs the fused kernel, 0 if the batched loop. */
int sloam_pose_plan_uses_fused(const sloam_pose_plan* plan);
int sloam_pose_plan_exec(sloam_pose_plan* plan, const float* d_points, const double* d_cyl,
                         const int32_t* d_cyl_offsets, const double* d_planes,
                         const int32_t* d_plane_offsets, const double* d_pose_in,
                         const int32_t* d_free_mask, double* d_pose_out, double* d_params_out,
                         int32_t* d_status, double* d_cost);

/* ---- range projection + ground / tree partition ---------------------------
 * Replaces: seg::Segmentation::_doProjection / _makeTensor / _mask / maskCloud /
 * _destaggerCloud (sloam/src/segmentation/inference.cpp:80-300,
 * sloam/include/segmentation/inference.h:57-79) with the per-point labels given
 * (0 other, 1 ground, 2 tree) instead of the ONNX network's argmax.
 * d_xyzi: n points of 4 floats (x, y, z, intensity) in input order; d_label: n
 * bytes.  Outputs (H = height, W = width):
 *   d_proj   (nullable) n x 2 int32: column, row (-1, -1 for non-finite points)
 *   d_range  (nullable) H*W float range image (nearest point; 0 = empty)
 *   d_winner (nullable) H*W int32 index of the pixel's nearest point (-1 empty)
 *   d_mask   H*W uint8: 255 tree, 1 ground, 0 other / invalid (range < 1 m)
 *   d_ground capacity n x 4 floats: points whose pixel is ground, input order;
 *            the count goes to d_n_ground (one int32, device)
 *   d_tree   (nullable) n x 4 floats: dense cloud, the point where its pixel
 *            is tree, NaN elsewhere; destagger shifts even rows by +32 columns
 *            (requires n == H*W).                                            */
typedef struct sloam_proj_cfg {
  int32_t height;        /* range image rows (64)                           */
  int32_t width;         /* range image columns (2048)                      */
  float fov_up_deg;      /* 22.5                                            */
  float fov_down_deg;    /* -22.5                                           */
  int32_t destagger;     /* 0 / 1                                           */
} sloam_proj_cfg;

int sloam_project_partition(sloam_ctx* ctx, const float* d_xyzi, const uint8_t* d_label,
                            int64_t n, const sloam_proj_cfg* cfg, int32_t* d_proj, float* d_range,
                            int32_t* d_winner, uint8_t* d_mask, float* d_ground,
                            int32_t* d_n_ground, float* d_tree);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, outputs and errors as
 * sloam_project_partition, but its scratch (z-buffer keys, point
 * projections, block counts) is a fixed array in device memory that each GPU
 * holds for the life of the process (about 8.4 MB), so the call allocates and
 * frees nothing, and it records no kernel timings.  Limits: height * width <=
 * 524288 pixels and n <= 524288 points, else SLOAM_ERR_TOO_LARGE.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (two
 * contexts, or two host threads, with work in flight at once) race; calls
 * ordered on one stream do not.                                             */
int sloam_project_partition_noalloc(sloam_ctx* ctx, const float* d_xyzi, const uint8_t* d_label,
                                    int64_t n, const sloam_proj_cfg* cfg, int32_t* d_proj,
                                    float* d_range, int32_t* d_winner, uint8_t* d_mask,
                                    float* d_ground, int32_t* d_n_ground, float* d_tree);

/* ---- ground planes: binning, bottom-k selection, per-cell plane fit -------
 * Replaces: sloam::sloam::binGroundPoints (sloam/src/core/sloam.cpp:330-386)
 * and Plane::Plane / Plane::computeModel (sloam/src/objects/plane.cpp:3-17,
 * 96-128) as called by computeModels (sloam.cpp:391-400).
 * d_points: capacity points of 4 floats (x, y, z, i); the first *d_count are
 * used (d_count NULL: all capacity).  Cells are radii_bins x theta_bins,
 * r-major (cell = r * theta_bins + t), as computeModels iterates them.
 * Per cell: d_cell_info 4 int32 {valid, kept points, binned points, 0};
 * d_cell_model 8 doubles {a, b, c, d (unit normal, z >= 0), centroid x, y, z,
 * 0}; d_features num_features x 4 floats: the cell's first num_features kept
 * points (bottom points by increasing z when the cell was reduced, input order
 * otherwise).  A cell is valid iff it keeps >= num_features (and > 0) points.
 * Optional (both NULL or both set): d_kept_offsets [cells + 1] and
 * d_kept_index [capacity]: cell c's binned points occupy slots
 * [kept_offsets[c], kept_offsets[c+1]); its kept points' input indices come
 * first, in the order above, and the remaining slots hold -1 (the cells of
 * binGroundPoints, sloam.cpp:330-386).                                       */
typedef struct sloam_ground_cfg {
  int32_t radii_bins;       /* groundRadiiBins                               */
  int32_t theta_bins;       /* groundThetaBins                               */
  double min_dist;          /* minGroundLidarDist                            */
  double max_dist;          /* maxGroundLidarDist                            */
  double retain_thresh;     /* groundRetainThresh (fraction kept per cell)   */
  int32_t num_features;     /* numGroundFeatures                             */
  float origin[3];          /* binning origin (the reference bins at SE3())  */
} sloam_ground_cfg;

int sloam_ground_planes(sloam_ctx* ctx, const float* d_points, int32_t capacity,
                        const int32_t* d_count, const sloam_ground_cfg* cfg, int32_t* d_cell_info,
                        double* d_cell_model, float* d_features, int32_t* d_kept_offsets,
                        int32_t* d_kept_index);

/* ---- batched ground-plane RANSAC (point-level, fixed hypotheses) ----------
 * No counterpart in the reference (its planes are per-cell SVD fits); the
 * inlier predicate is Plane::distance(const PointT&) < thr
 * (sloam/src/objects/plane.cpp:136-151) for a unit normal.  Set s holds points
 * d_xyz[offsets[s] .. offsets[s+1]) (3 floats each).  Hypothesis h of set s:
 * Philox4x32-10(counter {h, s, stream_tag, 0}, key seed) draws 3 distinct
 * points, the plane through them in FP32 (relative to the set's bounding-box
 * centre); counts of |n.p + d| < thr; best = max count, lowest index; FP64
 * PCA refit over its inliers.  Outputs as sloam_cyl_plan_exec, with the model
 * record {a, b, c, d (unit normal, z >= 0), centroid[3], refitted inliers}. */
typedef struct sloam_plane_fit_cfg {
  int32_t n_hyp;
  float thr;
  uint64_t seed;
  int32_t stream_tag;
  int32_t reserved;
} sloam_plane_fit_cfg;

typedef struct sloam_plane_plan sloam_plane_plan;

int sloam_plane_plan_create(sloam_ctx* ctx, const int32_t* h_offsets, int32_t n_sets,
                            const sloam_plane_fit_cfg* cfg, sloam_plane_plan** out);
void sloam_plane_plan_destroy(sloam_plane_plan* plan);
int sloam_plane_plan_exec(sloam_plane_plan* plan, const float* d_xyz, const int32_t* d_offsets,
                          int32_t* d_counts, int32_t* d_samples, uint32_t* d_mask,
                          int32_t* d_best, double* d_model);

/* ---- batched RANSAC line fit (cylinder axis from trellis vertices) --------
 * Replaces: the PCL SACSegmentation call of Cylinder::computeModel
 * (sloam/src/objects/cylinder.cpp:116-139: SACMODEL_LINE, SAC_RANSAC,
 * threshold 0.25, optimizeCoefficients) for a batch of trees.  Set s holds
 * points d_xyz[offsets[s] .. offsets[s+1]) (3 floats each: the vertex medians).
 * Sample-consensus semantics of the library: MT19937 seeded `seed` (12345),
 * swap-shuffled index array persisting across draws, samples redrawn until
 * all three coordinates of the two points differ (up to 1000 draws), adaptive
 * bound k = log(1 - probability) / log(1 - w^2), at most max_iterations
 * (50); optional refinement through the inliers' centroid along their
 * principal axis.  Outputs per set: d_coeff 6 floats (point, unit direction),
 * d_stats 4 int32 {ok, inliers, iterations, skipped}; d_inliers (nullable)
 * one byte per point.  Sets hold at most 4096 points.                      */
typedef struct sloam_line_fit_cfg {
  double threshold;        /* 0.25                                           */
  double probability;      /* 0.99                                           */
  int32_t max_iterations;  /* 50                                             */
  int32_t optimize;        /* 1                                              */
  uint32_t seed;           /* 12345                                          */
  int32_t reserved;
} sloam_line_fit_cfg;

int sloam_fit_lines(sloam_ctx* ctx, const float* d_xyz, const int32_t* d_offsets, int32_t n_sets,
                    const sloam_line_fit_cfg* cfg, float* d_coeff, int32_t* d_stats,
                    uint8_t* d_inliers);


/* ---- tree instances: organized connected components + beam vertices -------
 * Replaces: Instance::findClusters, findTrees, computeTreeVertex and
 * computeVertexProperties (sloam/src/segmentation/trellis.cpp:15-132; the
 * clustering is PCL's OrganizedConnectedComponentSegmentation with a
 * EuclideanClusterComparator, threshold 1.0 m, not depth dependent).
 * d_cloud: organized H x W cloud, 4 floats per pixel (x, y, z, intensity),
 * row-major; a pixel takes part iff x is finite (the tree cloud of
 * sloam_project_partition).  Components are numbered by their first pixel in
 * row-major order.  Each component with more than min_label_points pixels
 * is a candidate tree; its rows from the last to the first with more than
 * min_row_points pixels give a vertex: medians of x, y, z (element
 * floor(n/2) of each sorted axis), the points within max_dist_to_centroid
 * of the median in (z, y, x, column) order, valid with at least 2 points,
 * radius = radius_scale x the distance of the first and last kept points.
 * Trees with more than min_vertices valid vertices are kept, cut to the
 * first max_vertices (bottom up).
 * Outputs (device; capacities in entries):
 *   d_labels        (nullable) H*W int32 component id, -1 for empty pixels
 *   d_tree_offsets  [cap_trees + 1] into vertices
 *   d_vertex_int    [cap_vertices][4] treeId (component), row, points, valid
 *   d_vertex_f64    [cap_vertices][4] radius, median x, y, z
 *   d_point_offsets [cap_vertices + 1] into points
 *   d_points        [cap_points][4] kept points (x, y, z, intensity)
 *   d_counts        8 int32: trees, vertices, points, components, overflow
 *                   (1 when a capacity was too small: outputs are then
 *                   truncated), beams, reserved...
 * Bounds: cap_trees <= H*W / (min_label_points + 1), cap_vertices <=
 * cap_trees * max_vertices, cap_points <= H*W always suffice.  H, W <=
 * 65535 and H*W < 2^20.                                                      */
typedef struct sloam_trellis_cfg {
  float cluster_dist;          /* 1.0 m (trellis.cpp:23)                      */
  float max_dist_to_centroid;  /* 0.2 m (sloamNode.cpp:62)                     */
  int32_t min_label_points;    /* 80  (trellis.cpp:109)                       */
  int32_t min_row_points;      /* 3   (trellis.cpp:119)                       */
  int32_t min_vertices;        /* 16  (trellis.cpp:124)                       */
  int32_t max_vertices;        /* 56  (trellis.cpp:125)                       */
  double radius_scale;         /* 1.0 (trellis.cpp:98); 0.5 reproduces the
                                  reference's test fixtures                   */
} sloam_trellis_cfg;

int sloam_trellis(sloam_ctx* ctx, const float* d_cloud, int32_t height, int32_t width,
                  const sloam_trellis_cfg* cfg, int32_t* d_labels, int32_t cap_trees,
                  int32_t cap_vertices, int32_t cap_points, int32_t* d_tree_offsets,
                  int32_t* d_vertex_int, double* d_vertex_f64, int32_t* d_point_offsets,
                  float* d_points, int32_t* d_counts);


/* ---- keyframe front end (one plan, optionally one HIP graph) --------------
 * Projection + partition (sloam_project_partition), tree instances
 * (sloam_trellis) and ground-plane cells (sloam_ground_planes) of one
 * organized height x width scan, on buffers the plan owns: the per-scan part
 * of Segmentation::run / maskCloud, Instance::computeGraph and
 * sloam::computeModels' ground cells (sloam/src/segmentation/inference.cpp:
 * 80-300, trellis.cpp:134-141, sloam.cpp:388-414).  With use_graph the first
 * exec runs eagerly (sizing the scratch buffers) and captures the sequence
 * into a HIP graph; every later exec is one graph launch.  The caller writes
 * the scan into out.xyzi / out.label (device) before each exec.  With
 * host_results the sequence ends with a kernel that writes the used part of
 * the tree and ground-cell outputs into a pinned host buffer
 * (sloam_frontend_host_view), so they are in host memory as soon as the exec's
 * work has completed, without further copies.                               */
typedef struct sloam_frontend sloam_frontend;
typedef struct sloam_frontend_cfg {
  sloam_proj_cfg proj;
  sloam_trellis_cfg trellis;
  sloam_ground_cfg ground;
  int32_t use_graph;
  int32_t host_results;
  /* fit_lines: after the tree instances, every tree's axis line is fitted to
   * its vertex medians (as float) with `line`, as sloam_fit_lines would fit
   * that tree alone (Cylinder::computeModel's sample consensus,
   * sloam/src/objects/cylinder.cpp:116-139)                                  */
  sloam_line_fit_cfg line;
  int32_t fit_lines;
  int32_t reserved;
} sloam_frontend_cfg;
typedef struct sloam_frontend_out { /* device pointers owned by the plan */
  float* xyzi;               /* in:  H*W x 4                               */
  uint8_t* label;            /* in:  H*W                                   */
  uint8_t* mask;             /* H*W label mask                             */
  float* ground;             /* H*W x 4 ground points, count in n_ground   */
  int32_t* n_ground;
  float* tree;               /* H*W x 4 dense tree cloud                   */
  int32_t cap_trees, cap_vertices, cap_points;
  int32_t* tree_offsets;     /* sloam_trellis outputs                      */
  int32_t* vertex_int;
  double* vertex_f64;
  int32_t* point_offsets;
  float* points;
  int32_t* trellis_counts;   /* 8 int32                                    */
  int32_t n_cells;
  int32_t* cell_info;        /* sloam_ground_planes outputs                */
  double* cell_model;
  float* cell_features;
  int32_t* kept_offsets;
  int32_t* kept_index;
  float* line_coeff;         /* fit_lines: cap_trees x 6 (sloam_fit_lines)  */
  int32_t* line_stats;       /* fit_lines: cap_trees x 4                    */
} sloam_frontend_out;

typedef struct sloam_frontend_host { /* host pointers into the plan's pinned buffer */
  const int32_t* trellis_counts;  /* 8 int32, as sloam_trellis' d_counts       */
  const int32_t* tree_offsets;    /* trellis_counts[0] + 1                       */
  const int32_t* vertex_int;      /* 4 x trellis_counts[1]                       */
  const double* vertex_f64;       /* 4 x trellis_counts[1]                       */
  const int32_t* point_offsets;   /* trellis_counts[1] + 1                       */
  const float* points;            /* 4 x trellis_counts[2]                       */
  const int32_t* cell_info;       /* 4 x n_cells                                 */
  const double* cell_model;       /* 8 x n_cells                                 */
  const float* cell_features;     /* 4 x n_cells x num_features                  */
  const float* line_coeff;        /* fit_lines: 6 x trellis_counts[0]            */
  const int32_t* line_stats;      /* fit_lines: 4 x trellis_counts[0]            */
} sloam_frontend_host;

int sloam_frontend_create(sloam_ctx* ctx, const sloam_frontend_cfg* cfg, sloam_frontend** out);
void sloam_frontend_destroy(sloam_frontend* f);
int sloam_frontend_buffers(const sloam_frontend* f, sloam_frontend_out* out);
int sloam_frontend_exec(sloam_frontend* f);
/* The front end's sequence enqueued on its context stream as eager
 * launches, never captured by the plan itself: for callers that capture a
 * larger sequence (a whole keyframe) into a graph of their own.  Run it once
 * before such a capture (it sizes the context's scratch).                   */
int sloam_frontend_enqueue(sloam_frontend* f);
/* Host view of the results of the last exec (valid once its work has
 * completed, e.g. after sloam_sync); SLOAM_ERR_INVALID_ARG unless the plan was
 * created with host_results.  Counts past the capacities (overflow flag
 * trellis_counts[4]) are clamped to the capacities in the copied arrays. */
int sloam_frontend_host_view(const sloam_frontend* f, sloam_frontend_host* out);

/* ---- keyframe estimator on the device ----------------------------------------
 * Replaces: sloam::sloam::RunSloam (sloam/src/core/sloam.cpp:453-532) with
 * computeModels (:388-436), matchFeatures / matchModels (:258-328),
 * TwoStepOptimizePose / OptimizePose (:33-255) and projectModels (:438-451),
 * for trees modelled by cylinder fits of their points (the KeyframeRunner's
 * plain and learned tree models).  The front end's device outputs, the tree
 * fits (sloam_cyl_plan model records, one per tree in the front end's tree
 * order), the pose estimate and the map trees go in; the pose, the
 * landmarks projected at it and their map matches come out, in pinned host
 * memory, without a host round trip between the front end and the pose.  The
 * estimator keeps the reference's per-process state (firstScan_,
 * prevGPlanes_) on the device; sloam_kf_estimator_reset starts over.  The
 * model arithmetic is the drop-in classes' (same functions), so a keyframe
 * gives the host RunSloam's decisions and pose.
 * Capacities: cap_trees landmarks, cap_cells ground cells, cap_map map trees;
 * beyond them the result header's overflow word is 1 and the outputs are
 * truncated.                                                                */
typedef struct sloam_kf_params {
  double max_lidar_dist;         /* maxLidarDist                              */
  double default_tree_radius;    /* defaultTreeRadius                         */
  double max_tree_radius;        /* maxTreeRadius                             */
  double max_axis_theta;         /* maxAxisTheta (degrees)                    */
  double tree_match_thresh;      /* treeMatchThresh                           */
  double add_new_tree_thresh_dist; /* AddNewTreeThreshDist                    */
  double min_tree_models;        /* minTreeModels                             */
  double min_ground_models;      /* minGroundModels                           */
  double min_planes;             /* radii bins x theta bins x 0.1 (sloam.h)   */
  int32_t features_per_tree;     /* featuresPerTree                           */
  int32_t num_ground_features;   /* numGroundFeatures                         */
  int32_t two_step;              /* twoStepOptim                              */
  int32_t pose_precision;        /* as sloam_pose_cfg.precision (1: FP64)     */
} sloam_kf_params;

typedef struct sloam_kf_inputs { /* device pointers */
  const int32_t* trellis_counts; /* sloam_frontend_out.trellis_counts ...     */
  const int32_t* tree_offsets;
  const int32_t* vertex_int;
  const double* vertex_f64;
  const int32_t* point_offsets;
  const float* points;
  int32_t n_cells;               /* ground cells (sloam_frontend_out.n_cells) */
  int32_t num_features;          /* must equal num_ground_features            */
  const int32_t* cell_info;
  const double* cell_model;
  const float* cell_features;
  const double* tree_fits;       /* 8 doubles per tree: root, ray, radius, inliers */
  const double* pose;            /* 7: t, q (x y z w), a unit quaternion       */
  const double* map;             /* cap_map x 8: root, ray, radius, 0          */
  const int32_t* n_map;          /* map trees in use                          */
} sloam_kf_inputs;

typedef struct sloam_kf_result_view { /* pinned host memory */
  /* 32 int32: mode (0 first scan, 1 estimated, 2 returned false), planes,
   * landmarks, tree check, ground check, tree feature matches, ground feature
   * matches, tree solve index, ground solve index, solves, overflow, trees,
   * matched landmarks, matched planes, cylinder rows, plane rows, success,
   * ran (1: outputs written), LM iterations of the tree (or joint) solve and
   * of the ground solve (-1: not run) */
  const int32_t* head;
  const double* pose;      /* 7: T_Map_Curr (t, q x y z w)                   */
  const double* delta;     /* 7: T_Delta                                     */
  const int32_t* lm_tree;  /* per landmark: the front end's tree index       */
  const int32_t* lm_id;    /* per landmark: tree id (vertex [2]'s treeId)    */
  const int32_t* lm_match; /* per landmark: matchModels' map index or -1     */
  const double* lm_model;  /* per landmark: 8 doubles, root, ray, radius, 0  */
} sloam_kf_result_view;

typedef struct sloam_kf_estimator sloam_kf_estimator;
int sloam_kf_estimator_create(sloam_ctx* ctx, const sloam_kf_params* params, int32_t cap_trees,
                              int32_t cap_cells, int32_t cap_map, sloam_kf_estimator** out);
void sloam_kf_estimator_destroy(sloam_kf_estimator* e);
int sloam_kf_estimator_reset(sloam_kf_estimator* e);
int sloam_kf_estimator_exec(sloam_kf_estimator* e, const sloam_kf_inputs* in);
/* Results of the last exec (valid once its work has completed).           */
int sloam_kf_estimator_result(const sloam_kf_estimator* e, sloam_kf_result_view* out);
/* The front end's trees as a cluster batch for the cylinder fits, on the
 * device: d_offsets (cap_trees + 1) relative to the first tree's first
 * point, d_xyz (cap_points x 3 floats) their points, *d_n the trees in use
 * (at most cap_trees; the points past cap_points are not copied, which a
 * counted plan of the same capacities reports as overflow).  No counterpart
 * in the reference, whose trees are PCL clouds.                            */
int sloam_kf_tree_points(sloam_ctx* ctx, const sloam_frontend_out* fe, int32_t cap_trees,
                         int64_t cap_points, int32_t* d_offsets, float* d_xyz, int32_t* d_n);

/* ---- learned cylinder proposal (PointNet, BF16 MFMA) -----------------------
 * No counterpart in the reference.  It proposes a cylinder per cluster, so that
 * a fit needs a few verifying hypotheses instead of a full RANSAC (the
 * predicted model's output space is CylinderParameters {root, ray, radius},
 * sloam/include/objects/cylinder.h:14-22).
 * Input per cluster: up to SLOAM_PN_POINTS points, taken at indices
 * floor(j n / m), m = min(n, 512), shifted by the centre of their bounding box,
 * x and y scaled by 1 / (larger horizontal half extent), z by 1 / (vertical
 * half extent).  Network: shared per-point MLP
 * 3 -> 64 -> 128 -> 256 (ReLU), max-pool over the cluster's points, MLP
 * 256 -> 128 (ReLU) -> 7 = (axis direction, axis point nearest the box
 * centre, radius) in the normalised frame.  Weights are BF16 (rounded from
 * the FP32 master copy), layers 2 and 3 run on v_mfma_f32_16x16x32_bf16 with
 * FP32 accumulation, activations between them are BF16.
 * Parameters: one FP32 array, PyTorch order (weight [out][in], then bias):
 * W1 [64][3], b1, W2 [128][64], b2, W3 [256][128], b3, W4 [128][256], b4,
 * W5 [7][128], b5: SLOAM_PN_PARAMS floats, 16-byte aligned (the entry points
 * taking d_params return SLOAM_ERR_INVALID_ARG otherwise).                  */
#define SLOAM_PN_POINTS 512
#define SLOAM_PN_PARAMS 75399
#define SLOAM_PN_PROPOSAL_STRIDE 8

/* Proposals, one record of 8 floats per cluster (world frame): axis[3]
 * (unit), centre[3] (a point on the axis), radius, valid (1 / 0).          */
int sloam_pn_infer(sloam_ctx* ctx, const float* d_params, const float* d_xyz,
                   const int32_t* d_offsets, int32_t n_clusters, float* d_proposal);
/* The same for *d_n_clusters clusters (device memory, at most cap_clusters):
 * no host-side size, for graph capture (no counterpart in the reference).  */
int sloam_pn_infer_counted(sloam_ctx* ctx, const float* d_params, const float* d_xyz,
                           const int32_t* d_offsets, int32_t cap_clusters,
                           const int32_t* d_n_clusters, float* d_proposal);

/* Training loss against cylinder models (SLOAM_CYL_MODEL_STRIDE doubles per
 * cluster, e.g. the plain fit's output; clusters whose radius is <= 0 carry
 * no loss):  L_c = w_axis (1 - (a_hat . a)^2) + w_centre |c_hat - c*|^2 +
 * w_radius (r_hat - r*)^2 in the normalised frame (a: the target axis mapped
 * into it; c*: the point of that axis line nearest the box centre; r*: the
 * radius at the horizontal scale), L = mean over the batch.                 */
typedef struct sloam_pn_train_cfg {
  float lr;          /* Adam step size                                        */
  float beta1;       /* 0.9                                                   */
  float beta2;       /* 0.999                                                 */
  float eps;         /* 1e-8                                                  */
  float w_axis;      /* loss weights                                          */
  float w_centre;
  float w_radius;
  int32_t reserved;
} sloam_pn_train_cfg;

/* Loss per cluster (d_loss, n_clusters floats) and the batch-mean gradient
 * (d_grad, SLOAM_PN_PARAMS floats); parameters unchanged.                   */
int sloam_pn_loss_grad(sloam_ctx* ctx, const float* d_params, const float* d_xyz,
                       const int32_t* d_offsets, int32_t n_clusters, const double* d_targets,
                       const sloam_pn_train_cfg* cfg, float* d_loss, float* d_grad);

/* One Adam step (step = 1, 2, ... for the bias correction) on d_params with
 * moments d_m, d_v (SLOAM_PN_PARAMS floats each); d_loss (nullable) gets the
 * per-cluster losses before the step.                                       */
int sloam_pn_train_step(sloam_ctx* ctx, float* d_params, float* d_m, float* d_v, int32_t step,
                        const sloam_pn_train_cfg* cfg, const float* d_xyz,
                        const int32_t* d_offsets, int32_t n_clusters, const double* d_targets,
                        float* d_loss);

#ifdef __cplusplus
}
#endif
#endif /* SLOAM_GPU_H_ */
