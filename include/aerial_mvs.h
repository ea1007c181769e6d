/*
 * aerial_mvs — MI355X (gfx950) dense multi-view-stereo depth engine, C ABI.
 *
 * Drop-in for the PatchMatch depth-map stage of stitch-core
 * (crates/stitch-core/src/recon/dense.rs). The reference has no FFI of its own:
 * its seam is the private Rust function
 *     fn depth_map(dir, sp, ci, prm) -> Result<Option<(usize, usize, Vec<f32>)>>
 * (dense.rs:315-321), called once per view from reconstruct_dense
 * (dense.rs:530-540), plus the `.view` / `.depth` cache files (dense.rs:88-154,
 * 311-313, 481-508). Every entry point below names the reference code it
 * replaces. Plain C types only; all pointers are host pointers unless stated.
 *
 * Conventions (geometry.rs:1-5): world = local east/north/up metres, camera
 * x right / y down / z forward, R is world -> camera (row-major), a world point
 * X lands at R (X - C). Intrinsics are in the view's own pixels (already scaled
 * to the working size, dense.rs:138-145).
 *
 * Return codes: MVS_OK (0); MVS_NO_NEIGHBOURS (1) is the reference's Ok(None)
 * (dense.rs:335-337); negative values are errors whose text mvs_last_error()
 * returns (the reference's Err). There is no CPU fallback: without a usable
 * gfx950 device every compute entry point fails with MVS_ERR_DEVICE.
 *
 * Threading: an mvs_ctx owns one HIP stream and its device buffers; use it
 * from one thread at a time (the reference runs depth maps one view at a time,
 * dense.rs:528-530). Different contexts may run concurrently.
 */
#ifndef AERIAL_MVS_H
#define AERIAL_MVS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MVS_ABI_VERSION 1

enum {
    MVS_OK = 0,
    MVS_NO_NEIGHBOURS = 1,
    MVS_ERR_ARG = -1,
    MVS_ERR_DEVICE = -2,
    MVS_ERR_CANCELLED = -3,
    MVS_ERR_NOMEM = -4,
    MVS_ERR_IO = -5,
    MVS_ERR_FORMAT = -6
};

/* Pixel formats of a view's grey plane. U8 is converted with the reference's
 * grey formula for R = G = B (dense.rs:108); F32 holds the reference's own grey
 * values (the `.view` cache, dense.rs:104-111) and is used unchanged. */
enum { MVS_GRAY_U8 = 0, MVS_GRAY_F32 = 1 };

/* One working-size view (the reference's View + Camera, dense.rs:57-65,
 * sparse.rs:26-38, geometry.rs:13-53). */
typedef struct {
    int32_t w, h;
    int32_t stride;          /* pixels between rows of `pixels` (>= w) */
    int32_t format;          /* MVS_GRAY_U8 or MVS_GRAY_F32 */
    const void *pixels;      /* host memory, borrowed for the call */
    double f, cx, cy, k1, k2;
    double R[9];             /* world -> camera, row-major */
    double C[3];             /* camera centre */
} mvs_view;

/* Knobs of the depth stage. mvs_default_params() gives the reference's
 * constants: 5 taps at stride 2 (dense.rs:170), TEXTURE_MIN 0.008 (:173),
 * output threshold 0.45 (:471), Standard-profile iterations 4 (:41-46). */
typedef struct {
    int32_t taps;            /* window taps per side: 5 (reference), 7 or 11 */
    int32_t stride;          /* pixels between taps (reference 2) */
    int32_t iterations;      /* PatchMatch iterations */
    float texture_min;
    float cost_max;
    int32_t refine_steps;    /* Gauss-Newton refinement steps after PatchMatch
                              * (0 = none, the reference's behaviour; no
                              * reference counterpart) */
    int32_t mode;            /* MVS_MODE_EXACT (0): PatchMatch rounds operation for
                              * operation like the reference (bit-identical maps);
                              * MVS_MODE_FAST (1): contracted multiply-adds and
                              * hardware reciprocals, costs within 2e-5 of exact */
    int32_t reserved[1];     /* zero */
} mvs_params;

enum { MVS_MODE_EXACT = 0, MVS_MODE_FAST = 1 };

void mvs_default_params(mvs_params *p);

/* DenseParams::for_profile (dense.rs:32-55). kind: 0 Lite, 1 Standard, 2 High.
 * out = {size, iterations, neighbours, max_points}. Returns MVS_ERR_ARG for an
 * unknown kind. */
int mvs_dense_params_for_profile(int32_t kind, int64_t out[4]);

/* Neighbour choice of depth_map (dense.rs:324-334): keep cam.neighbours in
 * order whose baseline |C_j - C_ci| > 0.02 max(range0, 1), first `take` of them.
 * centres: ncams x 3 doubles. Writes indices to out, returns how many. */
int32_t mvs_select_neighbours(const double *centres, int32_t ncams, int32_t ci,
                              const int32_t *neighbours, int32_t nn, double range0,
                              int32_t take, int32_t *out);

typedef struct mvs_ctx mvs_ctx;

/* Create a context on HIP device `device`. On failure returns a negative code
 * and writes a message into err (if given). */
int mvs_create(int32_t device, mvs_ctx **out, char *err, size_t errlen);
void mvs_destroy(mvs_ctx *ctx);
const char *mvs_last_error(const mvs_ctx *ctx);
int mvs_synchronize(mvs_ctx *ctx);

/* Progress callback: stage "depth" per finished view, like Ctx::report
 * (dense.rs:539). Return nonzero to cancel. */
typedef int (*mvs_progress_fn)(void *user, const char *stage, int32_t done, int32_t total);

/* ---- one depth map from host views: replaces depth_map (dense.rs:316-479) ----
 * ref + nsrc sources (the already-filtered neighbours, dense.rs:324-341),
 * depth range = cam.depth_range (dense.rs:362). Outputs are caller-allocated
 * host arrays of w*h (normal: 3*w*h, xyz per pixel); out_normal / out_cost may
 * be NULL. out_depth follows dense.rs:468-477 (0 = invalid). cancel (nullable)
 * is polled between half-sweeps. progress (nullable) is called with stage
 * "patchmatch" before each half-sweep is enqueued (done = half-sweeps enqueued,
 * total = 2 x iterations) and once more with done = total when the maps are in
 * the output arrays; a nonzero return cancels (MVS_ERR_CANCELLED), like the
 * reference's per-view Ctx::check / Ctx::report pair (dense.rs:531, 539). */
int mvs_depth_map(mvs_ctx *ctx, const mvs_view *ref, const mvs_view *srcs, int32_t nsrc,
                  double range0, double range1, const mvs_params *prm, float *out_depth,
                  float *out_normal, float *out_cost, const volatile int32_t *cancel,
                  mvs_progress_fn progress, void *user);

/* ---- parity entry points (lock-step driver) ----
 * Cost of explicit (pixel, hypothesis) pairs: Ctx2::cost (dense.rs:181-263).
 * hyp4: n x {d, n0, n1, n2}. */
int mvs_ncc_cost(mvs_ctx *ctx, const mvs_view *ref, const mvs_view *srcs, int32_t nsrc,
                 double range0, double range1, const mvs_params *prm, int64_t n,
                 const int32_t *px, const int32_t *py, const float *hyp4, float *out_cost);

/* Fixed-pair NCC cost and its gradient: the forward and backward pass of the
 * refinement (no reference counterpart; the cost is Ctx2::cost's best-two
 * average, dense.rs:255-262, with the two sources given).
 * cost_i = 1 - (NCC_s0 + NCC_s1) / 2 for item i's plane hyp4[i] = {d, n0, n1,
 * n2} at pixel (px[i], py[i]) with sources pairs[2i], pairs[2i+1]; an item
 * whose pair is (-1, -1) gets its two best sources at that plane (ties to the
 * lower index), written back to pairs (-1, -1 stays when fewer than two
 * sources are defined). grad4 (nullable): n x 4 floats, grad_out[i] (nullable
 * = 1) times d cost_i / d(d, n0, n1, n2) with n free (not renormalised).
 * Undefined items get cost 2 and gradient 0. */
int mvs_ncc_pair(mvs_ctx *ctx, const mvs_view *ref, const mvs_view *srcs, int32_t nsrc,
                 double range0, double range1, const mvs_params *prm, int64_t n,
                 const int32_t *px, const int32_t *py, const float *hyp4, int32_t *pairs,
                 const float *grad_out, float *out_cost, float *grad4);

/* The Gauss-Newton terms of the refinement (no reference counterpart; the
 * cost is Ctx2::cost's best-two average, dense.rs:255-262, with the two
 * sources given) for explicit items, computed by the device code the
 * refinement itself runs on this view class: one lane per pixel for 8-bit
 * views on undistorted cameras, one lane per source otherwise (mvs_get_stats
 * reports which as split -1 / 1). Items and pairs as mvs_ncc_pair. Per item:
 * out_cost = 1 - (NCC_s0 + NCC_s1) / 2; grad4 (nullable, n x 4) = d cost /
 * d(d, n0, n1, n2); jtj10 (nullable, n x 10) = J^T J of the pair's stacked
 * residuals r_s = a/|a| - b_s/|b_s| (a, b_s the centred reference and sample
 * windows, 1 - NCC_s = |r_s|^2 / 2) w.r.t. (d, n0, n1, n2), packed upper
 * row-major (00 01 02 03 11 12 13 22 23 33): the matrix the Gauss-Newton
 * step solves with. The kernels work in depth and two tangent angles of the
 * normal; both outputs are mapped to (d, n) through that basis (the residuals
 * do not change along n). Undefined items: cost 2, zeros. */
int mvs_refine_terms(mvs_ctx *ctx, const mvs_view *ref, const mvs_view *srcs, int32_t nsrc,
                     double range0, double range1, const mvs_params *prm, int64_t n,
                     const int32_t *px, const int32_t *py, const float *hyp4, int32_t *pairs,
                     float *out_cost, float *grad4, float *jtj10);

/* Run PatchMatch in steps on a host state of w*h x {d, n0, n1, n2, cost}
 * (Hyp, dense.rs:156-161). first_step = -1 runs the initialisation
 * (dense.rs:365-402) before the sweeps; half-sweep s covers iteration s/2,
 * parity s%2 (dense.rs:413-467). textured (w*h, nullable) receives the mask.
 * evals (nullable) receives the number of cost evaluations performed. */
int mvs_patchmatch_steps(mvs_ctx *ctx, const mvs_view *ref, const mvs_view *srcs, int32_t nsrc,
                         double range0, double range1, const mvs_params *prm,
                         int32_t first_step, int32_t n_steps, float *state, uint8_t *textured,
                         int64_t *evals);

/* ---- resident survey: views uploaded once, depth maps kept in HBM ---- */
typedef struct mvs_survey mvs_survey;

int mvs_survey_create(mvs_ctx *ctx, int32_t nviews, mvs_survey **out);
void mvs_survey_destroy(mvs_survey *sv);
/* Upload view i (pixels converted to device grey once). */
int mvs_survey_set_view(mvs_survey *sv, int32_t i, const mvs_view *v);
/* Residency options of a survey (no reference counterpart: they let a survey
 * of thousands of views stay in one GPU's memory). They apply to views set
 * and maps computed after the call.
 * MVS_SURVEY_LEAN_VIEWS: an 8-bit view keeps its f32 grey and its 8-bit and
 *   f16 coefficient planes only (16 bytes per pixel instead of 32); jobs that
 *   touch such a view must run in fast mode on undistorted views (the
 *   one-lane-per-pixel kernels), others fail with MVS_ERR_ARG.
 * MVS_SURVEY_DEPTH_ONLY: maps keep the depth only (4 bytes per pixel instead
 *   of 20); downloading their normals or costs fails with MVS_ERR_ARG. */
#define MVS_SURVEY_LEAN_VIEWS 1
#define MVS_SURVEY_DEPTH_ONLY 2
int mvs_survey_set_options(mvs_survey *sv, int32_t options);

/* One reference view of the survey (the body of the loop dense.rs:530-540,
 * without file I/O). Neighbour indices are survey view indices, already
 * filtered. The maps stay on the device until mvs_survey_download. */
int mvs_survey_depth_map(mvs_survey *sv, int32_t ref, const int32_t *srcs, int32_t nsrc,
                         double range0, double range1, const mvs_params *prm);

/* Many reference views in one call; jobs are described in CSR form:
 * refs[j], srcs[src_off[j] .. src_off[j+1]), ranges[2j], ranges[2j+1].
 * Progress/cancel per view as in dense.rs:531-539. */
int mvs_survey_depth_maps(mvs_survey *sv, int32_t njobs, const int32_t *refs,
                          const int32_t *src_off, const int32_t *srcs, const double *ranges,
                          const mvs_params *prm, const volatile int32_t *cancel,
                          mvs_progress_fn progress, void *user);

/* As mvs_survey_depth_maps with parameters per job (prms[j]): window size,
 * stride, iterations, thresholds, refinement and mode may differ between the
 * views of one batch (no reference counterpart; the reference runs one
 * DenseParams per survey). Jobs are grouped into kernel classes (window,
 * iterations, neighbour-count bucket, projection, mode); each half-sweep is
 * one launch per class over all its views. */
int mvs_survey_depth_maps_ex(mvs_survey *sv, int32_t njobs, const int32_t *refs,
                             const int32_t *src_off, const int32_t *srcs, const double *ranges,
                             const mvs_params *prms, const volatile int32_t *cancel,
                             mvs_progress_fn progress, void *user);

int mvs_survey_download(mvs_survey *sv, int32_t view, float *depth, float *normal, float *cost);
/* A depth map computed earlier (the `.depth` cache, dense.rs:533-536) as view
 * `view`'s map, for consistency and fusion: w*h f32 in the view's size; its
 * normals and costs read as zero. */
int mvs_survey_set_depth(mvs_survey *sv, int32_t view, const float *depth);

/* Geometric consistency of view ref's depth map against its neighbours' maps
 * (fusion filter, dense.rs:575-615): neighbours = the first prm.neighbours+1 of
 * cam.neighbours that have a depth map (the caller passes that list). rule
 * MVS_CONS_REFERENCE: relative depth difference < 1% in >= 2 maps (>= 1 when
 * fewer than 3 neighbours). rule MVS_CONS_REPROJ (no reference counterpart):
 * the neighbour's own depth at the pixel the ref point lands on, back-projected
 * from that pixel's centre, reprojects within 1 px of the ref pixel centre, in
 * >= 3 maps. Writes per-pixel agree counts (-1 where depth <= 0) and returns the
 * count of consistent pixels via n_consistent. */
enum { MVS_CONS_REFERENCE = 0, MVS_CONS_REPROJ = 1 };
int mvs_survey_consistency(mvs_survey *sv, int32_t ref, const int32_t *nbrs, int32_t nn,
                           int32_t rule, int32_t *agree, int64_t *n_consistent);

/* mvs_survey_consistency for callers that cannot allocate device memory (no
 * reference counterpart; for code under HIP graph capture, tight serving
 * loops and harnesses that forbid hipMalloc). Same arguments, checks and
 * results; it calls no device allocator. Its scratch (the neighbour cameras,
 * the per-pixel agree counts and the counter) lives in one fixed device array
 * of about 32 MiB per GPU, set up by mvs_create and held until the process
 * ends: reference views above MVS_NOALLOC_MAX_PIXELS pixels or more than
 * MVS_NOALLOC_MAX_NBRS neighbours fail with MVS_ERR_ARG. Every call on a GPU
 * shares that GPU's array, so calls that overlap on one GPU (from different
 * contexts or threads) race; run them one at a time per GPU. */
#define MVS_NOALLOC_MAX_PIXELS (1 << 23)
#define MVS_NOALLOC_MAX_NBRS 16
int mvs_survey_consistency_noalloc(mvs_survey *sv, int32_t ref, const int32_t *nbrs, int32_t nn,
                                   int32_t rule, int32_t *agree, int64_t *n_consistent);

/* ---- fusion: consistent depths accumulated per voxel (dense.rs:542-651) ----
 * Colour plane of a survey view for the fused cloud (w*h*3 RGB, the `.view`
 * cache's rgb); views without one use their 8-bit grey. */
int mvs_survey_set_colour(mvs_survey *sv, int32_t view, const uint8_t *rgb);
/* Fuses the maps of refs[r] (views without a map are skipped, dense.rs:571):
 * every pixel that the neighbours nbrs[nbr_off[r] .. nbr_off[r+1]) with maps
 * confirm under `rule` (MVS_CONS_REFERENCE: dense.rs:585-612;
 * MVS_CONS_REPROJ) adds its world point and colour to voxel floor(X/voxel)
 * (dense.rs:617-628). Sums are exact fixed point, so the cloud does not depend
 * on the order views or pixels arrive in (the reference's f32 HashMap merge
 * does). n_voxels receives the number of voxels. */
int mvs_survey_fuse(mvs_survey *sv, int32_t nrefs, const int32_t *refs, const int32_t *nbr_off,
                    const int32_t *nbrs, int32_t rule, double voxel, int64_t *n_voxels);
/* The last fusion's cloud, one point per voxel in voxel-index order: mean
 * position (f32) and mean colour truncated to u8 (dense.rs:634-641). With
 * xyz and rgb NULL only n is written; cap = points the arrays hold. */
int mvs_survey_fused_points(mvs_survey *sv, float *xyz, uint8_t *rgb, int64_t cap, int64_t *n);
/* The fusion voxel edge (dense.rs:545-562): sparse point plan extent (1%-99%
 * percentiles of xs, ys), median of the cameras' mid depths (ranges: 2 per
 * camera), focal f of the first camera at full size width x height, working
 * size `size` and the profile's max_points. */
int mvs_fusion_voxel(const double *xs, const double *ys, int64_t npts, const double *ranges,
                     int32_t ncams, double f, int32_t size, int32_t width, int32_t height,
                     int64_t max_points, double *voxel);
/* Binary PLY as write_ply / read_ply (ply.rs:11-70): header, then per point
 * 3 x f32 LE and 3 x u8; written to the ".ply.tmp" sibling and renamed. */
int mvs_write_ply(const char *path, const float *xyz, const uint8_t *rgb, int64_t n,
                  const char *comment);
int mvs_read_ply(const char *path, float *xyz, uint8_t *rgb, int64_t cap, int64_t *n);

/* ---- timing of the hot kernels (HIP events on the context's stream) ---- */
typedef struct {
    int64_t sweep_launches;   /* half-sweeps (every kernel class of a batch) */
    double sweep_ms;          /* summed device time of the half-sweep kernels */
    int64_t init_launches;
    double init_ms;
    int64_t evals;            /* cost evaluations counted on the device */
    int64_t pixels;           /* reference pixels processed */
    int64_t init_evals;       /* of which in the initialisation (0 where the
                               * first half-sweep of each colour scores the
                               * initial planes: fast arithmetic on one lane
                               * per pixel, calls of at least two half-sweeps) */
    int64_t sweep_evals;      /* of which in the half-sweeps */
    int32_t split;            /* schedule of the last batch or cost list: lanes
                               * per source (-1: one lane per pixel, f16 coefficient
                               * plane; -2: one lane per pixel, 8-bit plane;
                               * -3: one lane per pixel, f32 plane) */
    int32_t gpc;              /* and pixel groups per work chunk */
    int64_t refine_launches;  /* Gauss-Newton refinement kernel */
    double refine_ms;
    int64_t refine_pixels;    /* pixels refined (PatchMatch cost below cost_max) */
    int64_t refine_accepted;  /* accepted refinement steps */
    int64_t graph_captures;   /* half-sweep graphs captured */
    int64_t graph_replays;    /* half-sweeps run as a graph launch */
    int64_t fuse_passes;      /* passes of mvs_survey_fuse over its views (a
                               * pass that overflows the voxel table is
                               * repeated with a larger one) */
    int64_t sweep_kernels;    /* half-sweep kernel launches: one per kernel
                               * class (window, projection, mode, and for the
                               * source-lane kernels the neighbour-count
                               * bucket) per half-sweep of a batch */
    int64_t sweep_skipped;    /* propagation candidates the one-lane-per-pixel
                               * half-sweeps did not score: the neighbour's
                               * last half-sweep left it unchanged, so this
                               * pixel scored it two half-sweeps before and did
                               * not take it (or took it and has it or a better
                               * one): it cannot be taken now, and the state
                               * is what scoring it would leave */
} mvs_stats;

int mvs_set_timing(mvs_ctx *ctx, int32_t enable);

/* Kernel schedule (no reference counterpart; with MVS_MODE_EXACT results do
 * not depend on it, with MVS_MODE_FAST they vary within the fast tolerance):
 * split = lanes sharing one source's window (0 = choose automatically; -1 =
 * one lane per reference pixel, for MVS_MODE_FAST with undistorted 8-bit
 * reference and source views, sampling an f16 coefficient plane; -2 = the same
 * sampling the 8-bit corner plane; -3 = the same for undistorted f32 grey
 * views, sampling the f32 coefficient plane; -4 = the same sampling
 * the f32 grey plane itself; a class the forced split does not apply to
 * takes its first candidate), gpc = pixel
 * groups per work chunk (0 = default), per_cu = persistent workgroups per
 * compute unit (0 = one workgroup per chunk). autotune = 1 times the candidate
 * schedules per kernel class and iteration on up to four views of the first
 * batch that has the class (twice each) and keeps the fastest. */
int mvs_set_schedule(mvs_ctx *ctx, int32_t split, int32_t gpc, int32_t per_cu,
                     int32_t autotune);
/* Half-sweep graphs (no reference counterpart; results do not depend on it):
 * with enable = 1 (the default) every half-sweep whose schedules are known
 * runs as a hipGraph of its launches, captured once and replayed by later
 * half-sweeps and calls with the same launch arguments. 0 launches directly
 * and frees the cached graphs. */
int mvs_set_graphs(mvs_ctx *ctx, int32_t enable);
/* The autotuner's choices so far, 9 int32 per entry: taps, source lanes,
 * pinhole (0 some camera distorted, 1 undistorted sources, 2 undistorted
 * sources and reference), log2 pixels, iteration, mode, split, gpc, per_cu.
 * Returns the entry count (writes at most cap entries). */
int32_t mvs_tuned_schedules(mvs_ctx *ctx, int32_t *out, int32_t cap);
int mvs_get_stats(mvs_ctx *ctx, mvs_stats *out);
int mvs_reset_stats(mvs_ctx *ctx);

/* ---- device self-checks (no reference counterpart) ----
 * MVS_SELFTEST_DIVISION: the kernels' shared-reciprocal division against IEEE
 * f32 division on n hashed operand triples; *mismatches = differing quotients
 * (0 expected). Lets a deployment confirm the arithmetic contract on its own
 * device and driver. */
enum { MVS_SELFTEST_DIVISION = 0 };
int mvs_selftest(mvs_ctx *ctx, int32_t which, int64_t n, uint64_t seed, int64_t *mismatches);

/* ---- `.depth` cache files (dense.rs:481-508) ---- */
/* Writes u32 LE w, u32 LE h, w*h f32 LE to path + ".tmp", then renames. */
int mvs_save_depth(const char *path, int32_t w, int32_t h, const float *depth);
/* Reads a `.depth` file; returns MVS_ERR_IO / MVS_ERR_FORMAT like load_depth's
 * None (wrong length). With depth == NULL only w and h are read. */
int mvs_load_depth(const char *path, int32_t *w, int32_t *h, float *depth, int64_t cap);

/* ---- `.view` cache files (dense.rs:92-154) ---- */
/* Writes u32 LE w, u32 LE h, w*h f32 LE grey, w*h*3 u8 RGB to path's ".tmp"
 * sibling, then renames (prepare_views, dense.rs:101-113). */
int mvs_save_view(const char *path, int32_t w, int32_t h, const float *gray, const uint8_t *rgb);
/* Reads a `.view` file (load_view, dense.rs:123-137): MVS_ERR_FORMAT unless the
 * length is exactly 8 + 7 w h ("damaged cache file"). gray (w*h) and rgb
 * (3*w*h) are nullable; with both NULL only w and h are read. cap = pixels. */
int mvs_load_view(const char *path, int32_t *w, int32_t *h, float *gray, uint8_t *rgb,
                  int64_t cap);
/* The working-size intrinsics of a view (dense.rs:139-145): f, cx, cy scaled
 * by w / full_width (k1, k2 unchanged). out = {f, cx, cy}. */
int mvs_view_intrinsics(double f, double cx, double cy, int32_t full_width, int32_t w,
                        double out[3]);

/* ---- surface model (recon/dsm.rs) ---- */
typedef struct {
    double x0, y0;   /* world coordinate of the south-west corner of cell (0, 0) */
    double cell;     /* cell edge, m */
    int64_t nx, ny;  /* grid; row 0 is the southernmost */
    float zmax;      /* highest height (-FLT_MAX without surface; NaN until built) */
    int32_t reserved;
} mvs_dsm_info;
/* build_dsm (dsm.rs:128-290) on the GPU from a cloud of n points (xyz f32, as
 * read from dense.ply; 100 <= n < 2^31), the fusion's voxel edge and the
 * profile (0 Lite, 1 Standard, 2 High: a budget of 3, 8, 20 M cells,
 * dsm.rs:121-128). Replaces the reference's build_dsm(dense_ply, voxel_m,
 * profile), which reads the cloud itself. The grid is built on the device and
 * kept there (until the next build on this context); info receives its
 * extent and zmax. With h != NULL the heights are also copied into h (cap
 * floats, at least nx * ny), row-major from the south, NaN = no surface;
 * with h == NULL, mvs_dsm_download copies them after the caller has sized
 * its array from info. MVS_ERR_ARG with "the point cloud is too small to
 * build a surface" below 100 points (dsm.rs:137-139). */
int mvs_build_dsm(mvs_ctx *ctx, const float *xyz, int64_t n, double voxel_m, int32_t profile,
                  mvs_dsm_info *info, float *h, int64_t cap);
/* The heights of the last grid mvs_build_dsm built on this context into h
 * (cap floats, at least nx * ny). Part of the same replacement. */
int mvs_dsm_download(mvs_ctx *ctx, float *h, int64_t cap);
/* Dsm::save / Dsm::load (dsm.rs:27-75): the `dsm.bin` layout ("DSM1", x0, y0,
 * cell f64, nx, ny u64, zmax f32, heights f32, little-endian; tmp + rename).
 * load: MVS_ERR_FORMAT for a damaged or truncated file; h (cap floats)
 * nullable to read the header only. */
int mvs_save_dsm(const char *path, const mvs_dsm_info *info, const float *h);
int mvs_load_dsm(const char *path, mvs_dsm_info *info, float *h, int64_t cap);
/* Dsm::height (dsm.rs:82-119): 1 with *out, 0 for None (outside), or an
 * error code. Host-only. */
int mvs_dsm_height(const mvs_dsm_info *info, const float *h, double x, double y, float *out);

/* ---- view preparation (dense.rs:92-121, imgops.rs:89-171) ---- */
/* shrink_to's output size for a w x h photo and working size max_side
 * (imgops.rs:145-171): unchanged when the long side fits, else an integer box
 * downsample by long / max_side (when >= 2) and a bilinear resize to the exact
 * size; scale = max_side / long (1 when unchanged). Host-only, no context. */
int mvs_prepared_size(int32_t w, int32_t h, int32_t max_side, int32_t *out_w, int32_t *out_h,
                      double *scale);
/* The per-photo work of prepare_views after decoding (dense.rs:100-112) on
 * the GPU: rgb is the decoded RGB8 photo (w x h, stride bytes per row >= 3 w),
 * max_side the profile's working size (DenseParams.size). Writes the
 * working-size grey (out_w x out_h f32, dense.rs:108) and RGB8 (3 out_w out_h
 * bytes) with out_w, out_h from mvs_prepared_size — the contents of a `.view`
 * file (mvs_save_view). Byte- and bit-identical to the reference's shrink_to
 * and grey formula. */
int mvs_prepare_view(mvs_ctx *ctx, const uint8_t *rgb, int32_t w, int32_t h, int32_t stride,
                     int32_t max_side, float *gray, uint8_t *rgb_out);

#ifdef __cplusplus
}
#endif
#endif
