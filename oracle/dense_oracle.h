/*
 * CPU oracle for the dense-MVS depth-map stage.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference's
 * PatchMatch depth-map algorithm (crates/stitch-core/src/recon/dense.rs) and of
 * the camera model it uses (crates/stitch-core/src/recon/geometry.rs). It is the
 * checker the GPU engine is compared against and the timed CPU baseline; it is
 * never linked into, loaded by or called from the product library.
 *
 * Arithmetic follows the reference operation by operation: f32 where the Rust
 * code uses f32, f64 where it uses f64, left-to-right evaluation, no fused
 * multiply-add (build with -ffp-contract=off, no fast-math).
 *
 * Generalisations beyond the reference (defaults reproduce it exactly):
 *   - window: `taps` per side at spacing `stride` (reference: 5 taps, stride 2,
 *     dense.rs:170);
 *   - any number of sources K >= 1 (reference: at most 8, dense.rs:217).
 */
#ifndef AERIAL_DENSE_ORACLE_H
#define AERIAL_DENSE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_TAPS 121 /* 11 x 11 */

/* A working-size view: grey plane plus camera (dense.rs:57-65, 123-154). */
typedef struct {
    int32_t w, h;
    const float *gray;          /* w*h grey levels, 0..1 (dense.rs:108) */
    double f, cx, cy, k1, k2;   /* intrinsics already scaled to w (dense.rs:139-145) */
    double R[9];                /* world -> camera rotation, row-major (geometry.rs:47-53) */
    double C[3];                /* camera centre, world (local ENU) metres */
} orc_view;

typedef struct {
    int32_t taps;        /* window taps per side, odd (reference 5) */
    int32_t stride;      /* tap spacing in pixels (reference 2) */
    int32_t iterations;  /* PatchMatch iterations (DenseParams.iterations) */
    float texture_min;   /* dense.rs:173 (0.008) */
    float cost_max;      /* dense.rs:471 (0.45) */
} orc_params;

typedef struct orc_ctx orc_ctx;

/* ---- scalar building blocks (each cites the reference line it restates) ---- */
void orc_gray_lut(float lut[256]);
float orc_gray_rgb(uint8_t r, uint8_t g, uint8_t b);
float orc_rnd(uint64_t *s);
uint64_t orc_seed_init(uint64_t i);
uint64_t orc_seed_refine(uint64_t i, uint64_t it, uint64_t parity);
void orc_random_normal(uint64_t *s, const float ray[2], const float around[3], float spread,
                       float out[3]);
void orc_distort(double f, double cx, double cy, double k1, double k2, double x, double y,
                 double out[2]);
void orc_undistort(double f, double cx, double cy, double k1, double k2, double u, double v,
                   double out[2]);
void orc_relative(const double Rr[9], const double Cr[3], const double Rs[9], const double Cs[3],
                  float R[9], float t[3]);
int orc_gray_at(const float *gray, int32_t w, int32_t h, float x, float y, float *out);
void orc_mean_sd(const float *v, int32_t n, float *m, float *sd);
int orc_select_neighbours(const double *centres, int32_t ncams, int32_t ci,
                          const int32_t *neighbours, int32_t nn, double range0, int32_t take,
                          int32_t *out);

/* ---- depth-map context ---- */
orc_ctx *orc_ctx_create(const orc_view *ref, const orc_view *srcs, int32_t nsrc,
                        const orc_params *prm, double range0, double range1);
void orc_ctx_destroy(orc_ctx *c);
const float *orc_ctx_rays(const orc_ctx *c);
void orc_ctx_relpose(const orc_ctx *c, int32_t s, float R[9], float t[3]);
void orc_ctx_limits(const orc_ctx *c, float out[4]); /* dmin, dmax, imin, imax */

/* 1 - NCC over the best two sources, 2.0 when undefined (dense.rs:181-263). */
float orc_cost(const orc_ctx *c, int32_t x, int32_t y, const float hyp[4]);
void orc_cost_batch(const orc_ctx *c, int64_t n, const int32_t *px, const int32_t *py,
                    const float *hyp4, float *out, int32_t nthreads);

/* state = 5 floats per pixel: d, n0, n1, n2, cost (Hyp, dense.rs:156-161). */
void orc_init(const orc_ctx *c, float *state, uint8_t *textured, int32_t nthreads);
int64_t orc_half_sweep(const orc_ctx *c, float *state, const uint8_t *textured, int32_t it,
                       int32_t parity, int32_t nthreads);
void orc_threshold(const orc_ctx *c, const float *state, float *depth);
int64_t orc_depth_map(const orc_ctx *c, int32_t iterations, int32_t nthreads, float *state,
                      uint8_t *textured, float *depth);

/* Fixed-pair NCC cost in f64 and its analytic gradient — the checks of the
 * GPU's backward pass and refinement (no reference counterpart: the reference
 * has no refinement). cost = 1 - (NCC_s0 + NCC_s1) / 2 for the plane
 * (d, n) with the given two sources; returns 0 when the cost is undefined
 * (window off the image, degenerate plane, failed or flat source). grad (may be
 * NULL) receives d cost / d(d, n0, n1, n2); cells (may be NULL) receives a
 * signature of the bilinear cells every sample fell in, so finite differences
 * can skip steps that cross a cell edge. */
int orc_pair_cost64(const orc_ctx *c, int32_t x, int32_t y, const double hyp[4], int32_t s0,
                    int32_t s1, double *cost, double grad[4], uint64_t *cells);
/* As orc_pair_cost64, and jtj (may be NULL) receives the 4 x 4 Gauss-Newton
 * matrix J^T J (row-major, w.r.t. d, n0, n1, n2) of the pair's stacked
 * residuals r_s = a/|a| - b_s/|b_s| (a, b_s the centred reference and sample
 * windows; 1 - NCC_s = |r_s|^2 / 2). No reference counterpart. */
int orc_pair_terms64(const orc_ctx *c, int32_t x, int32_t y, const double hyp[4], int32_t s0,
                     int32_t s1, double *cost, double grad[4], double jtj[16], uint64_t *cells);
/* The unit centred sample window b_s / |b_s| of source s (N values, tap
 * order), the part of r_s that depends on the plane: the finite-difference
 * check of J^T J differentiates it. 0 where undefined. */
int orc_pair_unit64(const orc_ctx *c, int32_t x, int32_t y, const double hyp[4], int32_t s,
                    double *unit, uint64_t *cells);

/* Geometric consistency of one reference depth map against neighbour maps
 * (dense.rs:585-615). Writes the agree count per pixel (-1 where depth <= 0)
 * and returns the number of pixels with agree >= need. */
typedef struct {
    const orc_view *view;   /* camera of the neighbour (gray unused) */
    const float *depth;     /* its depth map (w*h) */
} orc_nbr_depth;
int64_t orc_consistency(const orc_view *ref, const float *depth, const orc_nbr_depth *nbrs,
                        int32_t nnbr, int32_t *agree, int32_t nthreads);
/* Reprojection rule (no reference counterpart; the rule the 77-view survey
 * workload states): a neighbour agrees when the ref pixel's point lands on a
 * neighbour pixel (floor of the projection, as above) whose own depth, back-
 * projected from that pixel's centre and projected into the reference,
 * falls within 1 px of the ref pixel's centre; consistent when agree >= need. */
int64_t orc_consistency_reproj(const orc_view *ref, const float *depth, const orc_nbr_depth *nbrs,
                               int32_t nnbr, int32_t need, int32_t *agree, int32_t nthreads);

/* Fusion's point emission for one view (dense.rs:589-632): for every pixel
 * with agree >= need under `rule` (0 reference, 1 reprojection), its voxel key
 * floor(xw * inv_voxel) (3 x i32), world point (3 x f64) and pixel index.
 * Arrays hold w*h entries; returns the count. The per-voxel sums are the
 * caller's (the reference's f32 HashMap merge order is not deterministic). */
int64_t orc_fuse_points(const orc_view *ref, const float *depth, const orc_nbr_depth *nbrs,
                        int32_t nnbr, int32_t rule, int32_t need, double inv_voxel,
                        int32_t *keys, double *xw, int32_t *pix);
/* The fusion voxel edge (dense.rs:545-562). ranges: 2 per camera. */
double orc_fusion_voxel(const double *xs, const double *ys, int64_t npts, const double *ranges,
                        int32_t ncams, double f, int32_t size, int32_t width, int32_t height,
                        int64_t max_points);

/* View preparation (dense.rs:92-121 with imgops.rs:89-171): a decoded RGB8
 * photo (stride bytes per row) shrunk so its long side is at most max_side,
 * then its grey. orc_shrink_size gives the box factor f (0: none), the boxed
 * size, the output size and the scale new/old; sizes in, sizes out. */
void orc_shrink_size(int32_t w, int32_t h, int32_t max_side, int32_t *f, int32_t *bw,
                     int32_t *bh, int32_t *ow, int32_t *oh, double *scale);
void orc_box_down(const uint8_t *src, int32_t w, int32_t h, int32_t stride, int32_t f,
                  uint8_t *dst);
void orc_resize_bilinear(const uint8_t *src, int32_t sw, int32_t sh, int32_t stride,
                         uint8_t *dst, int32_t w, int32_t h);
/* the whole of it: out_rgb (ow*oh*3) and grey (ow*oh), sizes from orc_shrink_size */
void orc_prepare_view(const uint8_t *rgb, int32_t w, int32_t h, int32_t stride,
                      int32_t max_side, uint8_t *out_rgb, float *grey);

/* DSM from a point cloud (recon/dsm.rs:128-290). Cloud: n points xyz f32 (as
 * read from dense.ply). budget = cells of the profile (dsm.rs:122-128).
 * Returns 0, or -1 for fewer than 100 points. h (nx*ny, row 0 south) is
 * allocated by the oracle; free with orc_free. */
typedef struct {
    double x0, y0, cell;
    int64_t nx, ny;
    float zmax;
} orc_dsm;
int orc_build_dsm(const float *xyz, int64_t n, double voxel_m, int64_t budget, orc_dsm *out,
                  float **h);
void orc_free(void *p);
/* pieces, for the pins: dsm.rs:293-331 and :335-390 on an nx x ny grid */
void orc_remove_islands(float *h, int64_t nx, int64_t ny, int64_t min_cells);
void orc_fill_enclosed(float *h, int64_t nx, int64_t ny);
/* Dsm::height (dsm.rs:82-119): 1 and *out, or 0 for None */
int orc_dsm_height(const orc_dsm *d, const float *h, double x, double y, float *out);

#ifdef __cplusplus
}
#endif
#endif
