// Internal structures shared by the gfx950 kernels (mvs_kernels.hip) and the
// C-ABI layer (mvs_api.hip). Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace amvs {

// One source view as seen from a reference view: the relative pose of
// dense.rs:299-309 (built in f64 on the host, cast to f32) and the source's
// intrinsics. The f32 intrinsics are exact when `pinhole_exact` is set
// (k1 = k2 = 0 and f, cx, cy representable in f32); otherwise the *_lo parts
// carry the f64 residuals for the float-float distortion path.
struct SrcCam {
    float R[9];
    float t[3];
    float f, cx, cy;
    float f_lo, cx_lo, cy_lo;
    float k1, k1_lo, k2, k2_lo;
    const float *gray;  // device grey plane, row pitch = w
    // 8-bit inputs only (else null): per pixel the bytes of (x, y), (x+1, y),
    // (x, y+1), (x+1, y+1), low to high — a bilinear sample's four corners
    const uint32_t *quad;
    // 8-bit inputs only: per pixel the bilinear coefficients of the cell
    // with top-left corner (x, y), {g00 | A << 16, B | C << 16} as f16
    // (integers, exact), A = g10 - g00, B = g01 - g00, C = g11 - g10 - g01 +
    // g00: a sample is g00 + ax A + ay (B + ax C), from one 8-byte load
    const uint2 *quadh;
    // per pixel the grey of the same four corners (every view)
    const float4 *quadf;
    // f32 grey inputs only (else null): per pixel the same coefficients in
    // f32, {g00, A, B, C} with C = (g11 - g01) - A
    const float4 *quadd;
    int32_t w, h;
    int32_t pinhole_exact;
    int32_t pad;
};

// Everything a PatchMatch kernel needs about one reference view.
struct RefJob {
    const float *gray;    // w*h device grey
    const float2 *rays;   // w*h undistorted rays (f32 of the f64 fixed point)
    float4 *hyp;          // w*h {d, n0, n1, n2}
    float *cost;          // w*h
    uint8_t *textured;    // w*h
    // w*h: 1 where the pixel's last half-sweep changed its hypothesis (0 from
    // the initialisation on; untextured pixels never change)
    uint8_t *changed;
    // w*h: the two sources best at the pixel's plane (first | second << 4),
    // as the one-lane-per-pixel PatchMatch kernels scored that plane; 0xff
    // where the plane was set by any other kernel or the host (unknown)
    uint8_t *pair;
    int32_t *list;        // 2 * list_cap textured pixel indices, by parity
    int32_t *counts;      // [2] list lengths
    const SrcCam *srcs;   // nsrc
    int32_t w, h, nsrc, list_cap;
    float dmin, dmax, imin, imax;
    float up0, up1, up2;
    float texture_min, cost_max;
    int32_t stride;       // tap spacing
    int32_t half;         // window half footprint, pixels
    int32_t pinhole;      // every source has pinhole_exact set
    float ray_step;       // tap spacing / f of the reference camera
    int32_t lane_ok;      // undistorted reference and sources, at most 16: the
                          // one-lane-per-pixel refinement applies (1: 8-bit
                          // views, f16 corner plane; 2: f32 grey views, f32
                          // corner-difference plane)
};

// Arguments of one batched PatchMatch launch (k_pm in mvs_pm_impl.h).
struct PmArgs {
    const RefJob *jobs;   // device array of njobs jobs of one kernel class
    int32_t njobs;
    int32_t *prefix;      // 9 (njobs + 1) chunk and band offsets, by the prep kernel
    int32_t *counters;    // 8 per-XCD work counters, zeroed by the prep kernel
    int32_t it, parity;   // iteration and colour of a half-sweep
    float range;          // 0.5^it
    int32_t gpc;          // pixel groups per work chunk
    unsigned long long *evals;
    // Half-sweeps: skip_unchanged = 1 when both colours have had a half-sweep
    // since the state was set, so a propagation candidate from a neighbour
    // whose last half-sweep left it unchanged was scored by this pixel two
    // half-sweeps before (and kept or rejected); such candidates are counted
    // into *skipped rather than scored (one-lane-per-pixel kernels).
    int32_t skip_unchanged;
    unsigned long long *skipped;
    // Half-sweeps: eval_own = 1 in the first half-sweep of each colour when
    // the pixels' initial planes have no cost yet: each pixel scores its own
    // plane first, ahead of every candidate (one-lane-per-pixel kernels).
    int32_t eval_own;
    const int32_t *px, *py;  // explicit (pixel, hypothesis) list
    const float4 *hyp_in;
    float *out;
    int64_t n_items;
};

// Kernel-side neighbour description for the consistency filter (f64, as the
// reference's fusion, dense.rs:586-612).
struct NbrCam {
    double R[9];
    double C[3];
    double f, cx, cy, k1, k2;
    const float *depth;
    int32_t w, h;
};

struct RefCam64 {
    double R[9];
    double C[3];
    double f, cx, cy, k1, k2;
};

// ---- launchers (mvs_kernels.hip) ----
void launch_gray_from_u8(const uint8_t *src, int32_t w, int32_t h, int32_t stride, float *dst,
                         hipStream_t st);
void launch_quad_from_u8(const uint8_t *src, int32_t w, int32_t h, int32_t stride,
                         uint32_t *dst, hipStream_t st);
void launch_quadf(const float *gray, int32_t w, int32_t h, float4 *dst, hipStream_t st);
void launch_quadd(const float *gray, int32_t w, int32_t h, float4 *dst, hipStream_t st);
void launch_quadh_from_u8(const uint8_t *src, int32_t w, int32_t h, int32_t stride, uint2 *dst,
                          hipStream_t st);
void launch_rays(double f, double cx, double cy, double k1, double k2, int32_t w, int32_t h,
                 float2 *rays, hipStream_t st);
// write_hyp = 0 keeps a caller-provided state and only rebuilds mask + lists.
// job.counts holds 4 + 2 list_waves(w, h) ints: the list lengths, then the
// per-wavefront counts and offsets the lists are built from.
void launch_init(const RefJob &job, int32_t taps, int32_t write_hyp, hipStream_t st);
int64_t list_waves(int32_t w, int32_t h);
// One batched PatchMatch launch (prep + persistent kernel). mode: 0 evaluate
// the current hypotheses, 1 half-sweep, 2 explicit cost list. srcl: source
// lanes per pixel (4, 8, 16); split: lanes per source. Returns -1 for a class
// that is not instantiated.
enum { PM_EVAL_CURRENT = 0, PM_SWEEP = 1, PM_COST_LIST = 2 };
// persist: resident workgroups with atomic work fetch (half-sweeps only);
// otherwise one workgroup per band position. fast: the contracted-arithmetic
// kernels (MVS_MODE_FAST) instead of the reference-rounding ones. pinhole: 0
// some source distorted, 1 undistorted sources, 2 undistorted sources and
// reference. split = -1 selects the one-lane-per-pixel kernels (fast and
// pinhole = 2 only) sampling the f16 corner plane, split = -2 the same kernels
// sampling the 8-bit corner plane, split = -3 the same kernels on f32 grey
// views sampling the f32 corner-difference plane (fast = 2), split = -4 the
// same on f32 grey views sampling the grey plane itself (two row loads per
// sample, a quarter of the bytes per pixel; same values as split = -3), split
// = -5 the f16 corner plane staged in LDS tiles (same values as split = -1).
int launch_pm(int mode, int32_t taps, int32_t srcl, int32_t split, int32_t pinhole,
              bool persist, bool fast, const PmArgs &a, int32_t grid, hipStream_t st);
// One lane per reference pixel (mvs_pm_lane.hip); 64 pixels per group.
// plane: PL_U8 (8-bit corner plane), PL_F16 (f16 corner plane), PL_F32
// (f32 corner-difference plane of f32 grey views), PL_F32G (the f32 grey plane
// itself, read as two 8-byte row pairs per sample), PL_F16T (the f16 corner
// plane staged in LDS per batch and source where the windows fit a tile).
enum { PL_U8 = 0, PL_F16 = 1, PL_F32 = 2, PL_F32G = 3, PL_F16T = 4 };
// Zero floats after a view's grey plane (w*h floats): a row-pair read at the
// last column or row, weighted 0, stays inside the allocation on finite data.
__host__ __device__ inline int64_t grey_pad(int32_t w) { return (int64_t)w + 64; }
int launch_pml(int mode, int32_t taps, int plane, bool persist, const PmArgs &a, int32_t grid,
               hipStream_t st);
int launch_pm_t5(int mode, int32_t srcl, int32_t split, bool pinhole, bool persist,
                 const PmArgs &a, int32_t grid, hipStream_t st);
int launch_pm_t7(int mode, int32_t srcl, int32_t split, bool pinhole, bool persist,
                 const PmArgs &a, int32_t grid, hipStream_t st);
int launch_pm_t11(int mode, int32_t srcl, int32_t split, bool pinhole, bool persist,
                  const PmArgs &a, int32_t grid, hipStream_t st);
int launch_pm_t5f(int mode, int32_t srcl, int32_t split, bool pinhole, bool persist,
                  const PmArgs &a, int32_t grid, hipStream_t st);
int launch_pm_t7f(int mode, int32_t srcl, int32_t split, bool pinhole, bool persist,
                  const PmArgs &a, int32_t grid, hipStream_t st);
int launch_pm_t11f(int mode, int32_t srcl, int32_t split, bool pinhole, bool persist,
                   const PmArgs &a, int32_t grid, hipStream_t st);
// Tap splits instantiated for a window size (ascending), count returned.
int splits_for(int32_t taps, int32_t *out);
void launch_threshold(const RefJob &job, float *depth, float *normal, hipStream_t st);
void launch_consistency(const RefCam64 &ref, const float *depth, int32_t w, int32_t h,
                        const NbrCam *nbrs, int32_t nn, int32_t need, int32_t rule,
                        int32_t *agree, unsigned long long *n_ok, hipStream_t st);
// Clears n u64 counters with a kernel (no memset node, so the launch can sit
// in captured work).
void launch_zero_u64(unsigned long long *p, int32_t n, hipStream_t st);
void launch_selftest_div(int64_t n, uint64_t seed, unsigned long long *bad, hipStream_t st);
// Voxel accumulation table of the fusion (open addressing, linear probing).
// Keys pack the voxel index (21 bits per axis, offset 2^20); coordinate sums
// are fixed point (2^20 units per metre) in 64-bit two's complement, so the
// sums do not depend on the order points arrive in.
struct FuseTable {
    unsigned long long *keys;      // cap, empty = ~0
    unsigned long long *sum;       // 3 per slot
    unsigned int *col;             // 3 per slot: colour sums
    unsigned int *cnt;             // 1 per slot
    unsigned long long *counters;  // [0] voxels, [1] points stored, [2] points dropped
                                   // (table full), [3] points beyond the key range
    int64_t cap;                   // power of two
    int64_t limit;                 // occupancy past which a pass stops inserting
};
void launch_fuse_clear(const FuseTable &t, hipStream_t st);
// The fused cloud on the device (dense.rs:634-641): the n occupied slots
// compacted, sorted by voxel key (radix sort) and averaged into xyz (3n f32)
// and rgb (3n u8), in the order of the keys. 0 or -1 (HIP error).
struct FusePointsWork {
    unsigned long long *keys_in, *keys_out;  // n each
    uint32_t *slots_in, *slots_out;          // n each
    void *tmp;                               // tmp_bytes
    size_t tmp_bytes;
    unsigned long long *ctr;                 // 1
};
size_t fuse_points_tmp_bytes(int64_t n);
int fuse_points(const FuseTable &t, int64_t n, FusePointsWork &w, float *xyz, uint8_t *rgb,
                hipStream_t st);
// One reference view: its confirmed pixels into the table (dense.rs:589-632).
// Colour from rgb (w*h*3, nullable), else the pixel's 8-bit grey (quad byte 0,
// nullable), else the f32 grey x 255.
void launch_fuse(const RefCam64 &ref, const float *depth, int32_t w, int32_t h,
                 const NbrCam *nbrs, int32_t nn, int32_t rule, int32_t need, double inv_voxel,
                 const uint8_t *rgb, const uint32_t *quad, const float *gray, const FuseTable &t,
                 hipStream_t st);

// Gauss-Newton refinement of a job's accepted pixels (steps per pixel);
// counters (nullable) [0] += accepted steps, [1] += pixels refined. Jobs with
// lane_ok run the one-lane-per-pixel kernel (mvs_refine_lane.hip).
int launch_refine(int32_t taps, const RefJob &job, int32_t steps, unsigned long long *counters,
                  hipStream_t st);
int launch_refine_lane(int32_t taps, const RefJob &job, int32_t steps,
                       unsigned long long *counters, hipStream_t st);
// View preparation (mvs_prep.hip): shrink_to's sizes (box factor f, 0 =
// none; boxed size; output size; scale new/old) and the launches from a
// device copy of the photo to the working-size RGB8 and grey.
struct PrepSizes {
    int32_t f, bw, bh, ow, oh;
    double scale;
};
void prep_sizes(int32_t w, int32_t h, int32_t max_side, PrepSizes &z);
void launch_prepare_view(const uint8_t *src, int32_t w, int32_t h, int32_t stride,
                         const PrepSizes &z, uint8_t *boxed, uint8_t *out_rgb, float *grey,
                         hipStream_t st);

// DSM (mvs_dsm.hip): the grid of dsm.rs:129-150 and the device work buffers
// (ncell each; list_a/list_b int32; counters 3 x u32).
struct DsmGrid {
    double x0, y0, cell;
    int64_t nx, ny;
};
struct DsmWork {
    uint32_t *u32a, *u32b, *u32c, *u32d;
    float *snap;
    int32_t *parent, *list_a, *list_b;
    uint32_t *counters;
};
// Heights of the grid from n device points (xyz f32) into d_h; 0 or -1 (HIP error).
int dsm_build(const DsmGrid &g, const float *d_xyz, int64_t n, float *d_h, DsmWork &w,
              hipStream_t st, float *zmax);
// Bytes of device memory dsm_extent works in.
constexpr size_t kDsmSelBytes = 4 * 2048 * 4 + 64;
// The order statistics rank_lo / rank_hi (0-based, in f32::total_cmp order)
// of the x and of the y coordinates of n device points (n < 2^31): out =
// {x lo, x hi, y lo, y hi}. 0 or -1 (HIP error). Synchronises the stream.
int dsm_extent(const float *d_xyz, int64_t n, uint32_t rank_lo, uint32_t rank_hi, void *d_sel,
               hipStream_t st, float out[4]);

// The refinement's Gauss-Newton terms for an explicit (pixel, plane, pair)
// list, by the kernel launch_refine would use for the job (one lane per pixel
// when lane_ok): cost, cost gradient w.r.t. (d, n0, n1, n2) and the 4 x 4
// J^T J of the pair's stacked residuals (packed upper, 10 per item); grad and
// jtj nullable. 0, or -1 for an unsupported window.
int launch_refine_terms(int32_t taps, const RefJob &job, int64_t n, const int32_t *px,
                        const int32_t *py, const float4 *hyp, int32_t *pairs, float *cost,
                        float4 *grad, float *jtj, hipStream_t st);
int launch_terms_lane(int32_t taps, const RefJob &job, int64_t n, const int32_t *px,
                      const int32_t *py, const float4 *hyp, int32_t *pairs, float *cost,
                      float4 *grad, float *jtj, hipStream_t st);
// Fixed-pair cost and gradient of an explicit (pixel, plane, pair) list.
int launch_pair_list(int32_t taps, const RefJob &job, int64_t n, const int32_t *px,
                     const int32_t *py, const float4 *hyp, int32_t *pairs, const float *gout,
                     float *cost, float4 *grad, hipStream_t st);
// Source lanes per reference pixel for nsrc sources (one lane per source).
int lanes_for(int32_t nsrc);
bool taps_supported(int32_t taps);

}  // namespace amvs
