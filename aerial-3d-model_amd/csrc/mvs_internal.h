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
void launch_rays(double f, double cx, double cy, double k1, double k2, int32_t w, int32_t h,
                 float2 *rays, hipStream_t st);
// write_hyp = 0 keeps a caller-provided state and only rebuilds mask + lists.
void launch_init(const RefJob &job, int32_t taps, int32_t write_hyp, hipStream_t st);
// Evaluate the current hypothesis of every textured pixel (both lists).
int launch_eval_current(const RefJob &job, int32_t taps, unsigned long long *evals,
                        hipStream_t st);
int launch_sweep(const RefJob &job, int32_t taps, int32_t it, int32_t parity,
                 unsigned long long *evals, hipStream_t st);
int launch_cost_list(const RefJob &job, int32_t taps, int64_t n, const int32_t *px,
                     const int32_t *py, const float4 *hyp, float *out, hipStream_t st);
void launch_threshold(const RefJob &job, float *depth, float *normal, hipStream_t st);
void launch_consistency(const RefCam64 &ref, const float *depth, int32_t w, int32_t h,
                        const NbrCam *nbrs, int32_t nn, int32_t need, int32_t *agree,
                        unsigned long long *n_ok, hipStream_t st);
void launch_selftest_div(int64_t n, uint64_t seed, unsigned long long *bad, hipStream_t st);
// Lanes per reference pixel for nsrc sources (one lane per source).
int lanes_for(int32_t nsrc);
bool taps_supported(int32_t taps);

}  // namespace amvs
