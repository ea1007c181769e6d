// Gauss-Newton refinement with one lane per reference pixel, for views on
// undistorted cameras: 8-bit views sample the f16 bilinear-coefficient plane
// with the reference window as bytes (lane_ok 1), f32 grey views the f32
// coefficient plane with the window in f32 (lane_ok 2). It performs
// the same refinement as k_refine in mvs_refine.hip: LM-damped Gauss-Newton
// on (depth, two normal angles) of the 1 - NCC cost over the pair of sources
// that is best at the PatchMatch plane,
// the cost the reference optimises (dense.rs:255-262). No reference counterpart:
// the reference's depth stage stops after PatchMatch (dense.rs:468-477).
//
// Layout: each lane owns one pixel and evaluates both sources of its pair in
// turn. The best-pair search walks all sources in step, so its gathers are
// coherent like the PatchMatch sweep's; the pairs of neighbouring pixels
// mostly agree, which keeps the Gauss-Newton gathers coherent too. The
// window's rays are affine in the tap offsets (undistorted reference), so the
// plane depth dq = c / (n . r), the source-frame ray w = R r and the
// derivative factors are affine per tap rather than loaded. Samples come from
// the f16 bilinear-coefficient plane (g00 + ax A + ay (B + ax C) of the 8-bit
// image, whose A + ay C and B + ax C are the image gradient); the reference
// window lives in LDS as
// f16 integers b - mi with the exact correction delta = mean(b) - mi. All
// sums are in byte units; NCC and its derivatives are scale-invariant.
//
// With a_k = b_k - mean(b) and s_k the samples, per source (as in mvs_refine.hip):
//   NCC = a.s' / (|a| |s'|),  s' = s - mean(s),
//   dNCC/dtheta = (a.D / |a| - NCC s'.D / |s'|) / |s'|,  D_k = ds_k/dtheta,
//   J^T J = D^T P D / |s'|^2,  P = I - 11^T/N - s's'^T/|s'|^2.
#include <type_traits>

#include "mvs_device.h"

namespace amvs {
namespace {
using namespace dev;

typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef const __attribute__((address_space(1))) u32x2 gu32x2;
__device__ __forceinline__ gu32x2 *as_global_u32x2(const u32x2 *p) { return (gu32x2 *)p; }

constexpr int kMaxSrc = 16;
constexpr int pk(int n) { return n * (n + 1) / 2; }

template <int NDIR>
struct LTerms {
    float ncc = 0.0f, nb = 0.0f;
    float Dta[NDIR > 0 ? NDIR : 1];
    float Dtb[NDIR > 0 ? NDIR : 1];
    float DD[NDIR > 0 ? pk(NDIR) : 1];
    bool ok = false;
};

// This lane's pixel: centre ray, byte-mean correction and |a| (bytes).
struct LaneRef {
    float rcx = 0.0f, rcy = 0.0f, delta = 0.0f, na = 0.0f;
    bool ok = false;
};

// NCC terms of source S for the plane (d, n) at this lane's pixel; NDIR = 3
// adds the derivatives along depth and the two normal directions dirs.
template <int TAPS, int NDIR, class WT>
__device__ LTerms<NDIR> lane_terms(const SrcCam &S, const LaneRef &pr, float rstep, float d,
                                   const float n[3], const float (*dirs)[3],
                                   const WT (*Ah)[64], int lane) {
    constexpr int N = TAPS * TAPS, H = TAPS / 2;
    constexpr int ND = NDIR > 0 ? NDIR : 1;
    constexpr int NP = NDIR > 0 ? pk(NDIR) : 1;
    LTerms<NDIR> T;
    const float nc = fmaf(n[0], pr.rcx, fmaf(n[1], pr.rcy, n[2]));  // n . r_c
    const float cc = d * nc;
    const float ex = n[0] * rstep, ey = n[1] * rstep;  // n . r per tap column / row
    float wc[3], wx[3], wy[3];  // w = R r at the centre tap, per column, per row
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        wc[i] = fmaf(S.R[3 * i], pr.rcx, fmaf(S.R[3 * i + 1], pr.rcy, S.R[3 * i + 2]));
        wx[i] = S.R[3 * i] * rstep;
        wy[i] = S.R[3 * i + 1] * rstep;
    }
    const float F = S.f, CXh = S.cx - 0.5f, CYh = S.cy - 0.5f;
    const float wm1 = (float)(S.w - 1), hm1 = (float)(S.h - 1);
    // The window must lie on the plane's visible side (|n . r| >= 1e-6,
    // dq > 0) and project into the source (X_s.z > 0, inside the image) at
    // every tap; n . r and the scaled source point are affine over the window,
    // so checking the four corner taps suffices.
#pragma unroll
    for (int cr = 0; cr < 4; ++cr) {
        const float fdx = (cr & 1) ? (float)H : -(float)H;
        const float fdy = (cr & 2) ? (float)H : -(float)H;
        const float den = fmaf(fdy, ey, fmaf(fdx, ex, nc));
        if (!(fabsf(den) >= 1e-6f && cc * den > 0.0f)) return T;
        const float dq = cc * __builtin_amdgcn_rcpf(den);
        float X[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) X[i] = fmaf(dq, fmaf(fdy, wy[i], fmaf(fdx, wx[i], wc[i])), S.t[i]);
        if (!(X[2] > 0.0f)) return T;
        const float fi = F * __builtin_amdgcn_rcpf(X[2]);
        const float u = fmaf(fi, X[0], CXh), v = fmaf(fi, X[1], CYh);
        if (!(u >= 0.0f && v >= 0.0f && u < wm1 && v < hm1)) return T;
    }
    // r . dir per normal direction (affine) and d (r_c . dir)
    float rdc[ND], rdx[ND], rdy[ND], drc[ND];
#pragma unroll
    for (int q = 1; q < ND; ++q) {
        const float *dv = dirs[q - 1];
        rdc[q] = fmaf(dv[0], pr.rcx, fmaf(dv[1], pr.rcy, dv[2]));
        rdx[q] = dv[0] * rstep;
        rdy[q] = dv[1] * rstep;
        drc[q] = d * rdc[q];
    }
    constexpr bool F32 = std::is_same<WT, float>::value;
    const uint32_t sw = (uint32_t)S.w;
    const u32x2 *Qp = (const u32x2 *)S.quadh;
    const float4 *Qd = S.quadd;
    float c0 = 0.0f, s1 = 0.0f, s2 = 0.0f, sa = 0.0f;
    float sD[ND], sbD[ND], saD[ND], sDD[NP];
#pragma unroll
    for (int q = 0; q < ND; ++q) sD[q] = sbD[q] = saD[q] = 0.0f;
#pragma unroll
    for (int q = 0; q < NP; ++q) sDD[q] = 0.0f;
#pragma unroll 1
    for (int ty = 0; ty < TAPS; ++ty) {
        const float fdy = (float)(ty - H);
        const float denY = fmaf(fdy, ey, nc);
        float wY[3], rdY[ND];
#pragma unroll
        for (int i = 0; i < 3; ++i) wY[i] = fmaf(fdy, wy[i], wc[i]);
#pragma unroll
        for (int q = 1; q < ND; ++q) rdY[q] = fmaf(fdy, rdy[q], rdc[q]);
#pragma unroll
        for (int tx = 0; tx < TAPS; ++tx) {
            const float fdx = (float)(tx - H);
            const float den = fmaf(fdx, ex, denY);
            const float rden = __builtin_amdgcn_rcpf(den);
            const float dq = cc * rden;
            float w[3], X[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                w[i] = fmaf(fdx, wx[i], wY[i]);
                X[i] = fmaf(dq, w[i], S.t[i]);
            }
            const float inv = __builtin_amdgcn_rcpf(X[2]);
            const float q0 = X[0] * inv, q1 = X[1] * inv;
            // inside up to rounding; the clamp keeps every address in bounds
            const float fx = __builtin_amdgcn_fmed3f(fmaf(F, q0, CXh), 0.0f, wm1);
            const float fy = __builtin_amdgcn_fmed3f(fmaf(F, q1, CYh), 0.0f, hm1);
            const uint32_t x0 = (uint32_t)fx, y0 = (uint32_t)fy;
            const float ax = __builtin_amdgcn_fractf(fx), ay = __builtin_amdgcn_fractf(fy);
            // the cell's bilinear coefficients g00, A, B, C (f16 bytes of 8-bit
            // views, or f32 grey of f32 views): the sample is t1 + ay u with
            // t1 = g00 + ax A, u = B + ax C
            float cg, cA, cB, cC;
            if constexpr (F32) {
                const f32x4 e = as_global(Qd)[__umul24(y0, sw) + x0];
                cg = e.x;
                cA = e.y;
                cB = e.z;
                cC = e.w;
            } else {
                const u32x2 qv = as_global_u32x2(Qp)[__umul24(y0, sw) + x0];
                const uint32_t qa = qv.x, qb = qv.y;
                const h2 r0 = __builtin_bit_cast(h2, qa), r1 = __builtin_bit_cast(h2, qb);
                cg = (float)r0.x;
                cA = (float)r0.y;
                cB = (float)r1.x;
                cC = (float)r1.y;
            }
            const float t1 = fmaf(ax, cA, cg);
            const float u = fmaf(ax, cC, cB);
            const float val = fmaf(ay, u, t1);
            if (tx == 0 && ty == 0) c0 = val;
            const float sv = val - c0;  // shifted sample
            const float a = (float)Ah[ty * TAPS + tx][lane];
            s1 += sv;
            s2 = fmaf(sv, sv, s2);
            sa = fmaf(a, sv, sa);
            if constexpr (NDIR > 0) {
                // image gradient at the sample (per source pixel)
                const float Iv = u;
                const float Iu = fmaf(ay, cC, cA);
                // d(u, v)/d dq and the chain to depth and the normal angles
                const float fi = F * inv;
                const float du = fmaf(-q0, w[2], w[0]) * fi;
                const float dv = fmaf(-q1, w[2], w[1]) * fi;
                const float t0 = fmaf(Iu, du, Iv * dv) * rden;
                float Dk[ND];
                Dk[0] = t0 * nc;
#pragma unroll
                for (int q = 1; q < ND; ++q) {
                    const float rq = fmaf(fdx, rdx[q], rdY[q]);
                    Dk[q] = t0 * fmaf(-dq, rq, drc[q]);
                }
#pragma unroll
                for (int q = 0; q < ND; ++q) {
                    sD[q] += Dk[q];
                    sbD[q] = fmaf(sv, Dk[q], sbD[q]);
                    saD[q] = fmaf(a, Dk[q], saD[q]);
                }
                int m = 0;
#pragma unroll
                for (int p = 0; p < ND; ++p)
#pragma unroll
                    for (int q = p; q < ND; ++q, ++m) sDD[m] = fmaf(Dk[p], Dk[q], sDD[m]);
            }
        }
    }
    const float ms = s1 * (1.0f / (float)N);
    const float nb2 = fmaf(-ms, s1, s2);
    if (!(nb2 > 1e-12f)) return T;
    T.nb = sqrtf(nb2);
    T.ncc = fmaf(-pr.delta, s1, sa) / (pr.na * T.nb);
    if constexpr (NDIR > 0) {
#pragma unroll
        for (int q = 0; q < ND; ++q) {
            T.Dta[q] = fmaf(-pr.delta, sD[q], saD[q]);
            T.Dtb[q] = fmaf(-ms, sD[q], sbD[q]);
        }
        int m = 0;
#pragma unroll
        for (int p = 0; p < ND; ++p)
#pragma unroll
            for (int q = p; q < ND; ++q, ++m)
                T.DD[m] = sDD[m] - sD[p] * sD[q] * (1.0f / (float)N) - T.Dtb[p] * T.Dtb[q] / nb2;
    }
    T.ok = true;
    return T;
}

// NCC of source S for the plane (d, n) at this lane's pixel (lane_terms with
// NDIR = 0, the same definedness tests and sums), with the window's source
// points from the centre-scaled homography: scaled by the centre tap's depth
// dq_c, the source point of tap (dx, dy) is affine in (dx, dy),
//     X_c + dx (R e_x dq_c + t (n . e_x) / (n . r_c)) + dy (...),
// so a tap costs three multiply-adds and one reciprocal rather than the
// plane depth's reciprocal, its scaling and the projection's reciprocal.
// The taps clamp their positions only when some lane's window has a corner
// within 1/256 pixel of the sampling range's edge (rounding moves an inner
// tap by a few 1e-4 pixels at most).
template <int TAPS, class WT>
__device__ LTerms<0> lane_cost(const SrcCam &S, const LaneRef &pr, float rstep, float d,
                               const float n[3], const WT (*Ah)[64], int lane, bool act) {
    constexpr int N = TAPS * TAPS, H = TAPS / 2;
    constexpr bool F32 = std::is_same<WT, float>::value;
    LTerms<0> T;
    const float nc = fmaf(n[0], pr.rcx, fmaf(n[1], pr.rcy, n[2]));  // n . r_c
    const float cc = d * nc;
    const float ex = n[0] * rstep, ey = n[1] * rstep;
    float wc[3], wx[3], wy[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        wc[i] = fmaf(S.R[3 * i], pr.rcx, fmaf(S.R[3 * i + 1], pr.rcy, S.R[3 * i + 2]));
        wx[i] = S.R[3 * i] * rstep;
        wy[i] = S.R[3 * i + 1] * rstep;
    }
    const float F = S.f, CXh = S.cx - 0.5f, CYh = S.cy - 0.5f;
    const float wm1 = (float)(S.w - 1), hm1 = (float)(S.h - 1);
    constexpr float kInner = 1.0f / 256.0f;
    bool inner = true;
#pragma unroll
    for (int cr = 0; cr < 4; ++cr) {
        const float fdx = (cr & 1) ? (float)H : -(float)H;
        const float fdy = (cr & 2) ? (float)H : -(float)H;
        const float den = fmaf(fdy, ey, fmaf(fdx, ex, nc));
        act = act && fabsf(den) >= 1e-6f && cc * den > 0.0f;
        const float dq = cc * __builtin_amdgcn_rcpf(den);
        float X[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) X[i] = fmaf(dq, fmaf(fdy, wy[i], fmaf(fdx, wx[i], wc[i])), S.t[i]);
        act = act && X[2] > 0.0f;
        const float fi = F * __builtin_amdgcn_rcpf(X[2]);
        const float u = fmaf(fi, X[0], CXh), v = fmaf(fi, X[1], CYh);
        act = act && u >= 0.0f && v >= 0.0f && u < wm1 && v < hm1;
        inner = inner && u >= kInner && v >= kInner && u < wm1 - kInner && v < hm1 - kInner;
    }
    // the centre tap's depth and the source-space steps per tap column / row
    const float rnc = act ? __builtin_amdgcn_rcpf(nc) : 0.0f;
    const float dqc = cc * rnc;
    const float pc0 = pr.rcx * dqc, pc1 = pr.rcy * dqc;
    const float sx = dqc * rstep;
    const float tnx = ex * rnc, tny = ey * rnc;
    float B0[3], B1[3], B2[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float xc =
            fmaf(S.R[3 * i], pc0, fmaf(S.R[3 * i + 1], pc1, fmaf(S.R[3 * i + 2], dqc, S.t[i])));
        const float c1 = fmaf(S.t[i], tnx, S.R[3 * i] * sx);
        const float c2 = fmaf(S.t[i], tny, S.R[3 * i + 1] * sx);
        const float sc = i < 2 ? F : 1.0f;
        B0[i] = sc * xc;
        B1[i] = sc * c1;
        B2[i] = sc * c2;
    }
    const uint32_t sw = (uint32_t)S.w;
    const u32x2 *Qp = (const u32x2 *)S.quadh;
    const float4 *Qd = S.quadd;
    float c0 = 0.0f, s1 = 0.0f, s2 = 0.0f, sa = 0.0f;
    auto taps = [&](auto clamp_tag) {
        constexpr bool CLAMP = decltype(clamp_tag)::value;
#pragma unroll 1
        for (int ty = 0; ty < TAPS; ++ty) {
            const float fdy = (float)(ty - H);
            const float Y0 = fmaf(fdy, B2[0], B0[0]);
            const float Y1 = fmaf(fdy, B2[1], B0[1]);
            const float Y2 = fmaf(fdy, B2[2], B0[2]);
#pragma unroll
            for (int tx = 0; tx < TAPS; ++tx) {
                const float fdx = (float)(tx - H);
                const float P0 = tx == H ? Y0 : fmaf(fdx, B1[0], Y0);
                const float P1 = tx == H ? Y1 : fmaf(fdx, B1[1], Y1);
                const float P2 = tx == H ? Y2 : fmaf(fdx, B1[2], Y2);
                const float ip = __builtin_amdgcn_rcpf(P2);
                float fx = fmaf(P0, ip, CXh), fy = fmaf(P1, ip, CYh);
                if constexpr (CLAMP) {
                    fx = __builtin_amdgcn_fmed3f(fx, 0.0f, wm1);
                    fy = __builtin_amdgcn_fmed3f(fy, 0.0f, hm1);
                }
                const uint32_t x0 = (uint32_t)fx, y0 = (uint32_t)fy;
                const float ax = __builtin_amdgcn_fractf(fx), ay = __builtin_amdgcn_fractf(fy);
                // the sample t1 + ay u, t1 = g00 + ax A, u = B + ax C
                float t1, u;
                if constexpr (F32) {
                    const f32x4 e = as_global(Qd)[__umul24(y0, sw) + x0];
                    t1 = fmaf(ax, e.y, e.x);
                    u = fmaf(ax, e.w, e.z);
                } else {
                    const u32x2 qv = as_global_u32x2(Qp)[__umul24(y0, sw) + x0];
                    const h2 r0 = __builtin_bit_cast(h2, (uint32_t)qv.x);
                    const h2 r1 = __builtin_bit_cast(h2, (uint32_t)qv.y);
                    t1 = fmaf(ax, (float)r0.y, (float)r0.x);
                    u = fmaf(ax, (float)r1.y, (float)r1.x);
                }
                const float val = fmaf(ay, u, t1);
                if (tx == 0 && ty == 0) c0 = val;
                const float sv = val - c0;
                s1 += sv;
                s2 = fmaf(sv, sv, s2);
                sa = fmaf((float)Ah[ty * TAPS + tx][lane], sv, sa);
            }
        }
    };
    if (act) {
        if (__ballot(act && !inner) == 0ull)
            taps(std::false_type{});
        else
            taps(std::true_type{});
    }
    const float ms = s1 * (1.0f / (float)N);
    const float nb2 = fmaf(-ms, s1, s2);
    if (!(act && nb2 > 1e-12f)) return T;
    T.nb = sqrtf(nb2);
    T.ncc = fmaf(-pr.delta, s1, sa) / (pr.na * T.nb);
    T.ok = true;
    return T;
}

// a ranks before b: higher NCC, then lower source index (dense.rs:255-256).
__device__ inline bool lane_before(float na, int ia, float nb, int ib) {
    return na > nb || (na == nb && ia < ib);
}

__device__ bool lane_solve3(const float H[6], const float b[3], float x[3]) {
    const float a00 = H[0], a01 = H[1], a02 = H[2], a11 = H[3], a12 = H[4], a22 = H[5];
    const float c00 = a11 * a22 - a12 * a12, c01 = a02 * a12 - a01 * a22,
                c02 = a01 * a12 - a02 * a11;
    const float det = a00 * c00 + a01 * c01 + a02 * c02;
    if (!(det > 0.0f) || !isfinite(det)) return false;
    const float c11 = a00 * a22 - a02 * a02, c12 = a01 * a02 - a00 * a12,
                c22 = a00 * a11 - a01 * a01;
    const float id = 1.0f / det;
    x[0] = (c00 * b[0] + c01 * b[1] + c02 * b[2]) * id;
    x[1] = (c01 * b[0] + c11 * b[1] + c12 * b[2]) * id;
    x[2] = (c02 * b[0] + c12 * b[1] + c22 * b[2]) * id;
    return isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2]);
}

// This lane's reference window at pixel pix: the bytes b = rint(255 g) of the
// 8-bit view, a = b - mi into LDS as f16 integers (exact), the correction
// delta = mean(b) - mi, |a| and the centre ray. The caller checks that the
// window lies inside the view.
template <int TAPS, class WT>
__device__ LaneRef lane_ref(const RefJob &job, int32_t pix, WT (*Ah)[64], int lane) {
    constexpr int N = TAPS * TAPS, H = TAPS / 2;
    LaneRef pr;
    const int st = job.stride, w = job.w;
    if constexpr (std::is_same<WT, float>::value) {
        // f32 views: the grey less its window mean, f32 (delta = 0)
        float sg = 0.0f;
#pragma unroll 1
        for (int ty = 0; ty < TAPS; ++ty) {
            const float *row = job.gray + pix + (int64_t)((ty - H) * st) * w;
#pragma unroll
            for (int tx = 0; tx < TAPS; ++tx) sg += row[(tx - H) * st];
        }
        const float mg = sg * (1.0f / (float)N);
        float s2 = 0.0f;
#pragma unroll 1
        for (int ty = 0; ty < TAPS; ++ty) {
            const float *row = job.gray + pix + (int64_t)((ty - H) * st) * w;
#pragma unroll
            for (int tx = 0; tx < TAPS; ++tx) {
                const float av = row[(tx - H) * st] - mg;
                Ah[ty * TAPS + tx][lane] = av;
                s2 = fmaf(av, av, s2);
            }
        }
        pr.delta = 0.0f;
        pr.na = sqrtf(s2);
        const float2 rc = job.rays[pix];
        pr.rcx = rc.x;
        pr.rcy = rc.y;
        pr.ok = pr.na > 1e-6f;
        return pr;
    }
    float sb = 0.0f;
#pragma unroll 1
    for (int ty = 0; ty < TAPS; ++ty) {
        const float *row = job.gray + pix + (int64_t)((ty - H) * st) * w;
#pragma unroll
        for (int tx = 0; tx < TAPS; ++tx) sb += rintf(row[(tx - H) * st] * 255.0f);
    }
    const float mi = rintf(sb * (1.0f / (float)N));
    float s2 = 0.0f;
#pragma unroll 1
    for (int ty = 0; ty < TAPS; ++ty) {
        const float *row = job.gray + pix + (int64_t)((ty - H) * st) * w;
#pragma unroll
        for (int tx = 0; tx < TAPS; ++tx) {
            const float av = rintf(row[(tx - H) * st] * 255.0f) - mi;
            Ah[ty * TAPS + tx][lane] = (WT)av;
            s2 = fmaf(av, av, s2);
        }
    }
    pr.delta = (sb - (float)N * mi) * (1.0f / (float)N);
    // |a|^2 = sum (b - mi)^2 - N delta^2
    pr.na = sqrtf(fmaxf(fmaf(-(float)N * pr.delta, pr.delta, s2), 0.0f));
    const float2 rc = job.rays[pix];
    pr.rcx = rc.x;
    pr.rcy = rc.y;
    pr.ok = pr.na > 255e-6f;
    return pr;
}

// The two sources best at the plane (d, n) (highest NCC first, ties to the
// lower index, dense.rs:255-256), every source in step across the wavefront;
// i1 >= nsrc when fewer than two are defined.
template <int TAPS, class WT>
__device__ void lane_best_pair(const SrcCam *Ss, int nsrc, const LaneRef &pr, float rstep,
                               float d, const float n[3], const WT (*Ah)[64], int lane,
                               bool search, int &i0, int &i1) {
    float b0 = -INFINITY, b1 = -INFINITY;
    i0 = 1 << 20;
    i1 = 1 << 20;
    for (int k = 0; k < nsrc; ++k) {
        {
            const LTerms<0> T = lane_cost<TAPS, WT>(Ss[k], pr, rstep, d, n, Ah, lane, search);
            if (T.ok) {
                if (lane_before(T.ncc, k, b0, i0)) {
                    b1 = b0;
                    i1 = i0;
                    b0 = T.ncc;
                    i0 = k;
                } else if (lane_before(T.ncc, k, b1, i1)) {
                    b1 = T.ncc;
                    i1 = k;
                }
            }
        }
    }
}

// Gauss-Newton terms of the pair (A, B) at the plane (d, n): J^T J (packed
// upper, w.r.t. depth and the angles along dirs) and J^T r of the stacked
// residuals, and the pair cost 1 - (NCC_A + NCC_B) / 2. false where either
// source is undefined.
template <int TAPS, class WT>
__device__ bool lane_pair_gn(const SrcCam &A, const SrcCam &B, const LaneRef &pr, float rstep,
                             float d, const float n[3], const float dirs[2][3],
                             const WT (*Ah)[64], int lane, float Hm[6], float g[3],
                             float &cost) {
    const LTerms<3> TA = lane_terms<TAPS, 3, WT>(A, pr, rstep, d, n, dirs, Ah, lane);
    const LTerms<3> TB = lane_terms<TAPS, 3, WT>(B, pr, rstep, d, n, dirs, Ah, lane);
    if (!(TA.ok && TB.ok)) return false;
#pragma unroll
    for (int q = 0; q < 6; ++q) Hm[q] = TA.DD[q] / (TA.nb * TA.nb) + TB.DD[q] / (TB.nb * TB.nb);
#pragma unroll
    for (int q = 0; q < 3; ++q)
        g[q] = -(TA.Dta[q] / pr.na - TA.ncc * TA.Dtb[q] / TA.nb) / TA.nb -
               (TB.Dta[q] / pr.na - TB.ncc * TB.Dtb[q] / TB.nb) / TB.nb;
    cost = 1.0f - 0.5f * (TA.ncc + TB.ncc);
    return true;
}

// Waves per SIMD the refinement kernel is built for: two, which leaves the
// registers to issue a tap row's gathers together in the Gauss-Newton steps
// (at three waves the compiler issues them one at a time there: 77-view
// refinement 302 -> 281 ms at 7x7, 842 -> 585 ms at 11x11); the 11x11 window
// on f32 views, which needs more registers than two waves have, keeps the
// compiler's own choice.
template <int TAPS, class WT>
constexpr bool kRefineTwoWaves = TAPS <= 7 || !std::is_same<WT, float>::value;
template <int TAPS, class WT>
constexpr int kRefineWavesMin = kRefineTwoWaves<TAPS, WT> ? 2 : 1;
template <int TAPS, class WT>
constexpr int kRefineWavesMax = kRefineTwoWaves<TAPS, WT> ? 2 : 8;

// One lane per pixel of the job's textured-pixel lists (64 per wavefront);
// pixels whose PatchMatch cost is below cost_max are refined in place.
// counters (nullable): [0] += accepted steps, [1] += pixels refined.
template <int TAPS, class WT>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(kRefineWavesMin<TAPS, WT>, kRefineWavesMax<TAPS, WT>)))
void k_refine_lane(RefJob job, int32_t steps, unsigned long long *counters) {
    constexpr int N = TAPS * TAPS;
    __shared__ WT Ah[N][64];
    __shared__ SrcCam Ss[kMaxSrc];
    const int lane = threadIdx.x;
    const int64_t c0 = job.counts[0], total = c0 + job.counts[1];
    if ((int64_t)blockIdx.x * 64 >= total) return;
    {
        const int words = job.nsrc * (int)(sizeof(SrcCam) / 4);
        const uint32_t *src = (const uint32_t *)job.srcs;
        uint32_t *dst = (uint32_t *)Ss;
        for (int i = lane; i < words; i += 64) dst[i] = src[i];
    }
    __syncthreads();
    const int64_t item = (int64_t)blockIdx.x * 64 + lane;
    bool have = item < total;
    int32_t pix = 0;
    if (have) pix = item < c0 ? job.list[item] : job.list[job.list_cap + (item - c0)];
    const int w = job.w, y = pix / w, x = pix - y * w;
    float4 hp = have ? job.hyp[pix] : make_float4(1, 0, 0, -1);
    have = have && job.cost[pix] < job.cost_max && x >= job.half && y >= job.half &&
           x < w - job.half && y < job.h - job.half;
    LaneRef pr;
    if (have) pr = lane_ref<TAPS, WT>(job, pix, Ah, lane);
    const float rstep = job.ray_step;
    float nrm[3] = {hp.y, hp.z, hp.w};
    // best pair at the PatchMatch plane: every source in step
    int i0, i1;
    const bool search = have && pr.ok;
    lane_best_pair<TAPS, WT>(Ss, job.nsrc, pr, rstep, hp.x, nrm, Ah, lane, search, i0, i1);
    bool live = search && i1 < job.nsrc;
    const bool refined = live;
    const int p0 = live ? i0 : 0, p1 = live ? i1 : 0;
    const float lo = job.dmin * 0.7f, hi = job.dmax * 1.4f;
    float lambda = 1e-3f;
    unsigned long long acc = 0;
    // J^T J, the gradient and the cost at the current plane, and the tangent
    // directions they are taken along. A trial plane's pass forms the same
    // terms (except on the last step, which needs only its cost), so an
    // accepted step carries them to the next one and a rejected step keeps
    // the current ones: one window pass per source and step.
    float Hm[6], g[3], cost0 = 0.0f;
    float dirs[2][3];
    if (live) {
        tangent_dirs(nrm, dirs);
        live = lane_pair_gn<TAPS, WT>(Ss[p0], Ss[p1], pr, rstep, hp.x, nrm, dirs, Ah, lane, Hm, g,
                                      cost0);
    }
    for (int it = 0; it < steps; ++it) {
        if (!live) break;
        const float Hd[6] = {Hm[0] * (1.0f + lambda), Hm[1], Hm[2], Hm[3] * (1.0f + lambda),
                             Hm[4], Hm[5] * (1.0f + lambda)};
        const float rhs[3] = {-g[0], -g[1], -g[2]};
        float dx[3] = {0.0f, 0.0f, 0.0f};
        const bool solved = lane_solve3(Hd, rhs, dx);
        // bounded step: 2% of the depth, 0.1 rad per normal angle
        float d1 = hp.x, n1[3] = {nrm[0], nrm[1], nrm[2]};
        if (solved) {
            const float lim = 0.02f * hp.x;
            dx[0] = fminf(fmaxf(dx[0], -lim), lim);
            dx[1] = fminf(fmaxf(dx[1], -0.1f), 0.1f);
            dx[2] = fminf(fmaxf(dx[2], -0.1f), 0.1f);
            d1 = hp.x + dx[0];
#pragma unroll
            for (int c3 = 0; c3 < 3; ++c3)
                n1[c3] = nrm[c3] + dx[1] * dirs[0][c3] + dx[2] * dirs[1][c3];
            const float ln = sqrtf(n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2]);
#pragma unroll
            for (int c3 = 0; c3 < 3; ++c3) n1[c3] /= ln;
        }
        const bool cand = solved && (n1[0] * pr.rcx + n1[1] * pr.rcy + n1[2] < 0.0f) &&
                          d1 >= lo && d1 <= hi;
        bool better = false;
        if (cand && it + 1 < steps) {
            // the trial's cost with its own terms (the NCC sums are the same
            // arithmetic as the cost-only pass's)
            float dirs1[2][3], Hm1[6], g1[3], cost1 = 0.0f;
            tangent_dirs(n1, dirs1);
            better = lane_pair_gn<TAPS, WT>(Ss[p0], Ss[p1], pr, rstep, d1, n1, dirs1, Ah, lane,
                                            Hm1, g1, cost1) &&
                     cost1 < cost0;
            if (better) {
#pragma unroll
                for (int q = 0; q < 6; ++q) Hm[q] = Hm1[q];
#pragma unroll
                for (int q = 0; q < 3; ++q) g[q] = g1[q];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    dirs[0][q] = dirs1[0][q];
                    dirs[1][q] = dirs1[1][q];
                }
                cost0 = cost1;
            }
        } else if (cand) {
            const LTerms<0> CA = lane_cost<TAPS, WT>(Ss[p0], pr, rstep, d1, n1, Ah, lane, true);
            const LTerms<0> CB = lane_cost<TAPS, WT>(Ss[p1], pr, rstep, d1, n1, Ah, lane, true);
            better = CA.ok && CB.ok && 1.0f - 0.5f * (CA.ncc + CB.ncc) < cost0;
        }
        if (better) {
            hp = make_float4(d1, n1[0], n1[1], n1[2]);
            nrm[0] = n1[0];
            nrm[1] = n1[1];
            nrm[2] = n1[2];
            lambda = fmaxf(lambda * 0.1f, 1e-6f);
            ++acc;
        } else {
            lambda = fminf(lambda * 10.0f, 1e6f);
        }
    }
    if (have && acc) job.hyp[pix] = hp;
    if (counters) {
        unsigned long long e = acc, p = refined ? 1 : 0;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            e += __shfl_xor(e, off);
            p += __shfl_xor(p, off);
        }
        if (lane == 0 && e) atomicAdd(&counters[0], e);
        if (lane == 0 && p) atomicAdd(&counters[1], p);
    }
}

// The Gauss-Newton terms k_refine_lane forms, for an explicit list of (pixel,
// plane, source pair) items, one lane per item: the refinement's own window,
// pair search and terms, mapped to (d, n0, n1, n2). pairs: an item given
// (-1, -1) gets its best pair at that plane, written back. Undefined items:
// cost 2, gradient and J^T J 0.
template <int TAPS, class WT>
__global__ __launch_bounds__(64) void k_terms_lane(RefJob job, int64_t n,
                                                   const int32_t *__restrict__ px,
                                                   const int32_t *__restrict__ py,
                                                   const float4 *__restrict__ hyp, int32_t *pairs,
                                                   float *cost, float4 *grad, float *jtj) {
    constexpr int N = TAPS * TAPS;
    __shared__ WT Ah[N][64];
    __shared__ SrcCam Ss[kMaxSrc];
    const int lane = threadIdx.x;
    {
        const int words = job.nsrc * (int)(sizeof(SrcCam) / 4);
        const uint32_t *src = (const uint32_t *)job.srcs;
        uint32_t *dst = (uint32_t *)Ss;
        for (int i = lane; i < words; i += 64) dst[i] = src[i];
    }
    __syncthreads();
    const int64_t item = (int64_t)blockIdx.x * 64 + lane;
    const bool inlist = item < n;
    const int x = inlist ? px[item] : 0, y = inlist ? py[item] : 0;
    const float4 hp = inlist ? hyp[item] : make_float4(1, 0, 0, -1);
    const bool have = inlist && x >= job.half && y >= job.half && x < job.w - job.half &&
                      y < job.h - job.half;
    const int32_t pix = y * job.w + x;
    LaneRef pr;
    if (have) pr = lane_ref<TAPS, WT>(job, pix, Ah, lane);
    const float rstep = job.ray_step;
    const float nrm[3] = {hp.y, hp.z, hp.w};
    int p0 = inlist ? pairs[2 * item] : 0, p1 = inlist ? pairs[2 * item + 1] : 0;
    const bool choose = p0 < 0;
    int i0, i1;
    lane_best_pair<TAPS, WT>(Ss, job.nsrc, pr, rstep, hp.x, nrm, Ah, lane, have && pr.ok && choose,
                         i0, i1);
    if (choose) {
        p0 = i1 < job.nsrc ? i0 : -1;
        p1 = i1 < job.nsrc ? i1 : -1;
    }
    float Hm[6], g[3], c = 2.0f;
    float dirs[2][3];
    tangent_dirs(nrm, dirs);
    bool ok = false;
    if (have && pr.ok && p0 >= 0 && p1 >= 0 && p0 != p1)
        ok = lane_pair_gn<TAPS, WT>(Ss[p0], Ss[p1], pr, rstep, hp.x, nrm, dirs, Ah, lane, Hm, g, c);
    if (!inlist) return;
    if (choose) {
        pairs[2 * item] = p0;
        pairs[2 * item + 1] = p1;
    }
    float g4[4] = {0, 0, 0, 0}, h10[10];
#pragma unroll
    for (int m = 0; m < 10; ++m) h10[m] = 0.0f;
    if (ok) terms_to_dn(Hm, g, dirs, g4, h10);
    cost[item] = ok ? c : 2.0f;
    if (grad) grad[item] = make_float4(g4[0], g4[1], g4[2], g4[3]);
    if (jtj)
#pragma unroll
        for (int m = 0; m < 10; ++m) jtj[10 * item + m] = h10[m];
}

template <int TAPS>
void terms_lane_taps(const RefJob &job, int64_t n, const int32_t *px, const int32_t *py,
                     const float4 *hyp, int32_t *pairs, float *cost, float4 *grad, float *jtj,
                     hipStream_t st) {
    const unsigned blocks = (unsigned)((n + 63) / 64);
    if (blocks == 0) return;
    if (job.lane_ok == 2)
        hipLaunchKernelGGL((k_terms_lane<TAPS, float>), dim3(blocks), dim3(64), 0, st, job, n, px,
                           py, hyp, pairs, cost, grad, jtj);
    else
        hipLaunchKernelGGL((k_terms_lane<TAPS, _Float16>), dim3(blocks), dim3(64), 0, st, job, n,
                           px, py, hyp, pairs, cost, grad, jtj);
}

template <int TAPS>
void refine_lane_taps(const RefJob &job, int32_t steps, unsigned long long *counters,
                      hipStream_t st) {
    const unsigned blocks = (unsigned)((2 * (int64_t)job.list_cap + 63) / 64);
    if (job.lane_ok == 2)
        hipLaunchKernelGGL((k_refine_lane<TAPS, float>), dim3(blocks), dim3(64), 0, st, job, steps,
                           counters);
    else
        hipLaunchKernelGGL((k_refine_lane<TAPS, _Float16>), dim3(blocks), dim3(64), 0, st, job,
                           steps, counters);
}

}  // namespace

int launch_terms_lane(int32_t taps, const RefJob &job, int64_t n, const int32_t *px,
                      const int32_t *py, const float4 *hyp, int32_t *pairs, float *cost,
                      float4 *grad, float *jtj, hipStream_t st) {
    if (job.nsrc > kMaxSrc) return -1;
    switch (taps) {
        case 5: terms_lane_taps<5>(job, n, px, py, hyp, pairs, cost, grad, jtj, st); return 0;
        case 7: terms_lane_taps<7>(job, n, px, py, hyp, pairs, cost, grad, jtj, st); return 0;
        case 11: terms_lane_taps<11>(job, n, px, py, hyp, pairs, cost, grad, jtj, st); return 0;
        default: return -1;
    }
}

int launch_refine_lane(int32_t taps, const RefJob &job, int32_t steps,
                       unsigned long long *counters, hipStream_t st) {
    if (job.nsrc > kMaxSrc) return -1;
    switch (taps) {
        case 5: refine_lane_taps<5>(job, steps, counters, st); return 0;
        case 7: refine_lane_taps<7>(job, steps, counters, st); return 0;
        case 11: refine_lane_taps<11>(job, steps, counters, st); return 0;
        default: return -1;
    }
}

}  // namespace amvs
