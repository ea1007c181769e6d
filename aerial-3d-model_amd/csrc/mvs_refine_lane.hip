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
// scaled source points are affine per tap (a homography), and so are the
// factors that carry a sample's derivative along the plane vector to depth
// and the normal angles: the taps sum moments, and the derivative sums follow
// once per window (lane_pass). Samples come from
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

// One window pass of source S for the plane (d, n) at this lane's pixel: the
// NCC, and with DERIV its derivatives along depth and the two normal angles
// (tangent directions dirs). Lanes with act false (or whose window fails the
// definedness tests) return ok = false.
//
// The window must lie on the plane's visible side (|n . r| >= 1e-6, dq > 0)
// and project into the source (X_s.z > 0, inside the image) at every tap;
// n . r and the scaled source point are affine over the window, so checking
// the four corner taps suffices. The source points come from the
// centre-scaled homography: scaled by the centre tap's depth dq_c, the source
// point of tap (dx, dy) is affine in (dx, dy),
//     X_c + dx (R e_x dq_c + t (n . e_x) / (n . r_c)) + dy (...),
// so a tap costs three multiply-adds and one reciprocal. The taps clamp their
// positions only when some lane's window has a corner within 1/256 pixel of
// the sampling range's edge (rounding moves an inner tap by a few 1e-4 pixels
// at most).
//
// Derivatives. With the plane written m . X = 1 (m = n / (d n . r_c)), a tap
// samples the source at the projection of R r + t (m . r), so
//     ds/dm = d b r,  b = (Iu (Kx - u tz) + Iv (Ky - v tz)) / P_z,
// with (u, v) the source pixel position, (Iu, Iv) the image gradient there,
// Kx = F tx + (cx - 1/2) tz, Ky likewise and P_z the homography's third
// coordinate. The chain to theta = (depth, angle 1, angle 2) multiplies b by
// a factor affine in the tap offsets,
//     ds/dtheta_j = b g_j(dx, dy),  g_0 = -(n . r) / c,
//     g_q = (dx (dir_q.x rstep - (n.x rstep)(dir_q . r_c) / nc)
//          + dy (dir_q.y rstep - (n.y rstep)(dir_q . r_c) / nc)) / nc
// (c = d nc, nc = n . r_c; the angles move the plane about the centre
// point, so g_q vanishes there). The taps therefore only sum b, b^2, s b
// and a b against 1, dx, dy (and b^2 against their products); the
// derivative sums follow once per window from these moments.
template <int TAPS, class WT, bool DERIV>
__device__ LTerms<DERIV ? 3 : 0> lane_pass(const SrcCam &S, const LaneRef &pr, float rstep,
                                           float d, const float n[3], const float (*dirs)[3],
                                           const WT (*Ah)[64], int lane, bool act) {
    constexpr int N = TAPS * TAPS, H = TAPS / 2;
    constexpr bool F32 = std::is_same<WT, float>::value;
    LTerms<DERIV ? 3 : 0> T;
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
    const float tz = S.t[2];
    const float Kx = fmaf(F, S.t[0], CXh * tz), Ky = fmaf(F, S.t[1], CYh * tz);
    const uint32_t sw = (uint32_t)S.w;
    const u32x2 *Qp = (const u32x2 *)S.quadh;
    const float4 *Qd = S.quadd;
    float c0 = 0.0f, s1 = 0.0f, s2 = 0.0f, sa = 0.0f;
    // moments of b against (1, dx, dy) (m1), of s b (mb) and a b (ma), and of
    // b^2 against 1, dx, dy, dx^2, dx dy, dy^2 (m2)
    float m1[3] = {0.0f, 0.0f, 0.0f}, mb[3] = {0.0f, 0.0f, 0.0f}, ma[3] = {0.0f, 0.0f, 0.0f};
    float m2[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    auto taps = [&](auto clamp_tag) {
        constexpr bool CLAMP = decltype(clamp_tag)::value;
#pragma unroll 1
        for (int ty = 0; ty < TAPS; ++ty) {
            const float fdy = (float)(ty - H);
            const float Y0 = fmaf(fdy, B2[0], B0[0]);
            const float Y1 = fmaf(fdy, B2[1], B0[1]);
            const float Y2 = fmaf(fdy, B2[2], B0[2]);
            // this row's moments against 1 and dx (and dx^2 for b^2)
            float rb0 = 0.0f, rbx = 0.0f, rq0 = 0.0f, rqx = 0.0f, rqxx = 0.0f;
            float rs0 = 0.0f, rsx = 0.0f, ra0 = 0.0f, rax = 0.0f;
            // one tap row per trip: the row's positions and gathers, then
            // (a scheduling barrier keeps the compiler from sinking each
            // gather to its use) its samples and sums
            float fxs[TAPS], fys[TAPS], ips[TAPS];
            typedef std::conditional_t<F32, f32x4, u32x2> QT;
            QT qs[TAPS];
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
                if constexpr (F32)
                    qs[tx] = as_global(Qd)[__umul24(y0, sw) + x0];
                else
                    qs[tx] = as_global_u32x2(Qp)[__umul24(y0, sw) + x0];
                fxs[tx] = fx;
                fys[tx] = fy;
                ips[tx] = ip;
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int tx = 0; tx < TAPS; ++tx) {
                const float fdx = (float)(tx - H);
                const float fx = fxs[tx], fy = fys[tx], ip = ips[tx];
                const float ax = __builtin_amdgcn_fractf(fx), ay = __builtin_amdgcn_fractf(fy);
                // the cell's bilinear coefficients g00, A, B, C (f16 bytes of
                // 8-bit views, or f32 grey of f32 views): the sample is
                // t1 + ay u with t1 = g00 + ax A, u = B + ax C; the gradient
                // is (A + ay C, u)
                float t1, u, Iu;
                if constexpr (F32) {
                    const f32x4 e = qs[tx];
                    t1 = fmaf(ax, e.y, e.x);
                    u = fmaf(ax, e.w, e.z);
                    if constexpr (DERIV) Iu = fmaf(ay, e.w, e.y);
                } else {
                    const u32x2 qv = qs[tx];
                    const h2 r0 = __builtin_bit_cast(h2, (uint32_t)qv.x);
                    const h2 r1 = __builtin_bit_cast(h2, (uint32_t)qv.y);
                    t1 = fmaf(ax, (float)r0.y, (float)r0.x);
                    u = fmaf(ax, (float)r1.y, (float)r1.x);
                    if constexpr (DERIV) Iu = fmaf(ay, (float)r1.y, (float)r0.y);
                }
                const float val = fmaf(ay, u, t1);
                if (tx == 0 && ty == 0) c0 = val;
                const float sv = val - c0;  // shifted sample
                const float a = (float)Ah[ty * TAPS + tx][lane];
                s1 += sv;
                s2 = fmaf(sv, sv, s2);
                sa = fmaf(a, sv, sa);
                if constexpr (DERIV) {
                    const float b = fmaf(Iu, fmaf(-fx, tz, Kx), u * fmaf(-fy, tz, Ky)) * ip;
                    const float bb = b * b, sb = sv * b, ab = a * b;
                    rb0 += b;
                    rq0 += bb;
                    rs0 += sb;
                    ra0 += ab;
                    if (tx != H) {
                        rbx = fmaf(b, fdx, rbx);
                        rqx = fmaf(bb, fdx, rqx);
                        rqxx = fmaf(bb, fdx * fdx, rqxx);
                        rsx = fmaf(sb, fdx, rsx);
                        rax = fmaf(ab, fdx, rax);
                    }
                }
            }
            if constexpr (DERIV) {
                m1[0] += rb0;
                m1[1] += rbx;
                m1[2] = fmaf(fdy, rb0, m1[2]);
                mb[0] += rs0;
                mb[1] += rsx;
                mb[2] = fmaf(fdy, rs0, mb[2]);
                ma[0] += ra0;
                ma[1] += rax;
                ma[2] = fmaf(fdy, ra0, ma[2]);
                m2[0] += rq0;                         // 1
                m2[1] += rqx;                         // dx
                m2[2] = fmaf(fdy, rq0, m2[2]);        // dy
                m2[3] += rqxx;                        // dx^2
                m2[4] = fmaf(fdy, rqx, m2[4]);        // dx dy
                m2[5] = fmaf(fdy * fdy, rq0, m2[5]);  // dy^2
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
    if constexpr (DERIV) {
        // g_j = G[j][0] + dx G[j][1] + dy G[j][2]
        const float rcc = rnc / d;
        float G[3][3];
        G[0][0] = -1.0f / d;
        G[0][1] = -ex * rcc;
        G[0][2] = -ey * rcc;
#pragma unroll
        for (int q = 1; q < 3; ++q) {
            const float *dv = dirs[q - 1];
            const float k = fmaf(dv[0], pr.rcx, fmaf(dv[1], pr.rcy, dv[2])) * rnc;  // (dir . r_c) / nc
            G[q][0] = 0.0f;
            G[q][1] = fmaf(-ex, k, dv[0] * rstep) * rnc;
            G[q][2] = fmaf(-ey, k, dv[1] * rstep) * rnc;
        }
        const float M2[3][3] = {{m2[0], m2[1], m2[2]}, {m2[1], m2[3], m2[4]}, {m2[2], m2[4], m2[5]}};
        float sD[3], sbD[3], saD[3], GM[3][3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            sD[j] = fmaf(G[j][0], m1[0], fmaf(G[j][1], m1[1], G[j][2] * m1[2]));
            sbD[j] = fmaf(G[j][0], mb[0], fmaf(G[j][1], mb[1], G[j][2] * mb[2]));
            saD[j] = fmaf(G[j][0], ma[0], fmaf(G[j][1], ma[1], G[j][2] * ma[2]));
#pragma unroll
            for (int c = 0; c < 3; ++c)
                GM[j][c] = fmaf(G[j][0], M2[0][c], fmaf(G[j][1], M2[1][c], G[j][2] * M2[2][c]));
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            T.Dta[q] = fmaf(-pr.delta, sD[q], saD[q]);
            T.Dtb[q] = fmaf(-ms, sD[q], sbD[q]);
        }
        int m = 0;
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = p; q < 3; ++q, ++m) {
                const float sdd = fmaf(GM[p][0], G[q][0], fmaf(GM[p][1], G[q][1], GM[p][2] * G[q][2]));
                T.DD[m] = sdd - sD[p] * sD[q] * (1.0f / (float)N) - T.Dtb[p] * T.Dtb[q] / nb2;
            }
    }
    T.ok = true;
    return T;
}

// The NCC terms of source S at the plane (d, n) with their derivatives along
// depth and the two normal angles dirs.
template <int TAPS, class WT>
__device__ LTerms<3> lane_terms(const SrcCam &S, const LaneRef &pr, float rstep, float d,
                                const float n[3], const float (*dirs)[3], const WT (*Ah)[64],
                                int lane) {
    return lane_pass<TAPS, WT, true>(S, pr, rstep, d, n, dirs, Ah, lane, true);
}

// The NCC of source S at the plane (d, n) alone.
template <int TAPS, class WT>
__device__ LTerms<0> lane_cost(const SrcCam &S, const LaneRef &pr, float rstep, float d,
                               const float n[3], const WT (*Ah)[64], int lane, bool act) {
    return lane_pass<TAPS, WT, false>(S, pr, rstep, d, n, nullptr, Ah, lane, act);
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
    int j0 = 1 << 20, j1 = 1 << 20;
    if (__ballot(search) == 0ull) nsrc = 0;  // no lane of the wavefront searches
    for (int k = 0; k < nsrc; ++k) {
        const LTerms<0> T = lane_cost<TAPS, WT>(Ss[k], pr, rstep, d, n, Ah, lane, search);
        // (selects rather than branches keep the pair in registers)
        const bool first = T.ok && lane_before(T.ncc, k, b0, j0);
        const bool second = T.ok && !first && lane_before(T.ncc, k, b1, j1);
        b1 = first ? b0 : (second ? T.ncc : b1);
        j1 = first ? j0 : (second ? k : j1);
        b0 = first ? T.ncc : b0;
        j0 = first ? k : j0;
    }
    i0 = j0;
    i1 = j1;
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
    const LTerms<3> TA = lane_terms<TAPS, WT>(A, pr, rstep, d, n, dirs, Ah, lane);
    const LTerms<3> TB = lane_terms<TAPS, WT>(B, pr, rstep, d, n, dirs, Ah, lane);
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

// Waves per SIMD the refinement kernel is built for, with each tap row's
// gathers issued together (lane_pass): three up to 7x7 windows (77-view 7x7
// refinement 192 -> 164 ms against two), two for 11x11 on 8-bit views (at
// three its registers spill: 365 -> 370 ms); the 11x11 window on f32 views,
// which needs more registers than two waves have, keeps the compiler's own
// choice. (Left to itself the compiler aimed at three waves and issued the
// Gauss-Newton steps' gathers one at a time.)
template <int TAPS, class WT>
constexpr int kRefineWaves = TAPS <= 7 ? 3 : (std::is_same<WT, float>::value ? 0 : 2);
template <int TAPS, class WT>
constexpr int kRefineWavesMin = kRefineWaves<TAPS, WT> > 0 ? kRefineWaves<TAPS, WT> : 1;
template <int TAPS, class WT>
constexpr int kRefineWavesMax = kRefineWaves<TAPS, WT> > 0 ? kRefineWaves<TAPS, WT> : 8;

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
    // best pair at the PatchMatch plane: the one the sweep stored with the
    // plane (its two lowest source costs), else every source in step
    int i0, i1;
    const bool want = have && pr.ok;
    const uint32_t sp = want ? (uint32_t)job.pair[pix] : 0xffu;
    const int q0 = (int)(sp & 15u), q1 = (int)(sp >> 4);
    const bool stored = sp != 0xffu && q0 < job.nsrc && q1 < job.nsrc && q0 != q1;
    lane_best_pair<TAPS, WT>(Ss, job.nsrc, pr, rstep, hp.x, nrm, Ah, lane, want && !stored, i0,
                             i1);
    if (stored) {
        i0 = q0;
        i1 = q1;
    }
    bool live = want && i1 < job.nsrc;
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
    if (have && acc) {
        job.hyp[pix] = hp;
        job.pair[pix] = 0xff;  // (the refined plane's pair is not known)
    }
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
