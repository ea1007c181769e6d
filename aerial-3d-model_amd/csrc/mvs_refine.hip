// NCC backward pass and Gauss-Newton refinement of PatchMatch hypotheses
// (gfx950). No reference counterpart: the reference's depth stage stops after
// PatchMatch (crates/stitch-core/src/recon/dense.rs:468-477). This refines
// depth and normal of every pixel PatchMatch accepted, on the cost the
// reference optimises (1 - NCC averaged over the best two sources,
// dense.rs:255-262), with the two sources fixed at the PatchMatch optimum.
//
// Two lanes per pixel, one per source of the pair. A lane walks the window
// once per evaluation, recomputing its warped samples (nothing is stored), and
// accumulates shifted single-pass sums of the samples s_k, of their
// derivatives D_k = ds_k/dtheta and of the products needed below. With a_k the
// centred reference grey and b_k the centred samples:
//   NCC = a.b / (|a| |b|),   dNCC/ds_k = (a_k/|a| - NCC b_k/|b|) / |b|,
// and with the Gauss-Newton residual r = a/|a| - b/|b| (so 1 - NCC = |r|^2/2):
//   J^T J = D^T P D / |b|^2,  P = I - 11^T/N - b b^T/|b|^2,
//   J^T r = -dNCC/dtheta.
// Sums are shifted by the reference window mean so that the f32 variance of
// a bright, low-contrast window does not cancel.
#include "mvs_device.h"

namespace amvs {
namespace {
using namespace dev;

constexpr int pack(int n) { return n * (n + 1) / 2; }

// One source's NCC terms. NDIR derivative directions: the first is depth, the
// others are normal directions; NDIR = 0 evaluates the NCC only.
template <int NDIR>
struct Terms {
    float ncc = 0.0f, nb = 0.0f;          // NCC and |b|
    float Dta[NDIR > 0 ? NDIR : 1];        // sum a_k D_k
    float Dtb[NDIR > 0 ? NDIR : 1];        // sum b_k D_k
    float DD[NDIR > 0 ? pack(NDIR) : 1];  // D^T P D, packed upper (when asked for)
    bool ok = false;
};

// Reference window of one pixel: centre ray, mean grey and |a|.
struct PixelRef {
    float rc0 = 0.0f, rc1 = 0.0f, mr = 0.0f, na = 0.0f;
    bool ok = false;
};

template <int TAPS>
__device__ PixelRef pixel_ref(const RefJob &job, int x, int y) {
    constexpr int N = TAPS * TAPS;
    PixelRef p;
    const int w = job.w, st = job.stride, half = job.half;
    if (x < half || y < half || x >= w - half || y >= job.h - half) return p;
    const int64_t pidx = (int64_t)y * w + x;
    float s = 0.0f;
    for (int j = 0; j < N; ++j) {
        const int dy = (j / TAPS - TAPS / 2) * st, dx = (j % TAPS - TAPS / 2) * st;
        s += job.gray[pidx + (int64_t)dy * w + dx];
    }
    p.mr = s / (float)N;
    float v = 0.0f;
    for (int j = 0; j < N; ++j) {
        const int dy = (j / TAPS - TAPS / 2) * st, dx = (j % TAPS - TAPS / 2) * st;
        const float a = job.gray[pidx + (int64_t)dy * w + dx] - p.mr;
        v += a * a;
    }
    p.na = sqrtf(v);
    const float2 rc = job.rays[pidx];
    p.rc0 = rc.x;
    p.rc1 = rc.y;
    p.ok = p.na > 1e-6f;
    return p;
}

// NCC terms of source S for the plane (d, n) at pixel (x, y); dirs holds
// NDIR - 1 unit directions in normal space.
template <int TAPS, bool PINHOLE, int NDIR, bool WANT_DD>
__device__ Terms<NDIR> source_terms(const RefJob &job, const SrcCam &S, const PixelRef &pr,
                                    int x, int y, float d, const float n[3],
                                    const float (*dirs)[3]) {
    constexpr int N = TAPS * TAPS;
    constexpr int ND = NDIR > 0 ? NDIR : 1;
    constexpr int NP = NDIR > 0 ? pack(NDIR) : 1;
    Terms<NDIR> T;
    const int w = job.w, st = job.stride;
    const int64_t pidx = (int64_t)y * w + x;
    const float nc = n[0] * pr.rc0 + n[1] * pr.rc1 + n[2];
    const float cc = d * nc;
    const float ch[3] = {pr.rc0, pr.rc1, 1.0f};
    const ff FF{S.f, S.f_lo}, FCX{S.cx, S.cx_lo}, FCY{S.cy, S.cy_lo}, FK1{S.k1, S.k1_lo},
        FK2{S.k2, S.k2_lo};
    const int sw = S.w;
    const float wm1 = (float)(S.w - 1), hm1 = (float)(S.h - 1);
    float s1 = 0.0f, s2 = 0.0f, sa = 0.0f;
    float sD[ND], sbD[ND], saD[ND], sDD[NP];
#pragma unroll
    for (int q = 0; q < ND; ++q) sD[q] = sbD[q] = saD[q] = 0.0f;
#pragma unroll
    for (int q = 0; q < NP; ++q) sDD[q] = 0.0f;
    for (int j = 0; j < N; ++j) {
        const int dy = (j / TAPS - TAPS / 2) * st, dx = (j % TAPS - TAPS / 2) * st;
        const int64_t qi = pidx + (int64_t)dy * w + dx;
        const float2 ray = job.rays[qi];
        const float a = job.gray[qi] - pr.mr;
        const float rk[3] = {ray.x, ray.y, 1.0f};
        const float den = n[0] * rk[0] + n[1] * rk[1] + n[2];
        if (!(fabsf(den) >= 1e-6f)) return T;
        // hardware reciprocals: the refinement has no reference rounding to keep
        const float rden = __builtin_amdgcn_rcpf(den);
        const float dq = cc * rden;
        if (!(dq > 0.0f)) return T;
        // w = R r (the ray in the source frame); the window point is dq w + t
        float xs[3], wv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            wv[i] = fmaf(S.R[i * 3], rk[0], fmaf(S.R[i * 3 + 1], rk[1], S.R[i * 3 + 2]));
            xs[i] = fmaf(dq, wv[i], S.t[i]);
        }
        if (!(xs[2] > 0.0f)) return T;
        const float inv = __builtin_amdgcn_rcpf(xs[2]);
        const float q0 = xs[0] * inv, q1 = xs[1] * inv;
        float u, v;
        if (PINHOLE) {
            u = fmaf(S.f, q0, S.cx);
            v = fmaf(S.f, q1, S.cy);
        } else {
            project_general(q0, q1, FF, FCX, FCY, FK1, FK2, u, v);
        }
        const float fx = u - 0.5f, fy = v - 0.5f;
        if (!(fx >= 0.0f && fy >= 0.0f && fx < wm1 && fy < hm1)) return T;
        const int x0 = (int)fx, y0 = (int)fy;
        const float ax = fx - (float)x0, ay = fy - (float)y0;
        // the four corners in one 16-byte load from the corner-packed plane
        const f32x4 qv = as_global(S.quadf)[y0 * sw + x0];
        const float g00 = qv.x, g10 = qv.y, g01 = qv.z, g11 = qv.w;
        const float top = fmaf(ax, g10 - g00, g00), bot = fmaf(ax, g11 - g01, g01);
        const float sv = fmaf(ay, bot - top, top) - pr.mr;  // shifted sample
        s1 += sv;
        s2 = fmaf(sv, sv, s2);
        sa = fmaf(a, sv, sa);
        if constexpr (NDIR > 0) {
            // chain: d s / d dq = grad I . d(u,v)/d(q0,q1) . d(q0,q1)/d dq
            float ux, uy, vx, vy;
            if (PINHOLE) {
                ux = vy = S.f;
                uy = vx = 0.0f;
            } else {
                const float r2 = q0 * q0 + q1 * q1;
                const float D = 1.0f + S.k1 * r2 + S.k2 * r2 * r2;
                const float dD = 2.0f * S.k1 + 4.0f * S.k2 * r2;
                ux = S.f * (D + q0 * dD * q0);
                uy = S.f * q0 * dD * q1;
                vx = S.f * q1 * dD * q0;
                vy = S.f * (D + q1 * dD * q1);
            }
            const float Iu = fmaf(ay, (g11 - g01) - (g10 - g00), g10 - g00);
            const float Iv = bot - top;
            const float dq0 = (wv[0] - q0 * wv[2]) * inv, dq1 = (wv[1] - q1 * wv[2]) * inv;
            const float dI = Iu * (ux * dq0 + uy * dq1) + Iv * (vx * dq0 + vy * dq1);
            // d dq / d theta: depth, then the normal directions
            float Dk[ND];
            Dk[0] = dI * (nc * rden);
#pragma unroll
            for (int q = 1; q < NDIR; ++q) {
                float g = 0.0f;
#pragma unroll
                for (int c3 = 0; c3 < 3; ++c3) g += (d * ch[c3] - dq * rk[c3]) * dirs[q - 1][c3];
                Dk[q] = dI * (g * rden);
            }
#pragma unroll
            for (int q = 0; q < NDIR; ++q) {
                sD[q] += Dk[q];
                sbD[q] = fmaf(sv, Dk[q], sbD[q]);
                saD[q] = fmaf(a, Dk[q], saD[q]);
            }
            if constexpr (WANT_DD) {
                int m = 0;
#pragma unroll
                for (int p = 0; p < NDIR; ++p)
#pragma unroll
                    for (int q = p; q < NDIR; ++q, ++m) sDD[m] = fmaf(Dk[p], Dk[q], sDD[m]);
            }
        }
    }
    const float ms = s1 / (float)N;
    const float nb2 = s2 - ms * s1;
    if (!(nb2 > 1e-12f)) return T;
    T.nb = sqrtf(nb2);
    T.ncc = sa / (pr.na * T.nb);
    if constexpr (NDIR > 0) {
#pragma unroll
        for (int q = 0; q < NDIR; ++q) {
            T.Dta[q] = saD[q];
            T.Dtb[q] = sbD[q] - ms * sD[q];
        }
        if constexpr (WANT_DD) {
            int m = 0;
#pragma unroll
            for (int p = 0; p < NDIR; ++p)
#pragma unroll
                for (int q = p; q < NDIR; ++q, ++m)
                    T.DD[m] = sDD[m] - sD[p] * sD[q] / (float)N - T.Dtb[p] * T.Dtb[q] / nb2;
        }
    }
    T.ok = true;
    return T;
}

// a ranks before b: higher NCC, then lower source index (the order of a
// stable descending sort, dense.rs:255-256).
__device__ inline bool before(float na, int ia, float nb, int ib) {
    return na > nb || (na == nb && ia < ib);
}

__device__ inline void push2(float v, int i, float &b0, int &i0, float &b1, int &i1) {
    if (before(v, i, b0, i0)) {
        b1 = b0;
        i1 = i0;
        b0 = v;
        i0 = i;
    } else if (before(v, i, b1, i1)) {
        b1 = v;
        i1 = i;
    }
}

// The two best sources for plane (d, n) at a pixel, found by the lane pair
// (lane `side` scans sources side, side + 2, ...). Both lanes must call it.
// Returns (-1, -1) unless at least two sources give a defined NCC.
template <int TAPS, bool PINHOLE>
__device__ int2 best_pair(const RefJob &job, const PixelRef &pr, bool live, int x, int y,
                          float d, const float n[3], int side) {
    float b0 = -INFINITY, b1 = -INFINITY;
    int i0 = 1 << 20, i1 = 1 << 20;
    if (live) {
        for (int s = side; s < job.nsrc; s += 2) {
            const Terms<0> T = source_terms<TAPS, PINHOLE, 0, false>(job, job.srcs[s], pr, x, y,
                                                                     d, n, nullptr);
            if (T.ok) push2(T.ncc, s, b0, i0, b1, i1);
        }
    }
    const float o0 = __shfl_xor(b0, 1), o1 = __shfl_xor(b1, 1);
    const int j0 = __shfl_xor(i0, 1), j1 = __shfl_xor(i1, 1);
    push2(o0, j0, b0, i0, b1, i1);
    push2(o1, j1, b0, i0, b1, i1);
    if (i1 >= job.nsrc) return make_int2(-1, -1);
    return make_int2(i0, i1);
}

// Fixed-pair cost and gradient of an explicit list: the forward/backward
// pair. Two lanes per item. pairs: two source indices per item; an item given
// (-1, -1) gets its best pair at that plane, written back. grad (nullable) =
// grad_out (nullable: ones) * d cost / d(d, n0, n1, n2); cost 2 and gradient
// 0 where the cost is undefined.
template <int TAPS, bool PINHOLE>
__global__ __launch_bounds__(64) void k_pair_list(RefJob job, int64_t n, const int32_t *px,
                                                  const int32_t *py, const float4 *hyp,
                                                  int32_t *pairs, const float *gout, float *cost,
                                                  float4 *grad) {
    const int lane = threadIdx.x, side = lane & 1;
    const int64_t item = (int64_t)blockIdx.x * 32 + (lane >> 1);
    const bool have = item < n;
    const int x = have ? px[item] : 0, y = have ? py[item] : 0;
    const float4 hp = have ? hyp[item] : make_float4(1, 0, 0, -1);
    const float nn[3] = {hp.y, hp.z, hp.w};
    PixelRef pr;
    if (have) pr = pixel_ref<TAPS>(job, x, y);
    int2 pq = have ? make_int2(pairs[2 * item], pairs[2 * item + 1]) : make_int2(0, 0);
    const bool choose = pq.x < 0;  // the same for both lanes of an item
    const int2 best =
        best_pair<TAPS, PINHOLE>(job, pr, have && pr.ok && choose, x, y, hp.x, nn, side);
    if (choose) pq = best;
    const int s = side ? pq.y : pq.x;
    float ncc = 0.0f, g4[4] = {0, 0, 0, 0};
    bool ok = false;
    if (have && pr.ok && s >= 0 && s < job.nsrc && pq.x != pq.y) {
        const float dirs[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        const Terms<4> T =
            source_terms<TAPS, PINHOLE, 4, false>(job, job.srcs[s], pr, x, y, hp.x, nn, dirs);
        if (T.ok) {
            ok = true;
            ncc = T.ncc;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                g4[q] = (T.Dta[q] / pr.na - T.ncc * T.Dtb[q] / T.nb) / T.nb;
        }
    }
    const float ncc_o = __shfl_xor(ncc, 1);
    const bool ok_o = __shfl_xor((int)ok, 1) != 0;
    float go[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) go[q] = __shfl_xor(g4[q], 1);
    if (have && side == 0) {
        const bool both = ok && ok_o;
        cost[item] = both ? 1.0f - 0.5f * (ncc + ncc_o) : 2.0f;
        if (choose) {
            pairs[2 * item] = pq.x;
            pairs[2 * item + 1] = pq.y;
        }
        if (grad) {
            const float sc = both ? -0.5f * (gout ? gout[item] : 1.0f) : 0.0f;
            grad[item] = make_float4(sc * (g4[0] + go[0]), sc * (g4[1] + go[1]),
                                     sc * (g4[2] + go[2]), sc * (g4[3] + go[3]));
        }
    }
}

// This lane's source terms at the plane (d, n), summed with its partner
// lane's (the other source of the pair): J^T J (packed upper) and J^T r of the
// pair's stacked residuals w.r.t. depth and the angles along dirs, and the
// pair cost. Both lanes of a pair must call it. true: both sources defined.
template <int TAPS, bool PINHOLE>
__device__ bool pair_gn(const RefJob &job, const SrcCam &S, const PixelRef &pr, bool live, int x,
                        int y, float d, const float n[3], const float dirs[2][3], float H[6],
                        float g[3], float &cost) {
    Terms<3> T;
    if (live) T = source_terms<TAPS, PINHOLE, 3, true>(job, S, pr, x, y, d, n, dirs);
#pragma unroll
    for (int q = 0; q < 6; ++q) H[q] = T.ok ? T.DD[q] / (T.nb * T.nb) : 0.0f;
#pragma unroll
    for (int q = 0; q < 3; ++q)
        g[q] = T.ok ? -(T.Dta[q] / pr.na - T.ncc * T.Dtb[q] / T.nb) / T.nb : 0.0f;
    const float ncc = T.ok ? T.ncc : 0.0f;
    const bool ok_o = __shfl_xor((int)T.ok, 1) != 0;
    const float ncc_o = __shfl_xor(ncc, 1);
#pragma unroll
    for (int q = 0; q < 6; ++q) H[q] += __shfl_xor(H[q], 1);
#pragma unroll
    for (int q = 0; q < 3; ++q) g[q] += __shfl_xor(g[q], 1);
    cost = 1.0f - 0.5f * (ncc + ncc_o);
    return T.ok && ok_o;
}

// The Gauss-Newton terms k_refine forms, for an explicit list of (pixel,
// plane, source pair) items: two lanes per item, the refinement's own pair
// search and terms, mapped to (d, n0, n1, n2). pairs: an item given (-1, -1)
// gets its best pair at that plane, written back. Undefined items: cost 2,
// gradient and J^T J 0.
template <int TAPS, bool PINHOLE>
__global__ __launch_bounds__(64) void k_terms_pair(RefJob job, int64_t n, const int32_t *px,
                                                   const int32_t *py, const float4 *hyp,
                                                   int32_t *pairs, float *cost, float4 *grad,
                                                   float *jtj) {
    const int lane = threadIdx.x, side = lane & 1;
    const int64_t item = (int64_t)blockIdx.x * 32 + (lane >> 1);
    const bool have = item < n;
    const int x = have ? px[item] : 0, y = have ? py[item] : 0;
    const float4 hp = have ? hyp[item] : make_float4(1, 0, 0, -1);
    const float nn[3] = {hp.y, hp.z, hp.w};
    PixelRef pr;
    if (have) pr = pixel_ref<TAPS>(job, x, y);
    int2 pq = have ? make_int2(pairs[2 * item], pairs[2 * item + 1]) : make_int2(0, 0);
    const bool choose = pq.x < 0;  // the same for both lanes of an item
    const int2 best =
        best_pair<TAPS, PINHOLE>(job, pr, have && pr.ok && choose, x, y, hp.x, nn, side);
    if (choose) pq = best;
    const int s = side ? pq.y : pq.x;
    const bool live = have && pr.ok && s >= 0 && s < job.nsrc && pq.x != pq.y;
    float dirs[2][3];
    tangent_dirs(nn, dirs);
    float H[6], g[3], c;
    const bool ok = pair_gn<TAPS, PINHOLE>(job, job.srcs[live ? s : 0], pr, live, x, y, hp.x, nn,
                                           dirs, H, g, c);
    if (!have || side != 0) return;
    if (choose) {
        pairs[2 * item] = pq.x;
        pairs[2 * item + 1] = pq.y;
    }
    float g4[4] = {0, 0, 0, 0}, h10[10];
#pragma unroll
    for (int m = 0; m < 10; ++m) h10[m] = 0.0f;
    if (ok) terms_to_dn(H, g, dirs, g4, h10);
    cost[item] = ok ? c : 2.0f;
    if (grad) grad[item] = make_float4(g4[0], g4[1], g4[2], g4[3]);
    if (jtj)
#pragma unroll
        for (int m = 0; m < 10; ++m) jtj[10 * item + m] = h10[m];
}

// Solve the symmetric positive 3x3 system H x = b (H packed upper) by
// cofactors.
__device__ bool solve3(const float H[6], const float b[3], float x[3]) {
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

// Levenberg-Marquardt-damped Gauss-Newton on (depth, two normal angles) of
// every pixel whose PatchMatch cost is below the output threshold, on the
// pair of sources that is best at its PatchMatch plane. A step that does not
// lower the pair cost is rejected and the damping grows. Hypotheses are
// updated in place; the PatchMatch cost (the output mask) is kept. Two lanes
// per pixel, 32 pixels per wavefront, over the job's textured-pixel lists.
// counters (nullable): [0] += accepted steps, [1] += pixels refined.
template <int TAPS, bool PINHOLE>
__global__ __launch_bounds__(64) void k_refine(RefJob job, int32_t steps,
                                               unsigned long long *counters) {
    const int lane = threadIdx.x, side = lane & 1;
    const int64_t c0 = job.counts[0], total = c0 + job.counts[1];
    if ((int64_t)blockIdx.x * 32 >= total) return;
    const int64_t item = (int64_t)blockIdx.x * 32 + (lane >> 1);
    bool have = item < total;
    int32_t pix = 0;
    if (have) pix = item < c0 ? job.list[item] : job.list[job.list_cap + (item - c0)];
    const int y = pix / job.w, x = pix - y * job.w;
    float4 hp = have ? job.hyp[pix] : make_float4(1, 0, 0, -1);
    have = have && job.cost[pix] < job.cost_max;
    PixelRef pr;
    if (have) pr = pixel_ref<TAPS>(job, x, y);
    float nrm[3] = {hp.y, hp.z, hp.w};
    const int2 pq = best_pair<TAPS, PINHOLE>(job, pr, have && pr.ok, x, y, hp.x, nrm, side);
    const int s = side ? pq.y : pq.x;
    bool live = have && pr.ok && pq.x >= 0;
    const bool refined = live;
    const SrcCam &S = job.srcs[live ? s : 0];
    const float lo = job.dmin * 0.7f, hi = job.dmax * 1.4f;
    float lambda = 1e-3f;
    unsigned long long acc = 0;
    for (int it = 0; it < steps; ++it) {
        float dirs[2][3];
        tangent_dirs(nrm, dirs);
        // sum of J^T J and of the NCC gradients over the pair
        float H[6], g[3], cost0;
        live = pair_gn<TAPS, PINHOLE>(job, S, pr, live, x, y, hp.x, nrm, dirs, H, g, cost0);
        const float Hd[6] = {H[0] * (1.0f + lambda), H[1], H[2], H[3] * (1.0f + lambda), H[4],
                             H[5] * (1.0f + lambda)};
        const float rhs[3] = {-g[0], -g[1], -g[2]};
        float dx[3] = {0.0f, 0.0f, 0.0f};
        const bool solved = live && solve3(Hd, rhs, dx);
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
        const bool cand = solved && (n1[0] * pr.rc0 + n1[1] * pr.rc1 + n1[2] < 0.0f) &&
                          d1 >= lo && d1 <= hi;
        float ncc1 = 0.0f;
        bool ok1 = false;
        if (cand) {
            const Terms<0> T1 =
                source_terms<TAPS, PINHOLE, 0, false>(job, S, pr, x, y, d1, n1, nullptr);
            ok1 = T1.ok;
            ncc1 = T1.ok ? T1.ncc : 0.0f;
        }
        const bool ok1_o = __shfl_xor((int)ok1, 1) != 0;
        const float ncc1_o = __shfl_xor(ncc1, 1);
        const float cost1 = 1.0f - 0.5f * (ncc1 + ncc1_o);
        if (cand && ok1 && ok1_o && cost1 < cost0) {
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
    if (have && side == 0 && acc) job.hyp[pix] = hp;
    if (counters) {
        unsigned long long e = side == 0 ? acc : 0, p = (side == 0 && refined) ? 1 : 0;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            e += __shfl_xor(e, off);
            p += __shfl_xor(p, off);
        }
        if (lane == 0 && e) atomicAdd(&counters[0], e);
        if (lane == 0 && p) atomicAdd(&counters[1], p);
    }
}

template <int TAPS>
void refine_taps(const RefJob &job, int32_t steps, unsigned long long *counters,
                 hipStream_t st) {
    const unsigned blocks = (unsigned)((2 * (int64_t)job.list_cap + 31) / 32);
    if (job.pinhole)
        hipLaunchKernelGGL((k_refine<TAPS, true>), dim3(blocks), dim3(64), 0, st, job, steps,
                           counters);
    else
        hipLaunchKernelGGL((k_refine<TAPS, false>), dim3(blocks), dim3(64), 0, st, job, steps,
                           counters);
}

template <int TAPS>
void pair_list_taps(const RefJob &job, int64_t n, const int32_t *px, const int32_t *py,
                    const float4 *hyp, int32_t *pairs, const float *gout, float *cost,
                    float4 *grad, hipStream_t st) {
    const unsigned blocks = (unsigned)((n + 31) / 32);
    if (blocks == 0) return;
    if (job.pinhole)
        hipLaunchKernelGGL((k_pair_list<TAPS, true>), dim3(blocks), dim3(64), 0, st, job, n, px,
                           py, hyp, pairs, gout, cost, grad);
    else
        hipLaunchKernelGGL((k_pair_list<TAPS, false>), dim3(blocks), dim3(64), 0, st, job, n, px,
                           py, hyp, pairs, gout, cost, grad);
}

template <int TAPS>
void terms_pair_taps(const RefJob &job, int64_t n, const int32_t *px, const int32_t *py,
                     const float4 *hyp, int32_t *pairs, float *cost, float4 *grad, float *jtj,
                     hipStream_t st) {
    const unsigned blocks = (unsigned)((n + 31) / 32);
    if (blocks == 0) return;
    if (job.pinhole)
        hipLaunchKernelGGL((k_terms_pair<TAPS, true>), dim3(blocks), dim3(64), 0, st, job, n, px,
                           py, hyp, pairs, cost, grad, jtj);
    else
        hipLaunchKernelGGL((k_terms_pair<TAPS, false>), dim3(blocks), dim3(64), 0, st, job, n, px,
                           py, hyp, pairs, cost, grad, jtj);
}

}  // namespace

int launch_refine_terms(int32_t taps, const RefJob &job, int64_t n, const int32_t *px,
                        const int32_t *py, const float4 *hyp, int32_t *pairs, float *cost,
                        float4 *grad, float *jtj, hipStream_t st) {
    if (job.lane_ok) return launch_terms_lane(taps, job, n, px, py, hyp, pairs, cost, grad, jtj, st);
    switch (taps) {
        case 5: terms_pair_taps<5>(job, n, px, py, hyp, pairs, cost, grad, jtj, st); return 0;
        case 7: terms_pair_taps<7>(job, n, px, py, hyp, pairs, cost, grad, jtj, st); return 0;
        case 11: terms_pair_taps<11>(job, n, px, py, hyp, pairs, cost, grad, jtj, st); return 0;
        default: return -1;
    }
}

int launch_refine(int32_t taps, const RefJob &job, int32_t steps, unsigned long long *counters,
                  hipStream_t st) {
    if (job.lane_ok) return launch_refine_lane(taps, job, steps, counters, st);
    switch (taps) {
        case 5: refine_taps<5>(job, steps, counters, st); return 0;
        case 7: refine_taps<7>(job, steps, counters, st); return 0;
        case 11: refine_taps<11>(job, steps, counters, st); return 0;
        default: return -1;
    }
}

int launch_pair_list(int32_t taps, const RefJob &job, int64_t n, const int32_t *px,
                     const int32_t *py, const float4 *hyp, int32_t *pairs, const float *gout,
                     float *cost, float4 *grad, hipStream_t st) {
    switch (taps) {
        case 5: pair_list_taps<5>(job, n, px, py, hyp, pairs, gout, cost, grad, st); return 0;
        case 7: pair_list_taps<7>(job, n, px, py, hyp, pairs, gout, cost, grad, st); return 0;
        case 11: pair_list_taps<11>(job, n, px, py, hyp, pairs, gout, cost, grad, st); return 0;
        default: return -1;
    }
}

}  // namespace amvs
