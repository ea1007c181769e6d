// gfx950 kernels of the dense-MVS depth engine.
//
// Arithmetic contract: this file is compiled with -ffp-contract=off and IEEE
// f32 division/sqrt, so every expression rounds exactly as the reference's Rust
// code (crates/stitch-core/src/recon/dense.rs), which never fuses a*b+c and
// evaluates left to right. The only fused multiply-adds are the explicit fmaf()
// calls of the projection, where the reference computes in f64 and one rounding
// to f32 reproduces its result (see project_pinhole).
//
// Work decomposition of the PatchMatch kernels: a workgroup is one 64-lane
// wavefront; LANES consecutive lanes share one reference pixel and lane l owns
// source view l (one lane per source, LANES = next power of two >= K). For each
// hypothesis the pixel's lanes first split the window taps to build the
// plane-induced 3D points (stored once in LDS, read by every source lane as a
// broadcast), then each lane warps the window into its own source, samples the
// f32 grey plane bilinearly (L1/L2-cached gathers), and forms the two-pass NCC
// with sums in tap order. The best-two selection over sources is an exact
// min/min reduction over the group's lanes.
#include "mvs_internal.h"

#include <math.h>

namespace amvs {

namespace {

// ---------------------------------------------------------------------------
// scalar pieces of the reference
// ---------------------------------------------------------------------------

// dense.rs:273-279
__device__ __forceinline__ float rnd_next(uint64_t &s) {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return (float)(uint32_t)(s >> 40) * (1.0f / 16777216.0f);  // exact: / 2^24
}

// dense.rs:374
__device__ __forceinline__ uint64_t seed_init(uint64_t i) {
    return (i * 0x9E3779B97F4A7C15ull) | 1ull;
}

// dense.rs:446-448
__device__ __forceinline__ uint64_t seed_refine(uint64_t i, uint64_t it, uint64_t parity) {
    return (((i << 8) ^ (it * 2ull + parity)) * 0x2545F4914F6CDD1Dull) | 1ull;
}

// dense.rs:281-297, always called with Some(around).
__device__ __forceinline__ void random_normal(uint64_t &s, float r0, float r1, float b0, float b1,
                                              float b2, float spread, float &o0, float &o1,
                                              float &o2) {
    float t;
    t = (rnd_next(s) - 0.5f) * 2.0f;
    float n0 = b0 + t * spread;
    t = (rnd_next(s) - 0.5f) * 2.0f;
    float n1 = b1 + t * spread;
    t = (rnd_next(s) - 0.5f) * 2.0f;
    float n2 = b2 + t * spread;
    float facing = n0 * r0 + n1 * r1;
    facing = facing + n2;
    if (facing > 0.0f) {
        n0 = -n0;
        n1 = -n1;
        n2 = -n2;
    }
    float l2 = n0 * n0 + n1 * n1;
    l2 = l2 + n2 * n2;
    float l = fmaxf(sqrtf(l2), 1e-9f);
    o0 = n0 / l;
    o1 = n1 / l;
    o2 = n2 / l;
}

// a0/b and a1/b, each correctly rounded, sharing one reciprocal. These are the
// steps of the compiler's IEEE f32 division (reciprocal estimate, one
// Newton-Raphson refinement, two residual corrections) without the range
// scaling of v_div_scale/v_div_fixup, which is the identity for a normal
// divisor and a normal, finite quotient. Outside that range (denormal or
// overflowing quotients) the projected coordinate is off the image or equal to
// the principal point whichever way the quotient rounds; mvs_selftest checks
// the equality with `/` on the device.
__device__ __forceinline__ void div2_shared(const float a0, const float a1, const float b,
                                            float &q0, float &q1) {
    float r = __builtin_amdgcn_rcpf(b);
    const float e = fmaf(-b, r, 1.0f);
    r = fmaf(e, r, r);
    float q = a0 * r;
    float res = fmaf(-b, q, a0);
    q = fmaf(res, r, q);
    res = fmaf(-b, q, a0);
    q0 = fmaf(res, r, q);
    q = a1 * r;
    res = fmaf(-b, q, a1);
    q = fmaf(res, r, q);
    res = fmaf(-b, q, a1);
    q1 = fmaf(res, r, q);
}

// Global-address-space view of a device pointer (global_load rather than
// flat_load).
typedef const __attribute__((address_space(1))) float gfloat;
__device__ __forceinline__ gfloat *as_global(const float *p) { return (gfloat *)p; }

// float-float helpers for the general (distorted / non-f32 intrinsics) path
struct ff {
    float hi, lo;
};
__device__ __forceinline__ ff fast_two_sum(float a, float b) {
    float s = a + b;
    float e = b - (s - a);
    return {s, e};
}
__device__ __forceinline__ ff two_sum(float a, float b) {
    float s = a + b;
    float bb = s - a;
    float e = (a - (s - bb)) + (b - bb);
    return {s, e};
}
__device__ __forceinline__ ff two_prod(float a, float b) {
    float p = a * b;
    return {p, fmaf(a, b, -p)};
}
__device__ __forceinline__ ff ff_add(ff a, ff b) {
    ff s = two_sum(a.hi, b.hi);
    float e = s.lo + a.lo;
    e = e + b.lo;
    return fast_two_sum(s.hi, e);
}
__device__ __forceinline__ ff ff_mul(ff a, ff b) {
    ff p = two_prod(a.hi, b.hi);
    float c = a.hi * b.lo;
    c = c + a.lo * b.hi;
    return fast_two_sum(p.hi, p.lo + c);
}

// Projection of normalised (q0, q1) to source pixels. The reference computes
// geometry.rs:26-32 in f64 from the f32 quotient and casts to f32
// (dense.rs:231-234). With k1 = k2 = 0 its d is exactly 1 and f*x is exact in
// f64, so the result is RN32(f*x + cx) up to a double-rounding tie (f64 then
// f32) that needs |x| < 2^-5 and a 2^-28-rare bit pattern; one fmaf gives
// RN32 directly.
__device__ __forceinline__ void project_pinhole(const float q0, const float q1, float F,
                                                float CX, float CY, float &u, float &v) {
    u = fmaf(F, q0, CX);
    v = fmaf(F, q1, CY);
}

// General camera: distortion and intrinsics in float-float (about 44 bits),
// rounded once to f32.
__device__ __forceinline__ void project_general(const float q0, const float q1, const ff F,
                                               const ff CX, const ff CY, const ff K1,
                                               const ff K2, float &u, float &v) {
    ff r2 = ff_add(two_prod(q0, q0), two_prod(q1, q1));
    ff d = ff_add(ff{1.0f, 0.0f}, ff_mul(K1, r2));
    d = ff_add(d, ff_mul(ff_mul(K2, r2), r2));
    ff fd = ff_mul(F, d);
    ff ux = ff_add(ff_mul(fd, ff{q0, 0.0f}), CX);
    ff vy = ff_add(ff_mul(fd, ff{q1, 0.0f}), CY);
    u = ux.hi + ux.lo;
    v = vy.hi + vy.lo;
}

// ---------------------------------------------------------------------------
// per-view preparation
// ---------------------------------------------------------------------------

// dense.rs:108 for R = G = B.
__constant__ float c_gray_lut[256];

__global__ void k_gray_from_u8(const uint8_t *__restrict__ src, int32_t w, int32_t h,
                               int32_t stride, float *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)w * h) return;
    const int32_t y = (int32_t)(i / w), x = (int32_t)(i - (int64_t)y * w);
    dst[i] = c_gray_lut[src[(int64_t)y * stride + x]];
}

// dense.rs:343-349 with geometry.rs:34-45: f64 fixed point, cast to f32.
__global__ void k_rays(double f, double cx, double cy, double k1, double k2, int32_t w, int32_t h,
                       float2 *__restrict__ rays) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)w * h) return;
    const int64_t y = i / w, x = i - y * w;
    const double qx = ((double)x + 0.5 - cx) / f;
    const double qy = ((double)y + 0.5 - cy) / f;
    double px = qx, py = qy;
    for (int k = 0; k < 8; ++k) {
        const double r2 = px * px + py * py;
        double d = 1.0 + k1 * r2;
        d = d + (k2 * r2) * r2;
        px = qx / d;
        py = qy / d;
    }
    rays[i] = make_float2((float)px, (float)py);
}

// Initial hypotheses (dense.rs:365-381), texture mask (dense.rs:383-402) and
// the per-parity lists of textured pixels the sweeps iterate over.
__global__ __launch_bounds__(256) void k_init(RefJob job, int32_t taps, int32_t write_hyp) {
    const int32_t w = job.w, h = job.h;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < (int64_t)w * h;
    const int32_t y = valid ? (int32_t)(i / w) : 0;
    const int32_t x = valid ? (int32_t)(i - (int64_t)y * w) : 0;
    bool tex = false;
    if (valid) {
        if (write_hyp) {
            uint64_t s = seed_init((uint64_t)i);
            const float r = rnd_next(s);
            const float dd = job.imin + r * (job.imax - job.imin);
            const float d = 1.0f / dd;
            const float2 ray = job.rays[i];
            float n0, n1, n2;
            random_normal(s, ray.x, ray.y, job.up0, job.up1, job.up2, 0.3f, n0, n1, n2);
            job.hyp[i] = make_float4(d, n0, n1, n2);
            job.cost[i] = 2.0f;  // untextured pixels: their cost is 2.0 (see below)
        }
        const int32_t half = job.half;
        if (!(x < half || y < half || x >= w - half || y >= h - half)) {
            const int32_t r = taps / 2, st = job.stride;
            const float n = (float)(taps * taps);
            float sum = 0.0f;
            for (int32_t a = -r; a <= r; ++a)
                for (int32_t b = -r; b <= r; ++b)
                    sum = sum + job.gray[(int64_t)(y + a * st) * w + (x + b * st)];
            const float m = sum / n;
            float var = 0.0f;
            for (int32_t a = -r; a <= r; ++a)
                for (int32_t b = -r; b <= r; ++b) {
                    const float d = job.gray[(int64_t)(y + a * st) * w + (x + b * st)] - m;
                    var = var + d * d;
                }
            var = var / n;
            tex = sqrtf(var) >= job.texture_min;
        }
        job.textured[i] = tex ? 1 : 0;
    }
    // Wave-aggregated append to the parity lists.
    const int parity = (x + y) & 1;
    const uint64_t m0 = __ballot(valid && tex && parity == 0);
    const uint64_t m1 = __ballot(valid && tex && parity == 1);
    const int lane = threadIdx.x & 63;
    int b0 = 0, b1 = 0;
    if (lane == 0) {
        if (m0) b0 = atomicAdd(&job.counts[0], __popcll(m0));
        if (m1) b1 = atomicAdd(&job.counts[1], __popcll(m1));
    }
    b0 = __shfl(b0, 0);
    b1 = __shfl(b1, 0);
    if (valid && tex) {
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        const uint64_t m = parity ? m1 : m0;
        const int pos = (parity ? b1 : b0) + __popcll(m & below);
        job.list[(int64_t)parity * job.list_cap + pos] = (int32_t)i;
    }
}

// ---------------------------------------------------------------------------
// PatchMatch cost evaluation (dense.rs:181-263) and sweeps (dense.rs:413-467)
// ---------------------------------------------------------------------------

enum { MODE_EVAL_CURRENT = 0, MODE_SWEEP = 1, MODE_COST_LIST = 2 };

template <int TAPS, int LANES, bool PINHOLE, int MODE>
__global__ __launch_bounds__(64) void k_patchmatch(RefJob job, int32_t it, int32_t parity,
                                                   float range, unsigned long long *evals,
                                                   const int32_t *__restrict__ px,
                                                   const int32_t *__restrict__ py,
                                                   const float4 *__restrict__ hyp_in,
                                                   float *__restrict__ out, int64_t n_items) {
    constexpr int N = TAPS * TAPS;
    constexpr int PIX = 64 / LANES;
    constexpr int OWN = (N + LANES - 1) / LANES;
    constexpr uint64_t GMASK = LANES == 64 ? ~0ull : ((1ull << LANES) - 1ull);
    __shared__ float4 Ps[PIX][N];

    const int lane = threadIdx.x;
    const int g = lane / LANES, l = lane % LANES;
    const int32_t w = job.w, h = job.h, K = job.nsrc, stride = job.stride, half = job.half;

    int64_t total;
    if (MODE == MODE_EVAL_CURRENT)
        total = (int64_t)job.counts[0] + job.counts[1];
    else if (MODE == MODE_SWEEP)
        total = job.counts[parity];
    else
        total = n_items;
    if ((int64_t)blockIdx.x * PIX >= total) return;  // uniform over the wave

    const int64_t item = (int64_t)blockIdx.x * PIX + g;
    const bool have = item < total;
    int32_t x = 0, y = 0;
    if (have) {
        int32_t pix;
        if (MODE == MODE_EVAL_CURRENT) {
            const int64_t c0 = job.counts[0];
            pix = item < c0 ? job.list[item] : job.list[job.list_cap + (item - c0)];
        } else if (MODE == MODE_SWEEP) {
            pix = job.list[(int64_t)parity * job.list_cap + item];
        } else {
            pix = -1;
            x = px[item];
            y = py[item];
        }
        if (pix >= 0) {
            y = pix / w;
            x = pix - y * w;
        }
    }
    const int64_t pidx = (int64_t)y * w + x;
    const bool inside = have && x >= half && y >= half && x < w - half && y < h - half;

    // Reference patch statistics (dense.rs:213, mean_sd :266-271), sums in
    // tap order; every lane of the group computes them (identical values).
    float mr = 0.0f, sr = 0.0f;
    if (inside) {
        float s = 0.0f;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const int dy = (j / TAPS - TAPS / 2) * stride, dx = (j % TAPS - TAPS / 2) * stride;
            s = s + job.gray[pidx + (int64_t)dy * w + dx];
        }
        mr = s / (float)N;
        float v = 0.0f;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const int dy = (j / TAPS - TAPS / 2) * stride, dx = (j % TAPS - TAPS / 2) * stride;
            const float d = job.gray[pidx + (int64_t)dy * w + dx] - mr;
            v = v + d * d;
        }
        v = v / (float)N;
        sr = sqrtf(v);
    }
    const bool pixel_ok = inside && !(sr < job.texture_min);

    // Taps this lane owns in the plane phase: their rays and centred grey.
    float2 oray[OWN];
    float orefc[OWN];
#pragma unroll
    for (int o = 0; o < OWN; ++o) {
        const int j = l + LANES * o;
        oray[o] = make_float2(0.0f, 0.0f);
        orefc[o] = 0.0f;
        if (j < N && pixel_ok) {
            const int dy = (j / TAPS - TAPS / 2) * stride, dx = (j % TAPS - TAPS / 2) * stride;
            const int64_t qi = pidx + (int64_t)dy * w + dx;
            oray[o] = job.rays[qi];
            orefc[o] = job.gray[qi] - mr;
        }
    }
    const float2 rc = pixel_ok ? job.rays[pidx] : make_float2(0.0f, 0.0f);

    // This lane's source view.
    const bool has_src = l < K;
    const SrcCam &S = job.srcs[has_src ? l : 0];
    const float R0 = S.R[0], R1 = S.R[1], R2 = S.R[2], R3 = S.R[3], R4 = S.R[4], R5 = S.R[5],
                R6 = S.R[6], R7 = S.R[7], R8 = S.R[8];
    const float T0 = S.t[0], T1 = S.t[1], T2 = S.t[2];
    const float F = S.f, CX = S.cx, CY = S.cy;
    const ff FF{S.f, S.f_lo}, FCX{S.cx, S.cx_lo}, FCY{S.cy, S.cy_lo}, FK1{S.k1, S.k1_lo},
        FK2{S.k2, S.k2_lo};
    const float *__restrict__ G = S.gray;
    const int32_t sw = S.w;
    const float wm1 = (float)(S.w - 1), hm1 = (float)(S.h - 1);

    int64_t n_evals = 0;

    // Cost of hypothesis (d, n) at this group's pixel; `act` must be uniform
    // over the group. Every lane of the wave calls this (it holds the ballot,
    // barriers and the cross-lane reduction).
    auto eval = [&](bool act, float d, float n0, float n1, float n2) -> float {
        act = act && pixel_ok;
        bool pf = false;
        if (act) {
            // dense.rs:187-211: plane through X = ray*d with normal n; depth
            // of each window ray on that plane.
            const float xp0 = rc.x * d, xp1 = rc.y * d;
            float cc = n0 * xp0 + n1 * xp1;
            cc = cc + n2 * d;
#pragma unroll
            for (int o = 0; o < OWN; ++o) {
                const int j = l + LANES * o;
                if (j < N) {
                    float den = n0 * oray[o].x + n1 * oray[o].y;
                    den = den + n2;
                    const float dq = cc / den;
                    pf = pf || (fabsf(den) < 1e-6f) || !(dq > 0.0f);
                    Ps[g][j] = make_float4(oray[o].x * dq, oray[o].y * dq, dq, orefc[o]);
                }
            }
        }
        const uint64_t fb = __ballot(pf);
        const bool gfail = ((fb >> (g * LANES)) & GMASK) != 0ull;
        __syncthreads();
        float cs = has_src ? 2.0f : INFINITY;
        if (act && !gfail && has_src) {
            // dense.rs:218-253 for source l.
            // The reference stops at a source's first failing sample
            // (dense.rs:227-240) and scores it 2.0. Here every sample is
            // computed without branches (a failing one reads a safe address)
            // and `ok` collects the failures, which yields the same cost while
            // letting the gathers of successive taps overlap.
            const float4 *P = Ps[g];
            float sv[N];
            float sum = 0.0f;
            bool ok = true;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                const float4 p = P[j];
                float xs0 = R0 * p.x + R1 * p.y;
                xs0 = xs0 + R2 * p.z;
                xs0 = xs0 + T0;
                float xs1 = R3 * p.x + R4 * p.y;
                xs1 = xs1 + R5 * p.z;
                xs1 = xs1 + T1;
                float xs2 = R6 * p.x + R7 * p.y;
                xs2 = xs2 + R8 * p.z;
                xs2 = xs2 + T2;
                float q0, q1;
                div2_shared(xs0, xs1, xs2, q0, q1);
                float u, v;
                if (PINHOLE)
                    project_pinhole(q0, q1, F, CX, CY, u, v);
                else
                    project_general(q0, q1, FF, FCX, FCY, FK1, FK2, u, v);
                // gray_at, dense.rs:69-85: None unless 0 <= fx < w-1 and
                // 0 <= fy < h-1 (NaN fails too)
                const float fx = u - 0.5f, fy = v - 0.5f;
                const bool good =
                    (xs2 > 0.0f) && fx >= 0.0f && fy >= 0.0f && fx < wm1 && fy < hm1;
                ok = ok && good;
                const float sx = good ? fx : 0.0f, sy = good ? fy : 0.0f;
                const int x0 = (int)sx, y0 = (int)sy;
                // fx >= 0: fract is fx - floor(fx) = fx - x0, exactly
                const float ax = __builtin_amdgcn_fractf(sx), ay = __builtin_amdgcn_fractf(sy);
                gfloat *gp = as_global(G) + (y0 * sw + x0);
                const float g00 = gp[0], g10 = gp[1], g01 = gp[sw], g11 = gp[sw + 1];
                const float bx = 1.0f - ax, by = 1.0f - ay;
                float top = g00 * bx;
                top = top + g10 * ax;
                float bot = g01 * bx;
                bot = bot + g11 * ax;
                float val = top * by;
                val = val + bot * ay;
                sv[j] = val;
                sum = sum + val;
            }
            if (ok) {
                const float ms = sum / (float)N;
                float var = 0.0f, cov = 0.0f;
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    const float dd = sv[j] - ms;
                    var = var + dd * dd;
                    cov = cov + P[j].w * dd;
                }
                var = var / (float)N;
                const float ss = sqrtf(var);
                if (!(ss < 1e-3f)) {
                    float den = (float)N * sr;
                    den = den * ss;
                    cs = 1.0f - cov / den;
                }
            }
        }
        // Best two of the K source costs (dense.rs:255-262): exact min/min
        // all-reduce over the group's lanes.
        float m1 = cs, m2 = INFINITY;
#pragma unroll
        for (int off = 1; off < LANES; off <<= 1) {
            const float o1 = __shfl_xor(m1, off), o2 = __shfl_xor(m2, off);
            const float lo = fminf(m1, o1);
            m2 = fminf(fmaxf(m1, o1), fminf(m2, o2));
            m1 = lo;
        }
        float c = K >= 2 ? (m1 + m2) * 0.5f : m1;
        if (gfail || !pixel_ok) c = 2.0f;
        if (act && l == 0) ++n_evals;
        __syncthreads();
        return c;
    };

    if (MODE == MODE_COST_LIST) {
        const float4 hp = have ? hyp_in[item] : make_float4(1.0f, 0.0f, 0.0f, -1.0f);
        const float c = eval(have, hp.x, hp.y, hp.z, hp.w);
        if (have && l == 0) out[item] = c;
    } else if (MODE == MODE_EVAL_CURRENT) {
        const float4 hp = have ? job.hyp[pidx] : make_float4(1.0f, 0.0f, 0.0f, -1.0f);
        const float c = eval(have, hp.x, hp.y, hp.z, hp.w);
        if (have && l == 0) job.cost[pidx] = c;
    } else {
        float4 best = have ? job.hyp[pidx] : make_float4(1.0f, 0.0f, 0.0f, -1.0f);
        float bcost = have ? job.cost[pidx] : 2.0f;
        // Ten proposals in the reference's order, one evaluation site:
        // k < 8: propagation from the other colour (dense.rs:427-444) in CAND
        // order (-1,0),(1,0),(0,-1),(0,1),(-5,0),(5,0),(0,-5),(0,5); those pixels
        // are not written during this half-sweep, so reading live state equals
        // reading the reference's snapshot (:416).
        // k = 8, 9: random refinement narrowing each iteration (:445-462).
        uint64_t s = seed_refine((uint64_t)pidx, (uint64_t)it, (uint64_t)parity);
        const float span = job.imax - job.imin;
#pragma unroll 1
        for (int k = 0; k < 10; ++k) {
            float4 hp;
            bool act;
            if (k < 8) {
                const int dir = k & 3, mag = k < 4 ? 1 : 5;
                const int nx = x + (dir == 0 ? -mag : (dir == 1 ? mag : 0));
                const int ny = y + (dir == 2 ? -mag : (dir == 3 ? mag : 0));
                const bool inb = have && nx >= 0 && ny >= 0 && nx < w && ny < h;
                hp = make_float4(1.0f, 0.0f, 0.0f, -1.0f);
                if (inb) hp = job.hyp[(int64_t)ny * w + nx];
                const bool dup =
                    hp.x == best.x && hp.y == best.y && hp.z == best.z && hp.w == best.w;
                act = inb && !dup;
            } else {
                float step = (rnd_next(s) - 0.5f) * range;
                step = step * 0.25f;
                step = step * span;
                const float id = 1.0f / best.x + step;
                hp.x = (id > job.imin * 0.5f && id < job.imax * 1.5f) ? 1.0f / id : best.x;
                random_normal(s, rc.x, rc.y, best.y, best.z, best.w, 0.25f * range, hp.y, hp.z,
                              hp.w);
                act = have;
            }
            const float c = eval(act, hp.x, hp.y, hp.z, hp.w);
            if (act && c < bcost) {
                best = hp;
                bcost = c;
            }
        }
        if (have && l == 0) {
            job.hyp[pidx] = best;
            job.cost[pidx] = bcost;
        }
    }

    if (evals) {
        unsigned long long e = (unsigned long long)n_evals;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) e += __shfl_xor(e, off);
        if (lane == 0 && e) atomicAdd(evals, e);
    }
}

// Device self-check of div2_shared against IEEE `/`: n hashed (a0, a1, b)
// triples, half drawn log-uniformly over [2^-40, 2^40] with random signs on
// the numerators, half from the projection's working range (|a| < 4000,
// b in [0.05, 2000]). Counts quotients that differ in any bit.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float unit24(uint64_t h) { return (float)(uint32_t)(h >> 40) * 0x1p-24f; }

__global__ void k_selftest_div(int64_t n, uint64_t seed, unsigned long long *bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long nbad = 0;
    if (i < n) {
        const uint64_t h0 = mix64(seed ^ (uint64_t)i), h1 = mix64(h0), h2 = mix64(h1),
                       h3 = mix64(h2);
        float a0, a1, b;
        if (i & 1) {
            a0 = exp2f(80.0f * unit24(h0) - 40.0f) * ((h3 & 1) ? -1.0f : 1.0f);
            a1 = exp2f(80.0f * unit24(h1) - 40.0f) * ((h3 & 2) ? -1.0f : 1.0f);
            b = exp2f(80.0f * unit24(h2) - 40.0f);
        } else {
            a0 = 8000.0f * unit24(h0) - 4000.0f;
            a1 = 8000.0f * unit24(h1) - 4000.0f;
            b = 0.05f + 2000.0f * unit24(h2);
        }
        float q0, q1;
        div2_shared(a0, a1, b, q0, q1);
        const float r0 = a0 / b, r1 = a1 / b;
        nbad = (__float_as_uint(q0) != __float_as_uint(r0)) + (__float_as_uint(q1) != __float_as_uint(r1));
    }
    if (nbad) atomicAdd(bad, nbad);
}

// dense.rs:468-477, plus the normals of the final hypotheses.
__global__ void k_threshold(RefJob job, float *__restrict__ depth, float *__restrict__ normal) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)job.w * job.h) return;
    const float4 hp = job.hyp[i];
    const float c = job.cost[i];
    const float lo = job.dmin * 0.7f, hi = job.dmax * 1.4f;
    depth[i] = (c < job.cost_max && hp.x >= lo && hp.x <= hi) ? hp.x : 0.0f;
    if (normal) {
        normal[3 * i] = hp.y;
        normal[3 * i + 1] = hp.z;
        normal[3 * i + 2] = hp.w;
    }
}

// Geometric consistency (dense.rs:585-615) in f64 like the reference's fusion.
__global__ void k_consistency(RefCam64 rf, const float *__restrict__ depth, int32_t w, int32_t h,
                              const NbrCam *__restrict__ nbrs, int32_t nn, int32_t need,
                              int32_t *__restrict__ agree, unsigned long long *n_ok) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool good = false;
    if (i < (int64_t)w * h) {
        const float di = depth[i];
        if (di <= 0.0f) {
            agree[i] = -1;
        } else {
            const int64_t y = i / w, x = i - y * w;
            const double qx = ((double)x + 0.5 - rf.cx) / rf.f;
            const double qy = ((double)y + 0.5 - rf.cy) / rf.f;
            double px = qx, py = qy;
            for (int k = 0; k < 8; ++k) {
                const double r2 = px * px + py * py;
                double d = 1.0 + rf.k1 * r2;
                d = d + (rf.k2 * r2) * r2;
                px = qx / d;
                py = qy / d;
            }
            const double dd = (double)di;
            const double xc0 = px * dd, xc1 = py * dd, xc2 = dd;
            double xw[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                double acc = rf.R[0 * 3 + r] * xc0;
                acc = acc + rf.R[1 * 3 + r] * xc1;
                acc = acc + rf.R[2 * 3 + r] * xc2;
                xw[r] = acc + rf.C[r];
            }
            int32_t a = 0;
            for (int32_t k = 0; k < nn; ++k) {
                const NbrCam &v = nbrs[k];
                const double e0 = xw[0] - v.C[0], e1 = xw[1] - v.C[1], e2 = xw[2] - v.C[2];
                double xs[3];
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    double acc = v.R[r * 3 + 0] * e0;
                    acc = acc + v.R[r * 3 + 1] * e1;
                    acc = acc + v.R[r * 3 + 2] * e2;
                    xs[r] = acc;
                }
                if (xs[2] <= 0.0) continue;
                const double ux = xs[0] / xs[2], uy = xs[1] / xs[2];
                const double r2 = ux * ux + uy * uy;
                double d = 1.0 + v.k1 * r2;
                d = d + (v.k2 * r2) * r2;
                const double fd = v.f * d;
                const double u0 = fd * ux + v.cx, u1 = fd * uy + v.cy;
                const double fx = floor(u0), fy = floor(u1);
                if (!(fx >= 0.0) || !(fy >= 0.0) || !(fx < (double)v.w) || !(fy < (double)v.h))
                    continue;
                const double dn = (double)v.depth[(int64_t)fy * v.w + (int64_t)fx];
                if (dn > 0.0 && fabs((dn - xs[2]) / xs[2]) < 0.01) ++a;
            }
            agree[i] = a;
            good = a >= need;
        }
    }
    const uint64_t m = __ballot(good);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(n_ok, (unsigned long long)__popcll(m));
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------

template <int TAPS, int LANES, bool PINHOLE, int MODE>
void launch_pm(const RefJob &job, int64_t items, int32_t it, int32_t parity, float range,
               unsigned long long *evals, const int32_t *px, const int32_t *py,
               const float4 *hyp_in, float *out, hipStream_t st) {
    constexpr int PIX = 64 / LANES;
    const int64_t blocks = (items + PIX - 1) / PIX;
    if (blocks <= 0) return;
    hipLaunchKernelGGL((k_patchmatch<TAPS, LANES, PINHOLE, MODE>), dim3((unsigned)blocks),
                       dim3(64), 0, st, job, it, parity, range, evals, px, py, hyp_in, out,
                       items);
}

template <int MODE>
int dispatch_pm(const RefJob &job, int32_t taps, int64_t items, int32_t it, int32_t parity,
                float range, unsigned long long *evals, const int32_t *px, const int32_t *py,
                const float4 *hyp_in, float *out, hipStream_t st, bool pinhole) {
    const int lanes = lanes_for(job.nsrc);
#define AMVS_CASE(T, L)                                                                        \
    if (taps == T && lanes == L) {                                                             \
        if (pinhole)                                                                           \
            launch_pm<T, L, true, MODE>(job, items, it, parity, range, evals, px, py, hyp_in,  \
                                        out, st);                                              \
        else                                                                                   \
            launch_pm<T, L, false, MODE>(job, items, it, parity, range, evals, px, py, hyp_in, \
                                         out, st);                                             \
        return 0;                                                                              \
    }
    AMVS_CASE(5, 4)
    AMVS_CASE(5, 8)
    AMVS_CASE(5, 16)
    AMVS_CASE(7, 4)
    AMVS_CASE(7, 8)
    AMVS_CASE(7, 16)
#undef AMVS_CASE
    return -1;
}

}  // namespace

int lanes_for(int32_t nsrc) {
    if (nsrc <= 4) return 4;
    if (nsrc <= 8) return 8;
    if (nsrc <= 16) return 16;
    return -1;
}

bool taps_supported(int32_t taps) { return taps == 5 || taps == 7; }

static bool g_lut_ready = false;

void launch_gray_from_u8(const uint8_t *src, int32_t w, int32_t h, int32_t stride, float *dst,
                         hipStream_t st) {
    if (!g_lut_ready) {
        float lut[256];
        for (int i = 0; i < 256; ++i) {
            // dense.rs:108 with R = G = B = i, f32, left to right
            float v = 0.299f * (float)i + 0.587f * (float)i;
            v = v + 0.114f * (float)i;
            lut[i] = v / 255.0f;
        }
        (void)hipMemcpyToSymbol(HIP_SYMBOL(c_gray_lut), lut, sizeof lut);
        g_lut_ready = true;
    }
    const int64_t n = (int64_t)w * h;
    hipLaunchKernelGGL(k_gray_from_u8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, w,
                       h, stride, dst);
}

void launch_rays(double f, double cx, double cy, double k1, double k2, int32_t w, int32_t h,
                 float2 *rays, hipStream_t st) {
    const int64_t n = (int64_t)w * h;
    hipLaunchKernelGGL(k_rays, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, f, cx, cy, k1,
                       k2, w, h, rays);
}

void launch_init(const RefJob &job, int32_t taps, int32_t write_hyp, hipStream_t st) {
    const int64_t n = (int64_t)job.w * job.h;
    hipLaunchKernelGGL(k_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, job, taps,
                       write_hyp);
}

int launch_eval_current(const RefJob &job, int32_t taps, unsigned long long *evals,
                        hipStream_t st) {
    return dispatch_pm<MODE_EVAL_CURRENT>(job, taps, (int64_t)job.w * job.h, 0, 0, 1.0f, evals,
                                          nullptr, nullptr, nullptr, nullptr, st, job.pinhole != 0);
}

int launch_sweep(const RefJob &job, int32_t taps, int32_t it, int32_t parity,
                 unsigned long long *evals, hipStream_t st) {
    const float range = ldexpf(1.0f, -it);  // 0.5f32.powi(it), dense.rs:414
    return dispatch_pm<MODE_SWEEP>(job, taps, job.list_cap, it, parity, range, evals, nullptr,
                                   nullptr, nullptr, nullptr, st, job.pinhole != 0);
}

int launch_cost_list(const RefJob &job, int32_t taps, int64_t n, const int32_t *px,
                     const int32_t *py, const float4 *hyp, float *out, hipStream_t st) {
    return dispatch_pm<MODE_COST_LIST>(job, taps, n, 0, 0, 1.0f, nullptr, px, py, hyp, out, st,
                                       job.pinhole != 0);
}

void launch_selftest_div(int64_t n, uint64_t seed, unsigned long long *bad, hipStream_t st) {
    hipLaunchKernelGGL(k_selftest_div, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
                       seed, bad);
}

void launch_threshold(const RefJob &job, float *depth, float *normal, hipStream_t st) {
    const int64_t n = (int64_t)job.w * job.h;
    hipLaunchKernelGGL(k_threshold, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, job,
                       depth, normal);
}

void launch_consistency(const RefCam64 &ref, const float *depth, int32_t w, int32_t h,
                        const NbrCam *nbrs, int32_t nn, int32_t need, int32_t *agree,
                        unsigned long long *n_ok, hipStream_t st) {
    const int64_t n = (int64_t)w * h;
    hipLaunchKernelGGL(k_consistency, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ref,
                       depth, w, h, nbrs, nn, need, agree, n_ok);
}

}  // namespace amvs
