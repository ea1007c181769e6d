// PatchMatch cost evaluation (crates/stitch-core/src/recon/dense.rs:181-263)
// and red-black half-sweeps (:413-467) for gfx950, batched over reference
// views. Included by one translation unit per window size.
//
// Launch shape: persistent 64-lane workgroups (one wavefront each). The work of
// one launch is every active pixel of every job (reference view) in the batch,
// cut into chunks of `gpc` pixel groups; k_pm_prep turns the per-job pixel
// counts, which only the device knows, into chunk offsets and zeroes the work
// counters, so a launch sequence needs no host round trip and can be captured
// in a graph. Each job's chunks are cut into 8 bands, one per XCD (workgroups
// are dispatched round-robin over the XCDs, blockIdx % 8); XCD x walks band x
// of job 0, then band x of job 1, and so on. All XCDs therefore work on the
// same view at the same time (its source images stay in the shared Infinity
// Cache) while each XCD's L2 serves its own band. A workgroup whose XCD has
// run out of work steals from the other XCDs' sequences.
//
// Lane roles inside a wavefront: LANES = SRCL x SPLIT consecutive lanes share
// one reference pixel; SRCL >= K lanes own one source each and SPLIT lanes
// split that source's window into contiguous tap slices. The tap-ordered sums
// of the two-pass NCC (dense.rs:266-271, :249-252) are continued from lane to
// lane with shuffles, so the summation order - and every rounding - is the
// reference's, while each lane holds only its slice of the samples.
#pragma once

#include "mvs_device.h"

namespace amvs {
namespace {

constexpr int kNumXcd = 8;

enum { MODE_EVAL_CURRENT = 0, MODE_SWEEP = 1, MODE_COST_LIST = 2 };

__device__ __forceinline__ int64_t pm_items(const RefJob &job, const PmArgs &a, int mode) {
    if (mode == MODE_EVAL_CURRENT) return (int64_t)job.counts[0] + job.counts[1];
    if (mode == MODE_SWEEP) return job.counts[a.parity];
    return a.n_items;
}

// Chunk offsets per job, per-XCD band offsets and zeroed work counters for
// the launch that follows. prefix[0..njobs]: first chunk of each job;
// prefix[(1 + x)(njobs + 1) + j]: position of job j in XCD x's sequence.
__global__ void k_pm_prep(PmArgs a, int32_t mode, int32_t chunk) {
    const int32_t nj = a.njobs;
    if (threadIdx.x < kNumXcd) a.counters[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        int32_t acc = 0;
        a.prefix[0] = 0;
        for (int32_t j = 0; j < nj; ++j) {
            const int64_t items = pm_items(a.jobs[j], a, mode);
            acc += (int32_t)((items + chunk - 1) / chunk);
            a.prefix[j + 1] = acc;
        }
    }
    __syncthreads();
    if (threadIdx.x < kNumXcd) {
        const int32_t x = threadIdx.x;
        int32_t *bp = a.prefix + (int64_t)(1 + x) * (nj + 1);
        int32_t acc = 0;
        bp[0] = 0;
        for (int32_t j = 0; j < nj; ++j) {
            const int32_t n = a.prefix[j + 1] - a.prefix[j];
            acc += n * (x + 1) / kNumXcd - n * x / kNumXcd;
            bp[j + 1] = acc;
        }
    }
}

// PERSIST: a grid of resident workgroups that fetch band positions with
// atomics and steal across XCDs; otherwise a static grid where workgroup b
// takes position b / 8 of XCD (b mod 8)'s sequence (the grid is sized for
// the longest possible sequences and surplus workgroups exit at once).
// FAST: the same algorithm with contracted multiply-adds, the hardware
// reciprocal for the projection and plane divisions, and each tap slice's
// NCC sums formed independently and added across the slice lanes; results
// stay within the float tolerance of the exact kernels (FAST = false), which
// round operation for operation like the reference.
template <int TAPS, int SRCL, int SPLIT, bool PINHOLE, int MODE, bool PERSIST, bool FAST>
__global__ __launch_bounds__(64) void k_pm(PmArgs a) {
    using namespace dev;
    constexpr int N = TAPS * TAPS;
    constexpr int LANES = SRCL * SPLIT;
    constexpr int PIX = 64 / LANES;
    constexpr int OWN = (N + LANES - 1) / LANES;  // plane-phase taps per lane
    constexpr int SEG = (N + SPLIT - 1) / SPLIT;  // sampled taps per lane
    constexpr uint64_t GMASK = LANES == 64 ? ~0ull : ((1ull << LANES) - 1ull);
    constexpr uint64_t SMASK = (1ull << SPLIT) - 1ull;
    // Per pixel slot: window points of the current hypothesis (xyz) with the
    // centred reference grey (w), and the window rays.
    __shared__ float4 Ps[PIX][N];
    __shared__ float2 Rs[PIX][N];

    const int lane = threadIdx.x;
    const int g = lane / LANES, r = lane % LANES;
    const int sidx = r / SPLIT, part = r % SPLIT;
    const int src0 = g * LANES + sidx * SPLIT;  // first lane of this lane's source
    const int chunk = PIX * a.gpc;
    const int32_t nj = a.njobs;
    const int home = blockIdx.x % kNumXcd;
    unsigned long long n_evals = 0;

    for (int attempt = 0; attempt < (PERSIST ? kNumXcd : 1); ++attempt) {
        const int xcd = (home + attempt) % kNumXcd;
        const int32_t *bp = a.prefix + (int64_t)(1 + xcd) * (nj + 1);
        const int32_t total = bp[nj];
        bool first = true;
        while (true) {
            int32_t v = 0;
            if (PERSIST) {
                if (lane == 0) v = atomicAdd(&a.counters[xcd], 1);
                v = __builtin_amdgcn_readfirstlane(__shfl(v, 0));
            } else {
                if (!first) break;
                first = false;
                v = (int32_t)(blockIdx.x / kNumXcd);
            }
            if (v >= total) break;
            // job of sequence position v: the last j with bp[j] <= v
            int32_t jl = 0, jh = nj - 1;
            while (jl < jh) {
                const int32_t mid = (jl + jh + 1) >> 1;
                if (bp[mid] <= v)
                    jl = mid;
                else
                    jh = mid - 1;
            }
            const int32_t jj = __builtin_amdgcn_readfirstlane(jl);
            const int32_t njc = a.prefix[jj + 1] - a.prefix[jj];
            const int32_t c = a.prefix[jj] + njc * xcd / kNumXcd + (v - bp[jj]);
            const RefJob &job = a.jobs[jj];
            const int64_t items = pm_items(job, a, MODE);
            const int64_t base = (int64_t)(c - a.prefix[jj]) * chunk;
            const int32_t w = job.w, h = job.h, K = job.nsrc, stride = job.stride,
                          half = job.half;

            // This lane's source view.
            const bool has_src = sidx < K;
            const SrcCam &S = job.srcs[has_src ? sidx : 0];
            const float R0 = S.R[0], R1 = S.R[1], R2 = S.R[2], R3 = S.R[3], R4 = S.R[4],
                        R5 = S.R[5], R6 = S.R[6], R7 = S.R[7], R8 = S.R[8];
            const float T0 = S.t[0], T1 = S.t[1], T2 = S.t[2];
            const float F = S.f, CX = S.cx, CY = S.cy;
            const ff FF{S.f, S.f_lo}, FCX{S.cx, S.cx_lo}, FCY{S.cy, S.cy_lo},
                FK1{S.k1, S.k1_lo}, FK2{S.k2, S.k2_lo};
            const float K1 = S.k1, K2 = S.k2;
            const uint32_t *__restrict__ Q = S.quad;  // fast kernels only
            const float4 *__restrict__ QF = S.quadf;
            const int32_t sw = S.w;
            const float wm1 = (float)(S.w - 1), hm1 = (float)(S.h - 1);

            for (int b = 0; b < a.gpc; ++b) {
                const int64_t item0 = base + (int64_t)b * PIX;
                if (item0 >= items) break;  // uniform
                const int64_t item = item0 + g;
                const bool have = item < items;
                int32_t x = 0, y = 0;
                if (have) {
                    int32_t pix;
                    if (MODE == MODE_EVAL_CURRENT) {
                        const int64_t c0 = job.counts[0];
                        pix = item < c0 ? job.list[item] : job.list[job.list_cap + (item - c0)];
                    } else if (MODE == MODE_SWEEP) {
                        pix = job.list[(int64_t)a.parity * job.list_cap + item];
                    } else {
                        pix = -1;
                        x = a.px[item];
                        y = a.py[item];
                    }
                    if (pix >= 0) {
                        y = pix / w;
                        x = pix - y * w;
                    }
                }
                const int64_t pidx = (int64_t)y * w + x;
                const bool inside =
                    have && x >= half && y >= half && x < w - half && y < h - half;

                // Reference patch statistics (dense.rs:213, mean_sd :266-271),
                // sums in tap order; every lane computes them (equal values).
                float mr = 0.0f, sr = 0.0f;
                if (inside) {
                    float s = 0.0f;
#pragma unroll
                    for (int j = 0; j < N; ++j) {
                        const int dy = (j / TAPS - TAPS / 2) * stride,
                                  dx = (j % TAPS - TAPS / 2) * stride;
                        s = s + job.gray[pidx + (int64_t)dy * w + dx];
                    }
                    mr = s / (float)N;
                    float v = 0.0f;
#pragma unroll
                    for (int j = 0; j < N; ++j) {
                        const int dy = (j / TAPS - TAPS / 2) * stride,
                                  dx = (j % TAPS - TAPS / 2) * stride;
                        const float d = job.gray[pidx + (int64_t)dy * w + dx] - mr;
                        v = v + d * d;
                    }
                    v = v / (float)N;
                    sr = sqrtf(v);
                }
                const bool pixel_ok = inside && !(sr < job.texture_min);

                // Stage this lane's share of the window rays and centred grey.
#pragma unroll
                for (int o = 0; o < OWN; ++o) {
                    const int j = r + LANES * o;
                    if (j < N && pixel_ok) {
                        const int dy = (j / TAPS - TAPS / 2) * stride,
                                  dx = (j % TAPS - TAPS / 2) * stride;
                        const int64_t qi = pidx + (int64_t)dy * w + dx;
                        Rs[g][j] = job.rays[qi];
                        Ps[g][j].w = job.gray[qi] - mr;
                    }
                }
                const float2 rc = pixel_ok ? job.rays[pidx] : make_float2(0.0f, 0.0f);
                __syncthreads();
                // fast sums: this slice's sum of the centred reference grey
                float asum = 0.0f;
                if constexpr (FAST) {
                    if (pixel_ok) {
#pragma unroll
                        for (int i = 0; i < SEG; ++i)
                            if (SPLIT == 1 || part * SEG + i < N) asum += Ps[g][part * SEG + i].w;
                    }
                }

                // Cost of hypothesis (d, n) at this group's pixel; `act` is
                // uniform over the group. Every lane of the wave calls this.
                auto eval = [&](bool act, float d, float n0, float n1, float n2) -> float {
                    act = act && pixel_ok;
                    bool pf = false;
                    if (act) {
                        // dense.rs:187-211: plane through X = ray*d with normal
                        // n; depth of each window ray on that plane.
                        const float xp0 = rc.x * d, xp1 = rc.y * d;
                        float cc = n0 * xp0 + n1 * xp1;
                        cc = cc + n2 * d;
#pragma unroll
                        for (int o = 0; o < OWN; ++o) {
                            const int j = r + LANES * o;
                            if (j < N) {
                                const float2 ray = Rs[g][j];
                                float den = n0 * ray.x + n1 * ray.y;
                                den = den + n2;
                                // IEEE quotient for |den| >= 1e-6; smaller fails
                                float dq;
                                if constexpr (FAST)
                                    dq = cc * __builtin_amdgcn_rcpf(den);
                                else
                                    dq = div1(cc, den, recip_refined(den));
                                pf = pf || (fabsf(den) < 1e-6f) || !(dq > 0.0f);
                                float *pp = &Ps[g][j].x;
                                pp[0] = ray.x * dq;
                                pp[1] = ray.y * dq;
                                pp[2] = dq;
                            }
                        }
                    }
                    const uint64_t fb = __ballot(pf);
                    const bool gfail = ((fb >> (g * LANES)) & GMASK) != 0ull;
                    __syncthreads();
                    float cs = has_src ? 2.0f : INFINITY;
                    if (act && !gfail && has_src) {
                        // dense.rs:218-253 for source sidx, taps
                        // [part*SEG, part*SEG + SEG). The reference stops at a
                        // source's first failing sample and scores it 2.0; here
                        // every sample is computed without branches (a failing
                        // one reads a safe address) and `ok` collects failures,
                        // which gives the same cost while the gathers of
                        // successive taps overlap.
                        const float4 *P = Ps[g];
                        // exact: the samples, for the reference's two-pass sums;
                        // fast: single-pass sums shifted by the slice's first
                        // sample (c0), merged across slices below
                        float sv[FAST ? 1 : SEG];
                        float c0 = 0.0f, s1 = 0.0f, s2 = 0.0f, sa = 0.0f;
                        bool ok = true;
#pragma unroll
                        for (int i = 0; i < SEG; ++i) {
                            const int j = part * SEG + i;
                            const bool inw = SPLIT == 1 || j < N;
                            const float4 p = P[inw ? j : 0];
                            float xs2, u, v;
                            if constexpr (FAST) {
                                xs2 = fmaf(R6, p.x, fmaf(R7, p.y, fmaf(R8, p.z, T2)));
                                const float xs0 = fmaf(R0, p.x, fmaf(R1, p.y, fmaf(R2, p.z, T0)));
                                const float xs1 = fmaf(R3, p.x, fmaf(R4, p.y, fmaf(R5, p.z, T1)));
                                const float inv = __builtin_amdgcn_rcpf(xs2);
                                if (PINHOLE) {
                                    const float fi = F * inv;
                                    u = fmaf(fi, xs0, CX);
                                    v = fmaf(fi, xs1, CY);
                                } else {
                                    const float q0 = xs0 * inv, q1 = xs1 * inv;
                                    const float r2 = fmaf(q0, q0, q1 * q1);
                                    const float fd = F * fmaf(fmaf(K2, r2, K1), r2, 1.0f);
                                    u = fmaf(fd, q0, CX);
                                    v = fmaf(fd, q1, CY);
                                }
                            } else {
                                float xs0 = R0 * p.x + R1 * p.y;
                                xs0 = xs0 + R2 * p.z;
                                xs0 = xs0 + T0;
                                float xs1 = R3 * p.x + R4 * p.y;
                                xs1 = xs1 + R5 * p.z;
                                xs1 = xs1 + T1;
                                xs2 = R6 * p.x + R7 * p.y;
                                xs2 = xs2 + R8 * p.z;
                                xs2 = xs2 + T2;
                                float q0, q1;
                                div2_shared(xs0, xs1, xs2, q0, q1);
                                if (PINHOLE)
                                    project_pinhole(q0, q1, F, CX, CY, u, v);
                                else
                                    project_general(q0, q1, FF, FCX, FCY, FK1, FK2, u, v);
                            }
                            // gray_at, dense.rs:69-85: None unless 0 <= fx < w-1
                            // and 0 <= fy < h-1 (NaN fails too)
                            const float fx = u - 0.5f, fy = v - 0.5f;
                            const bool good = (xs2 > 0.0f) && fx >= 0.0f && fy >= 0.0f &&
                                              fx < wm1 && fy < hm1;
                            ok = ok && (good || !inw);
                            const float sx = good ? fx : 0.0f, sy = good ? fy : 0.0f;
                            const int x0 = (int)sx, y0 = (int)sy;
                            // fx >= 0: fract is fx - floor(fx) = fx - x0, exactly
                            const float ax = __builtin_amdgcn_fractf(sx),
                                        ay = __builtin_amdgcn_fractf(sy);
                            float val;
                            if constexpr (FAST) {
                                // the four corners as bytes in one 32-bit load;
                                // NCC is scale-invariant, so raw 0..255 values
                                // serve (the grey table is b / 255 to 1 ulp)
                                const uint32_t qd = as_global(Q)[y0 * sw + x0];
                                const float g00 = (float)(qd & 0xffu);
                                const float g10 = (float)((qd >> 8) & 0xffu);
                                const float g01 = (float)((qd >> 16) & 0xffu);
                                const float g11 = (float)(qd >> 24);
                                const float top = fmaf(ax, g10 - g00, g00);
                                const float bot = fmaf(ax, g11 - g01, g01);
                                val = fmaf(ay, bot - top, top);
                            } else {
                                // the four corners' grey in one 16-byte load
                                const f32x4 qv = as_global(QF)[y0 * sw + x0];
                                const float g00 = qv.x, g10 = qv.y, g01 = qv.z, g11 = qv.w;
                                const float bx = 1.0f - ax, by = 1.0f - ay;
                                float top = g00 * bx;
                                top = top + g10 * ax;
                                float bot = g01 * bx;
                                bot = bot + g11 * ax;
                                val = top * by;
                                val = val + bot * ay;
                            }
                            if constexpr (FAST) {
                                if (i == 0) c0 = val;
                                if (SPLIT == 1 || j < N) {
                                    const float dd = val - c0;
                                    s1 += dd;
                                    s2 = fmaf(dd, dd, s2);
                                    sa = fmaf(p.w, dd, sa);
                                }
                            } else {
                                sv[i] = val;
                            }
                        }
                        if (SPLIT > 1) {
                            // every slice of the source must have succeeded
                            const uint64_t bad = __ballot(!ok);
                            ok = ((bad >> src0) & SMASK) == 0ull;
                        }
                        float ms, var = 0.0f, cov = 0.0f;
                        if constexpr (FAST) {
                            // merge the slices' shifted sums (pairwise
                            // mean/variance update): slice count np, mean mp,
                            // M2 = s2 - s1^2/np; cov uses sum(a) of the slice
                            const int first = part * SEG;
                            const float np =
                                (float)(SPLIT == 1 ? N : (N - first < SEG ? N - first : SEG));
                            float tot = fmaf(np, c0, s1);
#pragma unroll
                            for (int off = 1; off < SPLIT; off <<= 1) tot += __shfl_xor(tot, off);
                            ms = tot * (1.0f / (float)N);
                            const float mp = c0 + s1 / np;
                            const float dm = mp - ms;
                            var = fmaxf(s2 - s1 * (s1 / np), 0.0f) + np * dm * dm;
                            cov = sa - (ms - c0) * asum;
#pragma unroll
                            for (int off = 1; off < SPLIT; off <<= 1) {
                                var += __shfl_xor(var, off);
                                cov += __shfl_xor(cov, off);
                            }
                            var = var * (1.0f / (float)N);
                            const float ss = __builtin_amdgcn_sqrtf(var);
                            // dense.rs:246-248's 1e-3 grey units, in byte units
                            if (ok && !(ss < 1e-3f * 255.0f))
                                cs = 1.0f - cov * __builtin_amdgcn_rcpf((float)N * sr * ss);
                        } else {
                            // mean: sum in tap order, continued slice to slice
                            float sum = 0.0f;
#pragma unroll
                            for (int p = 0; p < SPLIT; ++p) {
                                if (part == p) {
#pragma unroll
                                    for (int i = 0; i < SEG; ++i)
                                        if (SPLIT == 1 || p * SEG + i < N) sum = sum + sv[i];
                                }
                                if (SPLIT > 1) sum = __shfl(sum, src0 + p);
                            }
                            ms = sum / (float)N;
                            // variance and covariance, same order
#pragma unroll
                            for (int p = 0; p < SPLIT; ++p) {
                                if (part == p) {
#pragma unroll
                                    for (int i = 0; i < SEG; ++i) {
                                        if (SPLIT == 1 || p * SEG + i < N) {
                                            const float dd = sv[i] - ms;
                                            var = var + dd * dd;
                                            cov = cov + P[p * SEG + i].w * dd;
                                        }
                                    }
                                }
                                if (SPLIT > 1) {
                                    var = __shfl(var, src0 + p);
                                    cov = __shfl(cov, src0 + p);
                                }
                            }
                            var = var / (float)N;
                            const float ss = sqrtf(var);
                            if (ok && !(ss < 1e-3f)) {
                                float den = (float)N * sr;
                                den = den * ss;
                                cs = 1.0f - cov / den;
                            }
                        }
                    }
                    // Best two of the K source costs (dense.rs:255-262): exact
                    // min/min all-reduce over the group's sources.
                    float m1 = cs, m2 = INFINITY;
#pragma unroll
                    for (int off = SPLIT; off < LANES; off <<= 1) {
                        const float o1 = __shfl_xor(m1, off), o2 = __shfl_xor(m2, off);
                        const float lo2 = fminf(m1, o1);
                        m2 = fminf(fmaxf(m1, o1), fminf(m2, o2));
                        m1 = lo2;
                    }
                    float cst = K >= 2 ? (m1 + m2) * 0.5f : m1;
                    if (gfail || !pixel_ok) cst = 2.0f;
                    if (act && r == 0) ++n_evals;
                    __syncthreads();
                    return cst;
                };

                if (MODE == MODE_COST_LIST) {
                    const float4 hp = have ? a.hyp_in[item] : make_float4(1.0f, 0.0f, 0.0f, -1.0f);
                    const float cst = eval(have, hp.x, hp.y, hp.z, hp.w);
                    if (have && r == 0) a.out[item] = cst;
                } else if (MODE == MODE_EVAL_CURRENT) {
                    const float4 hp = have ? job.hyp[pidx] : make_float4(1.0f, 0.0f, 0.0f, -1.0f);
                    const float cst = eval(have, hp.x, hp.y, hp.z, hp.w);
                    if (have && r == 0) {
                        job.cost[pidx] = cst;
                        job.pair[pidx] = 0xff;  // (this kernel stores no pairs)
                    }
                } else {
                    const float4 orig = have ? job.hyp[pidx] : make_float4(1.0f, 0.0f, 0.0f, -1.0f);
                    float4 best = orig;
                    float bcost = have ? job.cost[pidx] : 2.0f;
                    // Ten proposals in the reference's order, one evaluation
                    // site. k < 8: propagation from the other colour
                    // (dense.rs:427-444) in CAND order (-1,0),(1,0),(0,-1),
                    // (0,1),(-5,0),(5,0),(0,-5),(0,5); those pixels are not
                    // written during this half-sweep, so reading live state
                    // equals reading the reference's snapshot (:416). k = 8, 9:
                    // random refinement narrowing each iteration (:445-462).
                    uint64_t s = seed_refine((uint64_t)pidx, (uint64_t)a.it, (uint64_t)a.parity);
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
                            const bool dup = hp.x == best.x && hp.y == best.y &&
                                             hp.z == best.z && hp.w == best.w;
                            act = inb && !dup;
                        } else {
                            float step = (rnd_next(s) - 0.5f) * a.range;
                            step = step * 0.25f;
                            step = step * span;
                            const float id = 1.0f / best.x + step;
                            hp.x = (id > job.imin * 0.5f && id < job.imax * 1.5f) ? 1.0f / id
                                                                                   : best.x;
                            random_normal(s, rc.x, rc.y, best.y, best.z, best.w, 0.25f * a.range,
                                          hp.y, hp.z, hp.w);
                            act = have;
                        }
                        const float cst = eval(act, hp.x, hp.y, hp.z, hp.w);
                        if (act && cst < bcost) {
                            best = hp;
                            bcost = cst;
                        }
                    }
                    if (have && r == 0) {
                        job.hyp[pidx] = best;
                        job.cost[pidx] = bcost;
                        // (float ==: a NaN component counts as a change)
                        const bool ch = !(best.x == orig.x && best.y == orig.y &&
                                          best.z == orig.z && best.w == orig.w);
                        job.changed[pidx] = ch;
                        if (ch) job.pair[pidx] = 0xff;  // (this kernel stores no pairs)
                    }
                }
            }
        }
    }

    if (a.evals) {
        unsigned long long e = n_evals;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) e += __shfl_xor(e, off);
        if (lane == 0 && e) atomicAdd(a.evals, e);
    }
}

// Prep + launch of one kernel class.
template <int TAPS, int SRCL, int SPLIT, bool PINHOLE, int MODE, bool PERSIST, bool FAST>
void launch_pm_class(const PmArgs &a, int32_t grid, hipStream_t st) {
    constexpr int PIX = 64 / (SRCL * SPLIT);
    hipLaunchKernelGGL(k_pm_prep, dim3(1), dim3(64), 0, st, a, (int32_t)MODE, PIX * a.gpc);
    hipLaunchKernelGGL((k_pm<TAPS, SRCL, SPLIT, PINHOLE, MODE, PERSIST, FAST>),
                       dim3((unsigned)grid), dim3(64), 0, st, a);
}

// Dispatch over the classes instantiated for one window size and arithmetic.
template <int TAPS, bool FAST, int... SPLITS>
struct PmDispatch {
    template <int SRCL, int SPLIT, bool PH>
    static void go(int mode, bool persist, const PmArgs &a, int32_t grid, hipStream_t st) {
        if (mode == MODE_EVAL_CURRENT)
            launch_pm_class<TAPS, SRCL, SPLIT, PH, MODE_EVAL_CURRENT, false, FAST>(a, grid, st);
        else if (mode == MODE_SWEEP && persist)
            launch_pm_class<TAPS, SRCL, SPLIT, PH, MODE_SWEEP, true, FAST>(a, grid, st);
        else if (mode == MODE_SWEEP)
            launch_pm_class<TAPS, SRCL, SPLIT, PH, MODE_SWEEP, false, FAST>(a, grid, st);
        else
            launch_pm_class<TAPS, SRCL, SPLIT, PH, MODE_COST_LIST, false, FAST>(a, grid, st);
    }
    template <int SPLIT>
    static bool try_split(int mode, int srcl, int split, bool ph, bool persist, const PmArgs &a,
                          int32_t grid, hipStream_t st) {
        if (split != SPLIT) return false;
        if (srcl == 4)
            ph ? go<4, SPLIT, true>(mode, persist, a, grid, st)
               : go<4, SPLIT, false>(mode, persist, a, grid, st);
        else if (srcl == 8)
            ph ? go<8, SPLIT, true>(mode, persist, a, grid, st)
               : go<8, SPLIT, false>(mode, persist, a, grid, st);
        else if (srcl == 16)
            ph ? go<16, SPLIT, true>(mode, persist, a, grid, st)
               : go<16, SPLIT, false>(mode, persist, a, grid, st);
        else
            return false;
        return true;
    }
    static int run(int mode, int srcl, int split, bool ph, bool persist, const PmArgs &a,
                   int32_t grid, hipStream_t st) {
        const bool done = (try_split<SPLITS>(mode, srcl, split, ph, persist, a, grid, st) || ...);
        return done ? 0 : -1;
    }
};

}  // namespace
}  // namespace amvs
