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
#include "mvs_device.h"
#include "mvs_geometry.h"

#include <math.h>

namespace amvs {

namespace {
using namespace dev;

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

// Corner-packed 8-bit plane for the fast kernels: pixel (x, y) holds the bytes
// of (x, y), (x+1, y), (x, y+1), (x+1, y+1) (clamped at the last column/row),
// so one aligned 32-bit load fetches a bilinear sample's four corners.
__global__ void k_quad_from_u8(const uint8_t *__restrict__ src, int32_t w, int32_t h,
                               int32_t stride, uint32_t *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)w * h) return;
    const int32_t y = (int32_t)(i / w), x = (int32_t)(i - (int64_t)y * w);
    const int32_t x1 = x + 1 < w ? x + 1 : x, y1 = y + 1 < h ? y + 1 : y;
    const uint8_t *r0 = src + (int64_t)y * stride, *r1 = src + (int64_t)y1 * stride;
    dst[i] = (uint32_t)r0[x] | ((uint32_t)r0[x1] << 8) | ((uint32_t)r1[x] << 16) |
             ((uint32_t)r1[x1] << 24);
}

// Bilinear-coefficient plane in f16 for the one-lane-per-pixel kernels: the
// sample at (x + ax, y + ay) of the cell whose top-left corner is (x, y) is
//     g00 + ax A + ay (B + ax C),
// A = g10 - g00, B = g01 - g00, C = g11 - g10 - g01 + g00 (g_ij the byte at
// (x + i, y + j), clamped at the last column/row). Pixel (x, y) holds
// {g00, A} and {B, C} as f16 (integers below 2^11, exact), so one 8-byte load
// gives a sample's three multiply-adds their operands; A + ay C and B + ax C
// are the sample's image gradient.
__global__ void k_quadh_from_u8(const uint8_t *__restrict__ src, int32_t w, int32_t h,
                                int32_t stride, uint2 *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)w * h) return;
    const int32_t y = (int32_t)(i / w), x = (int32_t)(i - (int64_t)y * w);
    const int32_t x1 = x + 1 < w ? x + 1 : x, y1 = y + 1 < h ? y + 1 : y;
    const uint8_t *r0 = src + (int64_t)y * stride, *r1 = src + (int64_t)y1 * stride;
    auto h16 = [](int v) -> uint32_t {
        return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)(float)v);
    };
    const int g00 = r0[x], g10 = r0[x1], g01 = r1[x], g11 = r1[x1];
    dst[i] = make_uint2(h16(g00) | (h16(g10 - g00) << 16),
                        h16(g01 - g00) | (h16(g11 - g10 - g01 + g00) << 16));
}

// Bilinear-coefficient plane of f32 grey views: pixel (x, y) holds {g00, A,
// B, C} as f32 with A = g10 - g00, B = g01 - g00, C = (g11 - g01) - A, each
// rounded once in that order (clamped at the last column/row); the kernels
// that read the grey plane itself form the same coefficients the same way.
__global__ void k_quadd(const float *__restrict__ g, int32_t w, int32_t h,
                        float4 *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)w * h) return;
    const int32_t y = (int32_t)(i / w), x = (int32_t)(i - (int64_t)y * w);
    const int32_t x1 = x + 1 < w ? x + 1 : x, y1 = y + 1 < h ? y + 1 : y;
    const float *r0 = g + (int64_t)y * w, *r1 = g + (int64_t)y1 * w;
    const float A = r0[x1] - r0[x];
    dst[i] = make_float4(r0[x], A, r1[x] - r0[x], (r1[x1] - r1[x]) - A);
}

// Corner-packed grey plane: pixel (x, y) holds the grey of (x, y), (x+1, y),
// (x, y+1), (x+1, y+1) (clamped at the last column/row), so one 16-byte load
// fetches a bilinear sample's four corners with their exact values.
__global__ void k_quadf(const float *__restrict__ g, int32_t w, int32_t h,
                        float4 *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)w * h) return;
    const int32_t y = (int32_t)(i / w), x = (int32_t)(i - (int64_t)y * w);
    const int32_t x1 = x + 1 < w ? x + 1 : x, y1 = y + 1 < h ? y + 1 : y;
    const float *r0 = g + (int64_t)y * w, *r1 = g + (int64_t)y1 * w;
    dst[i] = make_float4(r0[x], r0[x1], r1[x], r1[x1]);
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

// The threads of k_init walk the view in tiles of kListTileW x kListTileH
// pixels (tiles row-major, pixels row-major inside a tile), and the parity
// lists hold the textured pixels in exactly that order (per-wavefront counts,
// one scan, then a fill in walk order): the sweeps, which take list items in
// order, have a compact square of pixels in flight on each XCD instead of a
// band of full rows (while the hypotheses are still scattered, the source
// lines a square reaches along its epipolar segments are a fraction of a
// band's), and each 64-item chunk is the next 64 textured pixels of one
// parity in a 32-pixel-wide strip, about 4 rows of it: a compact block whose
// windows fall in or out of each source's coverage together and share cache
// lines. The order changes no result: every list reader takes (x, y) from the
// stored index. (32 x 256 measured best of tiles 16..256 wide, 64..1024 tall,
// on 2000x1500 views.)
#ifndef AMVS_LIST_TILE_W
#define AMVS_LIST_TILE_W 32
#endif
#ifndef AMVS_LIST_TILE_H
#define AMVS_LIST_TILE_H 256
#endif
constexpr int32_t kListTileW = AMVS_LIST_TILE_W;
constexpr int32_t kListTileH = AMVS_LIST_TILE_H;
static_assert((kListTileW * kListTileH) % 64 == 0 && kListTileH > 0, "list tiles hold whole wavefronts");

// Pixel (x, y) of position t of the tiled walk; false past the view's edge
// (x = y = 0 then).
__device__ __forceinline__ bool walk_pixel(int32_t w, int32_t h, int64_t t, int32_t &x,
                                           int32_t &y) {
    const int32_t ntx = (w + kListTileW - 1) / kListTileW;
    const int64_t tile = t / (kListTileW * kListTileH);
    const int32_t r = (int32_t)(t - tile * (kListTileW * kListTileH));
    const int32_t ty = (int32_t)(tile / ntx);
    const int32_t tx = (int32_t)(tile - (int64_t)ty * ntx);
    const int32_t yt = ty * kListTileH + r / kListTileW;
    const int32_t xt = tx * kListTileW + r % kListTileW;
    const bool valid = xt < w && yt < h;
    y = valid ? yt : 0;
    x = valid ? xt : 0;
    return valid;
}

// Initial hypotheses (dense.rs:365-381), texture mask (dense.rs:383-402) and
// each wavefront's count of textured pixels per parity (counts[4 + 2 wave +
// parity]); thread t takes pixel (x, y) of the tiled walk above.
__global__ __launch_bounds__(256) void k_init(RefJob job, int32_t taps, int32_t write_hyp) {
    const int32_t w = job.w, h = job.h;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t x, y;
    const bool valid = walk_pixel(w, h, t, x, y);
    const int64_t i = (int64_t)y * w + x;
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
        job.changed[i] = 0;
        job.pair[i] = 0xff;  // unknown until a pixel-lane kernel scores the plane
    }
    const int parity = (x + y) & 1;
    const uint64_t m0 = __ballot(valid && tex && parity == 0);
    const uint64_t m1 = __ballot(valid && tex && parity == 1);
    if ((threadIdx.x & 63) == 0) {
        const int64_t wave = t >> 6;
        job.counts[4 + 2 * wave] = __popcll(m0);
        job.counts[5 + 2 * wave] = __popcll(m1);
    }
}

// Exclusive scan of the per-wavefront counts in walk order (one workgroup),
// in place; the list lengths go to counts[0] (parity 0) and counts[1].
__global__ __launch_bounds__(1024) void k_list_scan(int32_t *counts, int64_t nwaves) {
    __shared__ int32_t part[2][1024];
    const int tid = (int)threadIdx.x;
    const int64_t per = (nwaves + 1023) / 1024;
    const int64_t b = tid * per, e = b + per < nwaves ? b + per : nwaves;
    int32_t s0 = 0, s1 = 0;
    for (int64_t i = b; i < e; ++i) {
        s0 += counts[4 + 2 * i];
        s1 += counts[5 + 2 * i];
    }
    part[0][tid] = s0;
    part[1][tid] = s1;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int32_t a0 = tid >= off ? part[0][tid - off] : 0;
        const int32_t a1 = tid >= off ? part[1][tid - off] : 0;
        __syncthreads();
        part[0][tid] += a0;
        part[1][tid] += a1;
        __syncthreads();
    }
    int32_t o0 = tid ? part[0][tid - 1] : 0, o1 = tid ? part[1][tid - 1] : 0;
    for (int64_t i = b; i < e; ++i) {
        const int32_t c0 = counts[4 + 2 * i], c1 = counts[5 + 2 * i];
        counts[4 + 2 * i] = o0;
        counts[5 + 2 * i] = o1;
        o0 += c0;
        o1 += c1;
    }
    if (tid == 1023) {
        counts[0] = part[0][1023];
        counts[1] = part[1][1023];
    }
}

// The parity lists: each wavefront writes its textured pixels of each parity
// at its scanned offset, in walk order.
__global__ __launch_bounds__(256) void k_list_fill(RefJob job) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t x, y;
    const bool valid = walk_pixel(job.w, job.h, t, x, y);
    const int64_t i = (int64_t)y * job.w + x;
    const bool tex = valid && job.textured[i] != 0;
    const int parity = (x + y) & 1;
    const uint64_t m0 = __ballot(tex && parity == 0);
    const uint64_t m1 = __ballot(tex && parity == 1);
    if (tex) {
        const int lane = (int)(threadIdx.x & 63);
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        const int64_t wave = t >> 6;
        const int32_t pos = job.counts[4 + 2 * wave + parity] + __popcll((parity ? m1 : m0) & below);
        job.list[(int64_t)parity * job.list_cap + pos] = (int32_t)i;
    }
}

// Device self-check of div2_shared against IEEE `/`: n hashed (a0, a1, b)
// triples, half drawn log-uniformly over [2^-40, 2^40] with random signs on
// the numerators, half from the projection's working range (|a| < 4000,
// b in [0.05, 2000]). Counts quotients that differ in any bit.
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
// Geometric consistency of one depth map, per pixel (agree_pixel in
// mvs_geometry.h): agree counts (-1 where depth <= 0) and the number of
// pixels with agree >= need.
__global__ void k_consistency(RefCam64 rf, const float *__restrict__ depth, int32_t w, int32_t h,
                              const NbrCam *__restrict__ nbrs, int32_t nn, int32_t need,
                              int32_t rule, int32_t *__restrict__ agree,
                              unsigned long long *n_ok) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool good = false;
    if (i < (int64_t)w * h) {
        double xw[3];
        const int32_t a = agree_pixel(rf, depth, w, nbrs, nn, rule, i, xw);
        agree[i] = a;
        good = a >= 0 && a >= need;
    }
    const uint64_t m = __ballot(good);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(n_ok, (unsigned long long)__popcll(m));
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------

}  // namespace

int lanes_for(int32_t nsrc) {
    if (nsrc <= 4) return 4;
    if (nsrc <= 8) return 8;
    if (nsrc <= 16) return 16;
    return -1;
}

bool taps_supported(int32_t taps) { return taps == 5 || taps == 7 || taps == 11; }

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

void launch_quad_from_u8(const uint8_t *src, int32_t w, int32_t h, int32_t stride,
                         uint32_t *dst, hipStream_t st) {
    const int64_t n = (int64_t)w * h;
    hipLaunchKernelGGL(k_quad_from_u8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src,
                       w, h, stride, dst);
}

void launch_quadh_from_u8(const uint8_t *src, int32_t w, int32_t h, int32_t stride, uint2 *dst,
                          hipStream_t st) {
    const int64_t n = (int64_t)w * h;
    hipLaunchKernelGGL(k_quadh_from_u8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src,
                       w, h, stride, dst);
}

void launch_quadd(const float *gray, int32_t w, int32_t h, float4 *dst, hipStream_t st) {
    const int64_t n = (int64_t)w * h;
    hipLaunchKernelGGL(k_quadd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, gray, w, h,
                       dst);
}

void launch_quadf(const float *gray, int32_t w, int32_t h, float4 *dst, hipStream_t st) {
    const int64_t n = (int64_t)w * h;
    hipLaunchKernelGGL(k_quadf, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, gray, w, h,
                       dst);
}

void launch_rays(double f, double cx, double cy, double k1, double k2, int32_t w, int32_t h,
                 float2 *rays, hipStream_t st) {
    const int64_t n = (int64_t)w * h;
    hipLaunchKernelGGL(k_rays, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, f, cx, cy, k1,
                       k2, w, h, rays);
}

__global__ void k_zero_i32(int32_t *p, int32_t n) {
    if ((int32_t)threadIdx.x < n) p[threadIdx.x] = 0;
}

int64_t list_waves(int32_t w, int32_t h) {
    const int64_t ntx = (w + kListTileW - 1) / kListTileW;
    const int64_t nty = (h + kListTileH - 1) / kListTileH;
    return ntx * nty * (kListTileW * kListTileH / 64);  // tiles cover the view
}

void launch_init(const RefJob &job, int32_t taps, int32_t write_hyp, hipStream_t st) {
    const int64_t nw = list_waves(job.w, job.h), n = 64 * nw;
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, job.counts, 4);
    hipLaunchKernelGGL(k_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, job, taps,
                       write_hyp);
    hipLaunchKernelGGL(k_list_scan, dim3(1), dim3(1024), 0, st, job.counts, nw);
    hipLaunchKernelGGL(k_list_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, job);
}

int launch_pm(int mode, int32_t taps, int32_t srcl, int32_t split, int32_t pinhole,
              bool persist, bool fast, const PmArgs &a, int32_t grid, hipStream_t st) {
    if (split < 0) {
        if (!fast || pinhole != 2 || split < -5) return -1;
        const int plane = split == -1 ? PL_F16
                          : split == -2 ? PL_U8
                          : split == -3 ? PL_F32
                          : split == -4 ? PL_F32G
                                        : PL_F16T;
        return launch_pml(mode, taps, plane, persist, a, grid, st);
    }
    if (a.eval_own) return -1;  // only the one-lane-per-pixel kernels score own planes
    if (fast) {
        if (taps == 5) return launch_pm_t5f(mode, srcl, split, pinhole != 0, persist, a, grid, st);
        if (taps == 7) return launch_pm_t7f(mode, srcl, split, pinhole != 0, persist, a, grid, st);
        if (taps == 11) return launch_pm_t11f(mode, srcl, split, pinhole != 0, persist, a, grid, st);
        return -1;
    }
    if (taps == 5) return launch_pm_t5(mode, srcl, split, pinhole != 0, persist, a, grid, st);
    if (taps == 7) return launch_pm_t7(mode, srcl, split, pinhole != 0, persist, a, grid, st);
    if (taps == 11) return launch_pm_t11(mode, srcl, split, pinhole != 0, persist, a, grid, st);
    return -1;
}

int splits_for(int32_t taps, int32_t *out) {
    if (taps == 5) {
        out[0] = 1;
        return 1;
    }
    if (taps == 7) {
        out[0] = 1;
        out[1] = 2;
        out[2] = 4;
        return 3;
    }
    if (taps == 11) {
        out[0] = 2;
        out[1] = 4;
        return 2;
    }
    return 0;
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

namespace {
__global__ void k_zero_u64(unsigned long long *p, int32_t n) {
    const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = 0ull;
}
}  // namespace

void launch_zero_u64(unsigned long long *p, int32_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, p, n);
}

void launch_consistency(const RefCam64 &ref, const float *depth, int32_t w, int32_t h,
                        const NbrCam *nbrs, int32_t nn, int32_t need, int32_t rule,
                        int32_t *agree, unsigned long long *n_ok, hipStream_t st) {
    const int64_t n = (int64_t)w * h;
    hipLaunchKernelGGL(k_consistency, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ref,
                       depth, w, h, nbrs, nn, need, rule, agree, n_ok);
}

}  // namespace amvs
