// View preparation on the GPU: the per-photo work of prepare_views
// (crates/stitch-core/src/recon/dense.rs:100-112) after decoding — shrink_to
// (crates/stitch-core/src/imgops.rs:145-171: integer box downsample, then an
// exact-size bilinear resize) and the grey of every pixel (dense.rs:108).
// Integer sums and f64/f32 arithmetic in the reference's order (no contraction
// in this build), so the output bytes and grey values equal the reference's.
#include "mvs_internal.h"

#include <math.h>

namespace amvs {
namespace {

// imgops.rs:89-118: each output pixel is the rounded mean of an f x f block
// ((sum + area/2) / area in u32). One thread per output pixel.
__global__ void k_box_down(const uint8_t *__restrict__ src, int32_t stride, int32_t f,
                           int32_t ow, int32_t oh, uint8_t *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)ow * oh) return;
    const int32_t oy = (int32_t)(i / ow), ox = (int32_t)(i - (int64_t)oy * ow);
    uint32_t a0 = 0, a1 = 0, a2 = 0;
    for (int32_t sy = oy * f; sy < oy * f + f; ++sy) {
        const uint8_t *row = src + (int64_t)sy * stride + (int64_t)ox * f * 3;
        for (int32_t k = 0; k < f; ++k) {
            a0 += row[3 * k];
            a1 += row[3 * k + 1];
            a2 += row[3 * k + 2];
        }
    }
    const uint32_t area = (uint32_t)(f * f);
    uint8_t *o = dst + 3 * i;
    o[0] = (uint8_t)((a0 + area / 2) / area);
    o[1] = (uint8_t)((a1 + area / 2) / area);
    o[2] = (uint8_t)((a2 + area / 2) / area);
}

// imgops.rs:120-143: f64 source coordinates of the output pixel centre,
// clamped to the image, bilinear weights, rounded half away from zero.
__global__ void k_resize_bilinear(const uint8_t *__restrict__ src, int32_t sw, int32_t sh,
                                  int32_t stride, uint8_t *__restrict__ dst, int32_t w,
                                  int32_t h) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)w * h) return;
    const int32_t y = (int32_t)(i / w), x = (int32_t)(i - (int64_t)y * w);
    const double sx = (double)sw / (double)w, sy = (double)sh / (double)h;
    double fy = ((double)y + 0.5) * sy - 0.5;
    fy = fy < 0.0 ? 0.0 : (fy > (double)(sh - 1) ? (double)(sh - 1) : fy);
    const int32_t y0 = (int32_t)fy;
    const int32_t y1 = y0 + 1 < sh - 1 ? y0 + 1 : sh - 1;
    const double ay = fy - (double)y0;
    double fx = ((double)x + 0.5) * sx - 0.5;
    fx = fx < 0.0 ? 0.0 : (fx > (double)(sw - 1) ? (double)(sw - 1) : fx);
    const int32_t x0 = (int32_t)fx;
    const int32_t x1 = x0 + 1 < sw - 1 ? x0 + 1 : sw - 1;
    const double ax = fx - (double)x0;
    const uint8_t *r0 = src + (int64_t)y0 * stride, *r1 = src + (int64_t)y1 * stride;
    uint8_t *o = dst + 3 * i;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double p00 = r0[3 * x0 + c], p10 = r0[3 * x1 + c];
        const double p01 = r1[3 * x0 + c], p11 = r1[3 * x1 + c];
        double top = p00 * (1.0 - ax);
        top = top + p10 * ax;
        double bot = p01 * (1.0 - ax);
        bot = bot + p11 * ax;
        double v = top * (1.0 - ay);
        v = v + bot * ay;
        o[c] = (uint8_t)round(v);
    }
}

// Packed copy of a strided RGB image (no shrinking needed).
__global__ void k_rgb_pack(const uint8_t *__restrict__ src, int32_t stride, int32_t w, int32_t h,
                           uint8_t *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)w * h * 3) return;
    const int64_t row = i / ((int64_t)w * 3), col = i - row * w * 3;
    dst[i] = src[row * stride + col];
}

// dense.rs:108: ((0.299 R + 0.587 G) + 0.114 B) / 255 in f32.
__global__ void k_view_grey(const uint8_t *__restrict__ rgb, int64_t n, float *__restrict__ grey) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = 0.299f * (float)rgb[3 * i] + 0.587f * (float)rgb[3 * i + 1];
    v = v + 0.114f * (float)rgb[3 * i + 2];
    grey[i] = v / 255.0f;
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void prep_sizes(int32_t w, int32_t h, int32_t max_side, PrepSizes &z) {
    // imgops.rs:145-171
    const int32_t lng = w > h ? w : h;
    z.f = 0;
    z.bw = z.ow = w;
    z.bh = z.oh = h;
    z.scale = 1.0;
    if (lng <= max_side) return;
    const int32_t fac = lng / max_side;
    if (fac >= 2) {
        z.f = fac;
        z.bw = w / fac;
        z.bh = h / fac;
    }
    z.scale = (double)max_side / (double)lng;
    const int32_t bl = z.bw > z.bh ? z.bw : z.bh;
    const double k = (double)max_side / (double)bl;
    z.ow = (int32_t)round((double)z.bw * k);
    z.oh = (int32_t)round((double)z.bh * k);
    if (z.ow < 1) z.ow = 1;
    if (z.oh < 1) z.oh = 1;
}

void launch_prepare_view(const uint8_t *src, int32_t w, int32_t h, int32_t stride,
                         const PrepSizes &z, uint8_t *boxed, uint8_t *out_rgb, float *grey,
                         hipStream_t st) {
    const uint8_t *b = src;
    int32_t bstride = stride;
    if (z.f >= 2) {
        hipLaunchKernelGGL(k_box_down, dim3(blocks_for((int64_t)z.bw * z.bh)), dim3(256), 0, st,
                           src, stride, z.f, z.bw, z.bh, boxed);
        b = boxed;
        bstride = z.bw * 3;
    }
    if (z.ow == z.bw && z.oh == z.bh) {
        if (z.f >= 2)
            (void)hipMemcpyAsync(out_rgb, boxed, (size_t)z.ow * z.oh * 3, hipMemcpyDeviceToDevice,
                                 st);
        else
            hipLaunchKernelGGL(k_rgb_pack, dim3(blocks_for((int64_t)w * h * 3)), dim3(256), 0, st,
                               src, stride, w, h, out_rgb);
    } else {
        hipLaunchKernelGGL(k_resize_bilinear, dim3(blocks_for((int64_t)z.ow * z.oh)), dim3(256), 0,
                           st, b, z.bw, z.bh, bstride, out_rgb, z.ow, z.oh);
    }
    const int64_t n = (int64_t)z.ow * z.oh;
    hipLaunchKernelGGL(k_view_grey, dim3(blocks_for(n)), dim3(256), 0, st, out_rgb, n, grey);
}

}  // namespace amvs
