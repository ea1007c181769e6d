// Fusion: depth maps that neighbouring views confirm, accumulated per voxel
// (crates/stitch-core/src/recon/dense.rs:589-632) into a device hash table.
//
// The reference sums f32 coordinates and colours per voxel in a HashMap,
// per view and then under a mutex in completion order, so its sums depend on
// thread timing. Here every confirmed pixel adds its world point in 64-bit
// fixed point (2^-20 m) and its colour as integers with atomics: the sums, and
// so the fused cloud, are the same whatever order the points arrive in.
#include "mvs_device.h"
#include "mvs_geometry.h"

namespace amvs {
namespace {

constexpr unsigned long long kEmpty = ~0ull;
constexpr int64_t kKeyHalf = 1ll << 20;
constexpr double kFix = 1048576.0;  // fixed-point units per metre

__global__ void k_fuse_clear(FuseTable t) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < t.cap; i += stride) {
        t.keys[i] = kEmpty;
        t.sum[3 * i] = 0;
        t.sum[3 * i + 1] = 0;
        t.sum[3 * i + 2] = 0;
        t.col[3 * i] = 0;
        t.col[3 * i + 1] = 0;
        t.col[3 * i + 2] = 0;
        t.cnt[i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x < 4) t.counters[threadIdx.x] = 0;
}

__global__ void k_fuse(RefCam64 rf, const float *__restrict__ depth, int32_t w, int32_t h,
                       const NbrCam *__restrict__ nbrs, int32_t nn, int32_t rule, int32_t need,
                       double inv_voxel, const uint8_t *__restrict__ rgb,
                       const uint32_t *__restrict__ quad, const float *__restrict__ gray,
                       FuseTable t) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)w * h) return;
    double xw[3];
    const int32_t agree = agree_pixel(rf, depth, w, nbrs, nn, rule, i, xw);
    if (agree < 0 || agree < need) return;  // dense.rs:614-616
    // dense.rs:617-621: floor(xw / voxel) per axis
    int64_t k3[3];
    bool in_range = true;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const double f = floor(xw[r] * inv_voxel);
        in_range = in_range && f >= (double)-kKeyHalf && f < (double)kKeyHalf;
        k3[r] = in_range ? (int64_t)f : 0;
    }
    if (!in_range) {
        atomicAdd(&t.counters[3], 1ull);
        return;
    }
    const unsigned long long key = ((unsigned long long)(k3[0] + kKeyHalf) << 42) |
                                   ((unsigned long long)(k3[1] + kKeyHalf) << 21) |
                                   (unsigned long long)(k3[2] + kKeyHalf);
    const unsigned long long mask = (unsigned long long)t.cap - 1ull;
    unsigned long long slot = dev::mix64(key) & mask;
    bool found = false;
    for (int64_t probe = 0; probe < t.cap; ++probe) {
        const unsigned long long prev = atomicCAS(&t.keys[slot], kEmpty, key);
        if (prev == kEmpty) atomicAdd(&t.counters[0], 1ull);
        if (prev == kEmpty || prev == key) {
            found = true;
            break;
        }
        slot = (slot + 1ull) & mask;
    }
    if (!found) {
        atomicAdd(&t.counters[2], 1ull);
        return;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
        atomicAdd(&t.sum[3 * slot + r], (unsigned long long)(long long)llrint(xw[r] * kFix));
    // dense.rs:622-628: the pixel's colour
    unsigned int c0, c1, c2;
    if (rgb) {
        c0 = rgb[3 * i];
        c1 = rgb[3 * i + 1];
        c2 = rgb[3 * i + 2];
    } else {
        const unsigned int g =
            quad ? (quad[i] & 0xffu)
                 : (unsigned int)fminf(fmaxf(gray[i] * 255.0f + 0.5f, 0.0f), 255.0f);
        c0 = c1 = c2 = g;
    }
    atomicAdd(&t.col[3 * slot], c0);
    atomicAdd(&t.col[3 * slot + 1], c1);
    atomicAdd(&t.col[3 * slot + 2], c2);
    atomicAdd(&t.cnt[slot], 1u);
    atomicAdd(&t.counters[1], 1ull);
}

}  // namespace

void launch_fuse_clear(const FuseTable &t, hipStream_t st) {
    const int64_t blocks = std::min<int64_t>((t.cap + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fuse_clear, dim3((unsigned)blocks), dim3(256), 0, st, t);
}

void launch_fuse(const RefCam64 &ref, const float *depth, int32_t w, int32_t h,
                 const NbrCam *nbrs, int32_t nn, int32_t rule, int32_t need, double inv_voxel,
                 const uint8_t *rgb, const uint32_t *quad, const float *gray, const FuseTable &t,
                 hipStream_t st) {
    const int64_t n = (int64_t)w * h;
    hipLaunchKernelGGL(k_fuse, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ref, depth, w,
                       h, nbrs, nn, rule, need, inv_voxel, rgb, quad, gray, t);
}

}  // namespace amvs
