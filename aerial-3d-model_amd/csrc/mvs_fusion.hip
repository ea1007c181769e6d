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

#include <rocprim/device/device_radix_sort.hpp>

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

// Longest probe sequence of a global insert. Below the table's occupancy
// limit (70%) linear probing under mix64 ends within a few tens of slots; an
// insert that finds no slot within this many is counted as dropped, which
// makes the host rebuild the table larger, so a full table costs each insert
// at most this many probes instead of a walk over the whole capacity.
constexpr int64_t kMaxProbe = 1024;

// Global insert of one voxel's aggregated sums (open addressing, bounded
// linear probing). Returns false when no slot was found.
__device__ bool fuse_insert(const FuseTable &t, unsigned long long key,
                            const unsigned long long s3[3], const unsigned int c3[3],
                            unsigned int n) {
    const unsigned long long mask = (unsigned long long)t.cap - 1ull;
    unsigned long long slot = dev::mix64(key) & mask;
    const int64_t probes = t.cap < kMaxProbe ? t.cap : kMaxProbe;
    for (int64_t probe = 0; probe < probes; ++probe) {
        const unsigned long long prev = atomicCAS(&t.keys[slot], kEmpty, key);
        if (prev == kEmpty) atomicAdd(&t.counters[0], 1ull);
        if (prev == kEmpty || prev == key) {
#pragma unroll
            for (int r = 0; r < 3; ++r) atomicAdd(&t.sum[3 * slot + r], s3[r]);
#pragma unroll
            for (int r = 0; r < 3; ++r) atomicAdd(&t.col[3 * slot + r], c3[r]);
            atomicAdd(&t.cnt[slot], n);
            return true;
        }
        slot = (slot + 1ull) & mask;
    }
    return false;
}

// One 16 x 16 pixel tile per workgroup. Confirmed pixels first meet in an LDS
// table of the tile's voxels (at most 256 distinct keys, so the 256-slot table
// never fills); each distinct voxel then goes to the global table once. The
// sums are integers, so the result does not depend on this grouping.
constexpr int kTile = 16;
constexpr int kTileSlots = kTile * kTile;

__global__ __launch_bounds__(kTileSlots) void k_fuse(
    RefCam64 rf, const float *__restrict__ depth, int32_t w, int32_t h,
    const NbrCam *__restrict__ nbrs, int32_t nn, int32_t rule, int32_t need, double inv_voxel,
    const uint8_t *__restrict__ rgb, const uint32_t *__restrict__ quad,
    const float *__restrict__ gray, FuseTable t) {
    __shared__ unsigned long long lkey[kTileSlots];
    __shared__ unsigned long long lsum[kTileSlots][3];
    __shared__ unsigned int lcol[kTileSlots][3];
    __shared__ unsigned int lcnt[kTileSlots];
    __shared__ unsigned int lrange;
    __shared__ unsigned int lfull;
    const int tid = threadIdx.y * kTile + threadIdx.x;
    lkey[tid] = kEmpty;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        lsum[tid][r] = 0;
        lcol[tid][r] = 0;
    }
    lcnt[tid] = 0;
    if (tid == 0) {
        lrange = 0;
        // Once the table holds `limit` voxels, or an insert has been dropped,
        // this pass will be repeated with a larger table: the workgroup still
        // aggregates its points (their count sizes the next table) but makes
        // no global inserts. Read once per workgroup, so a few workgroups in
        // flight may still insert past the limit; the bounded probe covers
        // them.
        const unsigned long long vox =
            __hip_atomic_load(&t.counters[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long drop =
            __hip_atomic_load(&t.counters[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        lfull = (vox >= (unsigned long long)t.limit || drop != 0ull) ? 1u : 0u;
    }
    __syncthreads();
    const int32_t x = blockIdx.x * kTile + threadIdx.x, y = blockIdx.y * kTile + threadIdx.y;
    if (x < w && y < h) {
        const int64_t i = (int64_t)y * w + x;
        double xw[3];
        const int32_t agree = agree_pixel(rf, depth, w, nbrs, nn, rule, i, xw);
        if (agree >= 0 && agree >= need) {  // dense.rs:614-616
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
                atomicAdd(&lrange, 1u);
            } else {
                const unsigned long long key = ((unsigned long long)(k3[0] + kKeyHalf) << 42) |
                                               ((unsigned long long)(k3[1] + kKeyHalf) << 21) |
                                               (unsigned long long)(k3[2] + kKeyHalf);
                // dense.rs:622-628: the pixel's colour
                unsigned int c[3];
                if (rgb) {
                    c[0] = rgb[3 * i];
                    c[1] = rgb[3 * i + 1];
                    c[2] = rgb[3 * i + 2];
                } else {
                    const unsigned int g =
                        quad ? (quad[i] & 0xffu)
                             : (unsigned int)fminf(fmaxf(gray[i] * 255.0f + 0.5f, 0.0f), 255.0f);
                    c[0] = c[1] = c[2] = g;
                }
                // LDS slot of the key: at most 256 keys in 256 slots, so the
                // probe ends within 256 steps
                unsigned int slot = (unsigned int)dev::mix64(key) & (kTileSlots - 1);
                for (int probe = 0; probe < kTileSlots; ++probe) {
                    const unsigned long long prev = atomicCAS(&lkey[slot], kEmpty, key);
                    if (prev == kEmpty || prev == key) break;
                    slot = (slot + 1) & (kTileSlots - 1);
                }
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    atomicAdd(&lsum[slot][r], (unsigned long long)(long long)llrint(xw[r] * kFix));
                    atomicAdd(&lcol[slot][r], c[r]);
                }
                atomicAdd(&lcnt[slot], 1u);
            }
        }
    }
    __syncthreads();
    if (tid == 0 && lrange) atomicAdd(&t.counters[3], (unsigned long long)lrange);
    const unsigned int n = lcnt[tid];
    bool stored = false;
    if (n) {
        const unsigned long long s3[3] = {lsum[tid][0], lsum[tid][1], lsum[tid][2]};
        const unsigned int c3[3] = {lcol[tid][0], lcol[tid][1], lcol[tid][2]};
        stored = lfull ? false : fuse_insert(t, lkey[tid], s3, c3, n);
        if (!stored) atomicAdd(&t.counters[2], (unsigned long long)n);  // points dropped
    }
    // points accumulated: one atomic per wavefront
    unsigned long long pts = stored ? n : 0;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) pts += __shfl_xor(pts, off);
    if ((tid & 63) == 0 && pts) atomicAdd(&t.counters[1], pts);
}

// The occupied slots as (key, slot) pairs; one counter atomic per wavefront.
// Writes stop at n (the voxel count); the count returned lets the caller check
// that it saw exactly n.
__global__ void k_fuse_compact(FuseTable t, int64_t n, unsigned long long *__restrict__ okeys,
                               uint32_t *__restrict__ oslots, unsigned long long *ctr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long key = i < t.cap ? t.keys[i] : kEmpty;
    const bool occ = key != kEmpty;
    const unsigned long long m = __ballot(occ);
    if (m == 0ull) return;
    const int lane = (int)__lane_id();
    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    unsigned long long base = 0;
    if (lane == __ffsll((long long)m) - 1) base = atomicAdd(ctr, (unsigned long long)__popcll(m));
    base = __shfl(base, __ffsll((long long)m) - 1);
    if (occ && (int64_t)(base + below) < n) {
        okeys[base + below] = key;
        oslots[base + below] = (uint32_t)i;
    }
}

// dense.rs:634-641: one point per voxel, the mean position (f64 from the
// fixed-point sums, then f32) and the mean colour truncated to bytes.
__global__ void k_fuse_points(FuseTable t, const uint32_t *__restrict__ slots, int64_t n,
                              float *__restrict__ xyz, uint8_t *__restrict__ rgb) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    const int64_t i = slots[o];
    const unsigned int k = t.cnt[i];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        xyz[3 * o + r] = (float)((double)(long long)t.sum[3 * i + r] / kFix / (double)k);
        rgb[3 * o + r] = (uint8_t)((float)t.col[3 * i + r] / (float)k);
    }
}

}  // namespace

int fuse_points(const FuseTable &t, int64_t n, FusePointsWork &w, float *xyz, uint8_t *rgb,
                hipStream_t st) {
    if (n <= 0) return 0;
    if (hipMemsetAsync(w.ctr, 0, 8, st) != hipSuccess) return -1;
    const int64_t blocks = (t.cap + 255) / 256;
    hipLaunchKernelGGL(k_fuse_compact, dim3((unsigned)blocks), dim3(256), 0, st, t, n,
                       w.keys_in, w.slots_in, w.ctr);
    size_t bytes = w.tmp_bytes;
    if (rocprim::radix_sort_pairs(w.tmp, bytes, w.keys_in, w.keys_out, w.slots_in, w.slots_out,
                                  (size_t)n, 0, 64, st) != hipSuccess)
        return -1;
    hipLaunchKernelGGL(k_fuse_points, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, t,
                       w.slots_out, n, xyz, rgb);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

size_t fuse_points_tmp_bytes(int64_t n) {
    size_t bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, bytes, (unsigned long long *)nullptr,
                                  (unsigned long long *)nullptr, (uint32_t *)nullptr,
                                  (uint32_t *)nullptr, (size_t)n, 0, 64) != hipSuccess)
        return 0;
    return bytes;
}

void launch_fuse_clear(const FuseTable &t, hipStream_t st) {
    const int64_t blocks = std::min<int64_t>((t.cap + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fuse_clear, dim3((unsigned)blocks), dim3(256), 0, st, t);
}

void launch_fuse(const RefCam64 &ref, const float *depth, int32_t w, int32_t h,
                 const NbrCam *nbrs, int32_t nn, int32_t rule, int32_t need, double inv_voxel,
                 const uint8_t *rgb, const uint32_t *quad, const float *gray, const FuseTable &t,
                 hipStream_t st) {
    const dim3 grid((unsigned)((w + kTile - 1) / kTile), (unsigned)((h + kTile - 1) / kTile));
    hipLaunchKernelGGL(k_fuse, grid, dim3(kTile, kTile), 0, st, ref, depth, w, h, nbrs, nn, rule,
                       need, inv_voxel, rgb, quad, gray, t);
}

}  // namespace amvs
