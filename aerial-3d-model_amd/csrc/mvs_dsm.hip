// Digital surface model from the dense cloud (crates/stitch-core/src/recon/
// dsm.rs:128-290) on the GPU: per-cell second-highest height, despike,
// island removal, enclosed-hole filling, small-hole filling and a 3x3 median.
// Every step reads a snapshot and writes per cell (or is an order-free
// reduction), so the result does not depend on thread order and equals the
// reference's sequential loops.
#include "mvs_internal.h"

#include <algorithm>
#include <math.h>
#include <string.h>

namespace amvs {
namespace {

// Monotone key of a non-NaN float for unsigned atomics (larger float, larger key).
__device__ __forceinline__ uint32_t fkey(float v) {
    const uint32_t b = __float_as_uint(v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float fval(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}
// f32::total_cmp as a signed integer key (the order the reference sorts by).
__device__ __forceinline__ int32_t tkey(float v) {
    int32_t x = __float_as_int(v);
    return x ^ (int32_t)(((uint32_t)(x >> 31)) >> 1);
}

struct Grid {
    double x0, y0, cell;
    int32_t nx, ny;
};

__device__ __forceinline__ bool point_cell(const Grid &g, const float *p, int32_t &k) {
    // dsm.rs:153-158: floor((p - corner) / cell) in f64, cells outside dropped
    const double i = floor(((double)p[0] - g.x0) / g.cell);
    const double j = floor(((double)p[1] - g.y0) / g.cell);
    if (i < 0.0 || j < 0.0 || i >= (double)g.nx || j >= (double)g.ny) return false;
    // a NaN coordinate passes the test above; Rust's `as usize` makes it 0
    k = (int32_t)(isnan(j) ? 0.0 : j) * g.nx + (int32_t)(isnan(i) ? 0.0 : i);
    return true;
}

// Pass 1: point count and highest height per cell.
__global__ void k_dsm_top1(Grid g, const float *__restrict__ xyz, int64_t n,
                           uint32_t *__restrict__ cnt, uint32_t *__restrict__ top1) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int32_t k;
    if (!point_cell(g, xyz + 3 * p, k)) return;
    atomicAdd(&cnt[k], 1u);
    const float z = xyz[3 * p + 2];
    if (!isnan(z)) atomicMax(&top1[k], fkey(z));  // NaN never beats a height
}

// Pass 2: how many points equal the highest, and the highest of the rest:
// the second entry of the reference's top-three (dsm.rs:160-169).
__global__ void k_dsm_top2(Grid g, const float *__restrict__ xyz, int64_t n,
                           const uint32_t *__restrict__ top1, uint32_t *__restrict__ eq,
                           uint32_t *__restrict__ top2) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int32_t k;
    if (!point_cell(g, xyz + 3 * p, k)) return;
    const float z = xyz[3 * p + 2];
    if (isnan(z)) return;
    const float t = fval(top1[k]);
    if (z == t)
        atomicAdd(&eq[k], 1u);
    else
        atomicMax(&top2[k], fkey(z));
}

// dsm.rs:171-178: none -> NaN, one -> it, more -> the second highest.
__global__ void k_dsm_heights(int64_t ncell, const uint32_t *__restrict__ cnt,
                              const uint32_t *__restrict__ top1, const uint32_t *__restrict__ eq,
                              const uint32_t *__restrict__ top2, float *__restrict__ h) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncell) return;
    const uint32_t c = cnt[k];
    h[k] = c == 0 ? NAN : (c == 1 || eq[k] >= 2 ? fval(top1[k]) : fval(top2[k]));
}

// dsm.rs:183-201: median (total order) of the finite values in the window of
// radius R; false when fewer than 5. The element of rank n/2 in the stable
// total-order sort is found by counting, fully unrolled, so the window stays
// in registers.
template <int R>
__device__ bool window_median(const float *__restrict__ h, int32_t nx, int32_t ny, int32_t i,
                              int32_t j, float &m) {
    constexpr int D = 2 * R + 1, W = D * D;
    float v[W];
    int32_t key[W];
    bool ok[W];
    int n = 0;
#pragma unroll
    for (int q = 0; q < W; ++q) {
        const int32_t a = i + q % D - R, b = j + q / D - R;
        const bool inb = a >= 0 && b >= 0 && a < nx && b < ny;
        const float x = inb ? h[(int64_t)b * nx + a] : NAN;
        ok[q] = isfinite(x);
        key[q] = tkey(x);
        v[q] = x;
        n += ok[q] ? 1 : 0;
    }
    if (n < 5) return false;
    const int want = n / 2;
    float res = 0.0f;
#pragma unroll
    for (int q = 0; q < W; ++q) {
        int rank = 0;
#pragma unroll
        for (int p = 0; p < W; ++p) {
            if (p == q) continue;
            // values ahead of q in the stable sort: smaller keys, or equal
            // keys met earlier in the reference's row-major window order
            const bool ahead = p < q ? key[p] <= key[q] : key[p] < key[q];
            rank += (ok[p] && ahead) ? 1 : 0;
        }
        if (ok[q] && rank == want) res = v[q];
    }
    m = res;
    return true;
}

// dsm.rs:202-216
__global__ void k_dsm_despike(int32_t nx, int32_t ny, const float *__restrict__ snap,
                              const uint32_t *__restrict__ cnt, float *__restrict__ h) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)nx * ny) return;
    const float s = snap[k];
    if (!isfinite(s)) return;
    const int32_t j = (int32_t)(k / nx), i = (int32_t)(k - (int64_t)j * nx);
    float m;
    if (window_median<2>(snap, nx, ny, i, j, m)) {
        if (fabsf(s - m) > 3.0f && cnt[k] <= 2) h[k] = m;
    } else if (cnt[k] <= 2) {
        h[k] = NAN;
    }
}

// dsm.rs:226-257: one pass of small-hole filling (8 neighbours, at least 3).
__global__ void k_dsm_small_holes(int32_t nx, int32_t ny, const float *__restrict__ snap,
                                  float *__restrict__ h) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)nx * ny) return;
    if (isfinite(snap[k])) return;
    const int32_t j = (int32_t)(k / nx), i = (int32_t)(k - (int64_t)j * nx);
    const int d[8][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}, {-1, -1}, {1, 1}, {-1, 1}, {1, -1}};
    float s = 0.0f;
    int c = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int32_t a = i + d[q][0], b = j + d[q][1];
        if (a >= 0 && b >= 0 && a < nx && b < ny) {
            const float x = snap[(int64_t)b * nx + a];
            if (isfinite(x)) {
                s += x;
                ++c;
            }
        }
    }
    if (c >= 3) h[k] = s / (float)c;
}

// dsm.rs:259-270
__global__ void k_dsm_median3(int32_t nx, int32_t ny, const float *__restrict__ snap,
                              float *__restrict__ h) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)nx * ny) return;
    if (!isfinite(snap[k])) return;
    const int32_t j = (int32_t)(k / nx), i = (int32_t)(k - (int64_t)j * nx);
    float m;
    if (window_median<1>(snap, nx, ny, i, j, m)) h[k] = m;
}

// ---- 4-connected components (lock-free union-find, roots hook to smaller
// indices, so parent chains always descend and every find ends) ----

// Loads and stores at device scope: other workgroups hook roots and shorten
// paths while this one walks, and a stale cached parent could make a retry
// loop spin.
__device__ __forceinline__ int32_t cc_parent(const int32_t *parent, int32_t x) {
    return __hip_atomic_load(parent + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The root of x, halving the path on the way: a node's parent is replaced by
// its grandparent, an ancestor, so every chain still descends to the root.
__device__ __forceinline__ int32_t cc_find(int32_t *parent, int32_t x) {
    int32_t p = cc_parent(parent, x);
    while (p != x) {
        const int32_t g = cc_parent(parent, p);
        if (g != p)
            __hip_atomic_store(parent + x, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = p;
        p = g;
    }
    return x;
}

// The root of x without writing (for passes where no root changes any more).
__device__ __forceinline__ int32_t cc_root(const int32_t *parent, int32_t x) {
    int32_t p = cc_parent(parent, x);
    while (p != x) {
        x = p;
        p = cc_parent(parent, x);
    }
    return x;
}

__device__ void cc_unite(int32_t *parent, int32_t a, int32_t b) {
    while (true) {
        a = cc_find(parent, a);
        b = cc_find(parent, b);
        if (a == b) return;
        if (a < b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        // hook root a under b, unless a stopped being a root meanwhile
        if (atomicCAS(&parent[a], a, b) == a) return;
    }
}

// member(k): finite heights (want_finite) or empty cells
__device__ __forceinline__ bool cc_member(const float *h, int64_t k, bool want_finite) {
    return isfinite(h[k]) == want_finite;
}

// Each member starts under its left neighbour when that is a member, else
// under its upper one, else as a root (both have smaller indices); -1 for
// other cells.
__global__ void k_cc_init(int32_t nx, int32_t ny, const float *__restrict__ h, bool want_finite,
                          int32_t *__restrict__ parent) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)nx * ny) return;
    // branch-free: both neighbours are read at clamped, in-bounds indices and
    // the parent is a select (hipcc 7.2 dropped the root case of the nested
    // if/else form for some lanes)
    const int32_t kk = (int32_t)k, i = (int32_t)(k % nx);
    const int32_t kl = i > 0 ? kk - 1 : kk, ku = kk >= nx ? kk - nx : kk;
    const bool m = cc_member(h, kk, want_finite);
    const bool left = i > 0 && cc_member(h, kl, want_finite);
    const bool up = kk >= nx && cc_member(h, ku, want_finite);
    const int32_t p = left ? kl : (up ? ku : kk);
    parent[k] = m ? p : -1;
}

// Pointer jumping on the initial row chains (no root changes yet): each pass
// replaces a parent by its parent, so after log2(width) passes every member
// points at the start of its run or close to it, and the finds that follow
// are short.
__global__ void k_cc_jump(int64_t ncell, int32_t *parent) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncell) return;
    const int32_t p = cc_parent(parent, (int32_t)k);
    if (p < 0) return;
    const int32_t g = cc_parent(parent, p);
    if (g != p) __hip_atomic_store(parent + k, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The adjacencies the initial parents do not cover: a member whose left and
// upper neighbours are both members is joined to the upper one — only at the
// first cell of each stretch where two rows' runs touch: further along, the
// cell's left neighbour is already joined to its own upper neighbour, and the
// row links carry that over (k ~ k-1 ~ k-1-nx ~ k-nx).
__global__ void k_cc_hook(int32_t nx, int32_t ny, const float *__restrict__ h, bool want_finite,
                          int32_t *parent) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)nx * ny || k < nx || !cc_member(h, k, want_finite)) return;
    const int32_t i = (int32_t)(k % nx);
    if (i > 0 && cc_member(h, k - 1, want_finite) && cc_member(h, k - nx, want_finite) &&
        !cc_member(h, k - 1 - nx, want_finite))
        cc_unite(parent, (int32_t)k, (int32_t)(k - nx));
}

// Path halving from every member (no root changes any more): the walks of
// the flatten pass that follows become short.
__global__ void k_cc_compress(int64_t ncell, int32_t *parent) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncell || cc_parent(parent, (int32_t)k) < 0) return;
    (void)cc_find(parent, (int32_t)k);
}

// Every member points at its root afterwards; sizes per root. No root
// changes here, and each thread writes only its own cell (the root, an
// ancestor of every value it replaces), so concurrent walks stay valid.
__global__ void k_cc_flatten(int64_t ncell, int32_t *parent, uint32_t *__restrict__ size) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncell || cc_parent(parent, (int32_t)k) < 0) return;
    const int32_t r = cc_root(parent, (int32_t)k);
    __hip_atomic_store(parent + k, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (size) {
        // most lanes of a wave share a root (one large component): they add
        // their count with one atomic, the others one each
        const int32_t lead = __builtin_amdgcn_readfirstlane(r);
        const unsigned long long same = __ballot(r == lead);
        if (r != lead)
            atomicAdd(&size[r], 1u);
        else if ((int)__lane_id() == __ffsll((long long)same) - 1)
            atomicAdd(&size[lead], (uint32_t)__popcll(same));
    }
}

// dsm.rs:293-331: components smaller than min_cells become empty.
__global__ void k_cc_drop_small(int64_t ncell, const int32_t *parent,
                                const uint32_t *__restrict__ size, uint32_t min_cells,
                                float *__restrict__ h) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncell || parent[k] < 0) return;
    if (size[parent[k]] < min_cells) h[k] = NAN;  // flattened: parent is the root
}

// dsm.rs:337-359: empty components touching the border are "outside".
__global__ void k_cc_mark_outside(int32_t nx, int32_t ny, const int32_t *parent,
                                  uint32_t *__restrict__ outside) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nb = 2 * (int64_t)nx + 2 * (int64_t)ny;
    if (t >= nb) return;
    int64_t k;
    if (t < nx)
        k = t;
    else if (t < 2 * (int64_t)nx)
        k = (int64_t)(ny - 1) * nx + (t - nx);
    else if (t < 2 * (int64_t)nx + ny)
        k = (t - 2 * (int64_t)nx) * nx;
    else
        k = (t - 2 * (int64_t)nx - ny) * nx + nx - 1;
    if (parent[k] >= 0) outside[parent[k]] = 1u;  // flattened: parent is the root
}

// The enclosed empty cells: the first to-do list of dsm.rs:360-362.
__global__ void k_cc_enclosed(int64_t ncell, const int32_t *parent,
                              const uint32_t *__restrict__ outside, int32_t *__restrict__ todo,
                              uint32_t *__restrict__ ntodo) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncell || parent[k] < 0) return;
    if (!outside[parent[k]]) todo[atomicAdd(ntodo, 1u)] = (int32_t)k;  // flattened
}

// One pass of dsm.rs:363-389: every listed cell with a finite 4-neighbour in
// the snapshot takes their mean; the others go to the next list.
__global__ void k_fill_pass(int32_t nx, int32_t ny, const float *__restrict__ snap,
                            float *__restrict__ h, const int32_t *__restrict__ todo,
                            const uint32_t *__restrict__ ntodo, int32_t *__restrict__ next,
                            uint32_t *__restrict__ nnext) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)*ntodo) return;
    const int32_t k = todo[t];
    const int32_t j = k / nx, i = k - j * nx;
    const int d[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
    float s = 0.0f;
    int c = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int32_t a = i + d[q][0], b = j + d[q][1];
        if (a >= 0 && b >= 0 && a < nx && b < ny) {
            const float x = snap[(int64_t)b * nx + a];
            if (isfinite(x)) {
                s += x;
                ++c;
            }
        }
    }
    if (c > 0)
        h[k] = s / (float)c;
    else
        next[atomicAdd(nnext, 1u)] = k;
}

// The highest finite height: a grid-stride pass per thread, then the
// block's maximum, one atomic per block.
__global__ __launch_bounds__(256) void k_dsm_zmax(int64_t ncell, const float *__restrict__ h,
                                                  uint32_t *__restrict__ zk) {
    __shared__ float wmax[4];
    float v = -INFINITY;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < ncell;
         k += (int64_t)gridDim.x * blockDim.x)
        if (isfinite(h[k])) v = fmaxf(v, h[k]);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float m = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
        if (m > -INFINITY) atomicMax(zk, fkey(m));
    }
}

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

// ---- order statistics of the coordinates (dsm.rs:140-143) by radix select
// over the total-order keys: three digit passes (11, 11 and 10 bits from the
// top), each a histogram of the keys that match the digits found so far.
// Targets: 0 x lo, 1 x hi, 2 y lo, 3 y hi. ----

constexpr int kSelBins = 2048;

struct SelState {
    uint32_t prefix[4];  // the key's digits found so far
    uint32_t rank[4];    // rank still to find among the keys with that prefix
};

__device__ __forceinline__ uint32_t ukey(float v) { return (uint32_t)tkey(v) ^ 0x80000000u; }

__device__ __forceinline__ uint32_t sel_digit(uint32_t u, int pass) {
    return pass == 0 ? u >> 21 : (pass == 1 ? (u >> 10) & 2047u : u & 1023u);
}
__device__ __forceinline__ uint32_t sel_mask(int pass) {
    return pass == 0 ? 0u : (pass == 1 ? 0xFFE00000u : 0xFFFFFC00u);
}

__global__ __launch_bounds__(256) void k_sel_hist(const float *__restrict__ xyz, int64_t n,
                                                  int pass, const SelState *__restrict__ st,
                                                  uint32_t *__restrict__ hist) {
    __shared__ uint32_t lh[4 * kSelBins];
    for (int b = threadIdx.x; b < 4 * kSelBins; b += 256) lh[b] = 0u;
    __syncthreads();
    const uint32_t mask = sel_mask(pass);
    const uint32_t p0 = st->prefix[0], p1 = st->prefix[1], p2 = st->prefix[2],
                   p3 = st->prefix[3];
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * 256) {
        const uint32_t ux = ukey(xyz[3 * p]), uy = ukey(xyz[3 * p + 1]);
        const uint32_t dx = sel_digit(ux, pass), dy = sel_digit(uy, pass);
        if ((ux & mask) == p0) atomicAdd(&lh[dx], 1u);
        if ((ux & mask) == p1) atomicAdd(&lh[kSelBins + dx], 1u);
        if ((uy & mask) == p2) atomicAdd(&lh[2 * kSelBins + dy], 1u);
        if ((uy & mask) == p3) atomicAdd(&lh[3 * kSelBins + dy], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < 4 * kSelBins; b += 256)
        if (lh[b]) atomicAdd(&hist[b], lh[b]);
}

// One block per target: the digit whose bin holds the rank, then the bins
// cleared for the next pass. The rank is below the matching keys' count, so
// the walk ends inside the histogram.
__global__ __launch_bounds__(256) void k_sel_scan(int pass, SelState *st,
                                                  uint32_t *__restrict__ hist) {
    const int t = blockIdx.x;
    uint32_t *h = hist + t * kSelBins;
    __shared__ uint32_t part[256];
    __shared__ uint32_t found[2];
    uint32_t v[8], s = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        v[i] = h[threadIdx.x * 8 + i];
        s += v[i];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t r = st->rank[t];
        uint32_t c = 0;
        while (c < 255 && r >= part[c]) {
            r -= part[c];
            ++c;
        }
        found[0] = c;
        found[1] = r;
    }
    __syncthreads();
    if (threadIdx.x == found[0]) {
        // the first bin of this thread's eight that holds the rank
        uint32_t r = found[1], d = 8;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if (d == 8u) {
                if (r < v[i] || i == 7)
                    d = (uint32_t)i;
                else
                    r -= v[i];
            }
        }
        const uint32_t digit = threadIdx.x * 8 + d;
        const int shift = pass == 0 ? 21 : (pass == 1 ? 10 : 0);
        st->prefix[t] |= digit << shift;
        st->rank[t] = r;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) h[threadIdx.x * 8 + i] = 0u;
}

__global__ void k_sel_init(SelState *st, uint32_t rank_lo, uint32_t rank_hi) {
    const int t = threadIdx.x;
    if (t < 4) {
        st->prefix[t] = 0u;
        st->rank[t] = (t & 1) ? rank_hi : rank_lo;
    }
}

unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

uint32_t dsm_fkey_host(float v) {
    uint32_t b;
    memcpy(&b, &v, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

float dsm_fval_host(uint32_t k) {
    const uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float v;
    memcpy(&v, &b, 4);
    return v;
}

int dsm_extent(const float *d_xyz, int64_t n, uint32_t rank_lo, uint32_t rank_hi, void *d_sel,
               hipStream_t st, float out[4]) {
    uint32_t *hist = (uint32_t *)d_sel;
    SelState *state = (SelState *)(hist + 4 * kSelBins);
    hipLaunchKernelGGL(k_fill_u32, dim3(nblk(4 * kSelBins)), dim3(256), 0, st, hist,
                       (int64_t)4 * kSelBins, 0u);
    hipLaunchKernelGGL(k_sel_init, dim3(1), dim3(64), 0, st, state, rank_lo, rank_hi);
    const unsigned blocks = (unsigned)std::min<int64_t>(nblk(n), 1024);
    for (int pass = 0; pass < 3; ++pass) {
        hipLaunchKernelGGL(k_sel_hist, dim3(blocks), dim3(256), 0, st, d_xyz, n, pass, state,
                           hist);
        hipLaunchKernelGGL(k_sel_scan, dim3(4), dim3(256), 0, st, pass, state, hist);
    }
    SelState h;
    if (hipMemcpyAsync(&h, state, sizeof(h), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return -1;
    for (int t = 0; t < 4; ++t) {
        // invert ukey: back to the total-order key, then to the float's bits
        const int32_t k = (int32_t)(h.prefix[t] ^ 0x80000000u);
        const int32_t x = k ^ (int32_t)(((uint32_t)(k >> 31)) >> 1);
        memcpy(&out[t], &x, 4);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int dsm_build(const DsmGrid &g, const float *d_xyz, int64_t n, float *d_h, DsmWork &w,
              hipStream_t st, float *zmax) {
    const Grid gg{g.x0, g.y0, g.cell, (int32_t)g.nx, (int32_t)g.ny};
    const int64_t nc = g.nx * g.ny;
    uint32_t *cnt = w.u32a, *top1 = w.u32b, *eq = w.u32c, *top2 = w.u32d;
    hipLaunchKernelGGL(k_fill_u32, dim3(nblk(nc)), dim3(256), 0, st, cnt, nc, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(nblk(nc)), dim3(256), 0, st, eq, nc, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(nblk(nc)), dim3(256), 0, st, top1, nc,
                       dsm_fkey_host(-INFINITY));
    hipLaunchKernelGGL(k_fill_u32, dim3(nblk(nc)), dim3(256), 0, st, top2, nc,
                       dsm_fkey_host(-INFINITY));
    if (n > 0) {
        hipLaunchKernelGGL(k_dsm_top1, dim3(nblk(n)), dim3(256), 0, st, gg, d_xyz, n, cnt, top1);
        hipLaunchKernelGGL(k_dsm_top2, dim3(nblk(n)), dim3(256), 0, st, gg, d_xyz, n, top1, eq,
                           top2);
    }
    hipLaunchKernelGGL(k_dsm_heights, dim3(nblk(nc)), dim3(256), 0, st, nc, cnt, top1, eq, top2,
                       d_h);
    float *snap = w.snap;
    const size_t hb = (size_t)nc * sizeof(float);
    // despike
    if (hipMemcpyAsync(snap, d_h, hb, hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
    hipLaunchKernelGGL(k_dsm_despike, dim3(nblk(nc)), dim3(256), 0, st, gg.nx, gg.ny, snap, cnt,
                       d_h);
    // islands (dsm.rs:220): 4-connected components of finite cells
    int32_t *parent = w.parent;
    // ceil(log2(nx + ny)) + 1 passes: the initial chains run along rows and
    // up the first cells of runs (longer chains are only shortened, and every
    // walk still ends)
    int jumps = 1;
    while ((1ll << (jumps - 1)) < g.nx + g.ny) ++jumps;
    uint32_t *size = w.u32b;  // top1 is no longer needed
    hipLaunchKernelGGL(k_fill_u32, dim3(nblk(nc)), dim3(256), 0, st, size, nc, 0u);
    hipLaunchKernelGGL(k_cc_init, dim3(nblk(nc)), dim3(256), 0, st, gg.nx, gg.ny, d_h, true,
                       parent);
    for (int pass = 0; pass < jumps; ++pass)
        hipLaunchKernelGGL(k_cc_jump, dim3(nblk(nc)), dim3(256), 0, st, nc, parent);
    hipLaunchKernelGGL(k_cc_hook, dim3(nblk(nc)), dim3(256), 0, st, gg.nx, gg.ny, d_h, true,
                       parent);
    hipLaunchKernelGGL(k_cc_compress, dim3(nblk(nc)), dim3(256), 0, st, nc, parent);
    hipLaunchKernelGGL(k_cc_flatten, dim3(nblk(nc)), dim3(256), 0, st, nc, parent, size);
    const double mc = 25.0 / (g.cell * g.cell);
    const uint32_t min_cells = mc >= 4294967295.0 ? 0xffffffffu : (uint32_t)mc;
    hipLaunchKernelGGL(k_cc_drop_small, dim3(nblk(nc)), dim3(256), 0, st, nc, parent, size,
                       min_cells, d_h);
    // enclosed holes (dsm.rs:222): components of empty cells off the border
    uint32_t *outside = w.u32c;  // eq is no longer needed
    hipLaunchKernelGGL(k_fill_u32, dim3(nblk(nc)), dim3(256), 0, st, outside, nc, 0u);
    hipLaunchKernelGGL(k_cc_init, dim3(nblk(nc)), dim3(256), 0, st, gg.nx, gg.ny, d_h, false,
                       parent);
    for (int pass = 0; pass < jumps; ++pass)
        hipLaunchKernelGGL(k_cc_jump, dim3(nblk(nc)), dim3(256), 0, st, nc, parent);
    hipLaunchKernelGGL(k_cc_hook, dim3(nblk(nc)), dim3(256), 0, st, gg.nx, gg.ny, d_h, false,
                       parent);
    hipLaunchKernelGGL(k_cc_compress, dim3(nblk(nc)), dim3(256), 0, st, nc, parent);
    hipLaunchKernelGGL(k_cc_flatten, dim3(nblk(nc)), dim3(256), 0, st, nc, parent,
                       (uint32_t *)nullptr);
    hipLaunchKernelGGL(k_cc_mark_outside, dim3(nblk(2 * (g.nx + g.ny))), dim3(256), 0, st, gg.nx,
                       gg.ny, parent, outside);
    uint32_t *counts = w.counters;  // [0] list a, [1] list b, [2] zmax key
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, counts, 3, 0u);
    hipLaunchKernelGGL(k_cc_enclosed, dim3(nblk(nc)), dim3(256), 0, st, nc, parent, outside,
                       w.list_a, counts);
    uint32_t left = 0;
    if (hipMemcpyAsync(&left, counts, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return -1;
    int32_t *cur = w.list_a, *nxt = w.list_b;
    uint32_t *ccur = counts, *cnxt = counts + 1;
    for (int guard = 0; left > 0 && guard < 10000; ++guard) {
        if (hipMemcpyAsync(snap, d_h, hb, hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, cnxt, 1, 0u);
        hipLaunchKernelGGL(k_fill_pass, dim3(nblk(left)), dim3(256), 0, st, gg.nx, gg.ny, snap,
                           d_h, cur, ccur, nxt, cnxt);
        if (hipMemcpyAsync(&left, cnxt, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return -1;
        int32_t *t = cur;
        cur = nxt;
        nxt = t;
        uint32_t *tc = ccur;
        ccur = cnxt;
        cnxt = tc;
    }
    // small holes: 8 passes (dsm.rs:225-257)
    for (int pass = 0; pass < 8; ++pass) {
        if (hipMemcpyAsync(snap, d_h, hb, hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
        hipLaunchKernelGGL(k_dsm_small_holes, dim3(nblk(nc)), dim3(256), 0, st, gg.nx, gg.ny, snap,
                           d_h);
    }
    // 3x3 median
    if (hipMemcpyAsync(snap, d_h, hb, hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
    hipLaunchKernelGGL(k_dsm_median3, dim3(nblk(nc)), dim3(256), 0, st, gg.nx, gg.ny, snap, d_h);
    // zmax (dsm.rs:271-275): f32::MIN when there is no surface
    hipLaunchKernelGGL(k_dsm_zmax, dim3((unsigned)std::min<int64_t>(nblk(nc), 1024)), dim3(256), 0,
                       st, nc, d_h, counts + 2);
    uint32_t zk = 0;
    if (hipMemcpyAsync(&zk, counts + 2, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return -1;
    *zmax = zk ? dsm_fval_host(zk) : -3.40282347e38f;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace amvs
