// PatchMatch with one lane per reference pixel, for the fast arithmetic
// (MVS_MODE_FAST) on undistorted cameras: the cost of dense.rs:181-263 and the
// half-sweep of :413-467, with the rounding rules of the fast source-lane
// kernels (contracted multiply-adds, hardware reciprocals, sums shifted by the
// window's first sample), laid out so that every gather instruction reads one
// source image: the wavefront's 64 lanes hold 64 pixels of one reference view
// and step through the sources together, so one gather touches the few cache
// lines under 64 neighbouring pixels rather than lines of eight images.
//
// Window geometry. For an undistorted reference camera the window rays are
// affine in the tap offsets, r = r_c + dx e_x + dy e_y (e = tap spacing / f).
// On the plane n . X = c (c = d n . r_c, dense.rs:187-190) the window point of
// ray r is X = r c / (n . r), and in the source camera
//     X_s = R X + t = (c / (n . r)) (R + t n^T / c) r = dq H r,
// dq = c / (n . r) the window depth of dense.rs:201-211. Scaled by the centre
// tap's depth dq_c, the source point is affine in (dx, dy):
//     (dq_c / dq) X_s = X_c + dx dq_c H e_x + dy dq_c H e_y,
// X_c the centre tap's source point (formed as the source-lane kernels form
// every tap's, for their accuracy), dq_c H e = R e dq_c + t (n . e) / (n . r_c).
// Three multiply-adds per tap give it; its third coordinate has the sign of
// X_s.z (dq > 0), and the pixel is f times the first two over the third plus
// the principal point, added last so that its rounding stays that of one add.
// dense.rs requires |n . r| >= 1e-6 and dq > 0 at every tap; n . r is affine
// over the window, so both hold at every tap when they hold at the four
// corner taps.
#include <type_traits>

#include "mvs_pm_impl.h"

typedef uint32_t amvs_u32x2 __attribute__((ext_vector_type(2)));
// Structured buffer loads (the LLVM intrinsics; clang has no builtin for
// them): the address is base + vindex * stride, so a pixel index needs no
// scaling to bytes; num_records counts records.
__device__ amvs_u32x2 amvs_struct_load_b64(__amdgpu_buffer_rsrc_t rsrc, int32_t vindex,
                                           int32_t voffset, int32_t soffset, int32_t aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.v2i32");
__device__ uint32_t amvs_struct_load_b32(__amdgpu_buffer_rsrc_t rsrc, int32_t vindex,
                                         int32_t voffset, int32_t soffset, int32_t aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.i32");
typedef float amvs_f32x4 __attribute__((ext_vector_type(4)));
__device__ amvs_f32x4 amvs_struct_load_f4(__amdgpu_buffer_rsrc_t rsrc, int32_t vindex,
                                          int32_t voffset, int32_t soffset, int32_t aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.v4f32");
// Raw buffer load: the address is base + voffset + soffset (bytes).
__device__ amvs_u32x2 amvs_raw_load_b64(__amdgpu_buffer_rsrc_t rsrc, int32_t voffset,
                                        int32_t soffset, int32_t aux)
    __asm("llvm.amdgcn.raw.ptr.buffer.load.v2i32");

namespace amvs {
namespace {

typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
// Word 3 of a buffer resource on gfx9-family chips (32-bit data format, no
// swizzle): reads past num_records return zero.
constexpr int32_t kBufferWord3 = 0x00020000;

// Half-sweeps of the first iterations (random planes) score their
// compacted candidates in order of depth, in kDepthBins bins of inverse depth
// over the job's range: a batch of 64 then holds candidates of similar depth
// from the wavefront's block of pixels, whose windows land close together in
// each source (shared cache lines, and in or out of the source's coverage
// together). Later iterations list each pixel's candidates together (they
// agree among themselves by then).
#ifndef AMVS_SORT_ITERS
#define AMVS_SORT_ITERS 2
#endif
constexpr int kSortIters = AMVS_SORT_ITERS;
// The finer the bins, the closer neighbouring lanes' depths (77-view
// half-sweeps: 16 bins 1807 ms, 64 bins 1757, 128 bins 1749, 256 bins 1747
// with a register spilled).
#ifndef AMVS_DEPTH_BINS
#define AMVS_DEPTH_BINS 128
#endif
constexpr int kDepthBins = AMVS_DEPTH_BINS;  // a power of two, at most 256
constexpr int kBinBits = kDepthBins <= 16    ? 4
                         : kDepthBins <= 32  ? 5
                         : kDepthBins <= 64  ? 6
                         : kDepthBins <= 128 ? 7
                                             : 8;  // per candidate in a 64-bit word
constexpr int kBinsPerLane = kDepthBins > 64 ? kDepthBins / 64 : 1;
static_assert(kDepthBins == 1 << kBinBits || kDepthBins < 16, "bins: a power of two");

// PL_F16T: the f16 coefficient plane staged in LDS per (batch, source) when the
// batch's windows fit a tile of kTileCap texels (8 bytes each) and lie
// inside the source: a few coalesced loads replace the per-tap gathers,
// and the taps read LDS. Batches that do not fit gather as PL_F16 does.
constexpr int kTileCap = 512;

// A window whose four corner samples lie this far (pixels) inside the
// source's sampling range keeps every tap inside it: rounding moves an inner
// tap's position by a few 1e-4 pixels at most (ulps of coordinates below
// 4096), so such windows need no clamp.
constexpr float kInnerMargin = 1.0f / 256.0f;

// Pixel of propagation candidate k of (x, y) (dense.rs:403-412, CAND order:
// (-1,0), (1,0), (0,-1), (0,1), then the same at distance 5).
__device__ __forceinline__ void cand_pos(int k, int x, int y, int &nx, int &ny) {
    const int dir = k & 3, mag = k < 4 ? 1 : 5;
    nx = x + (dir == 0 ? -mag : (dir == 1 ? mag : 0));
    ny = y + (dir == 2 ? -mag : (dir == 3 ? mag : 0));
}

// dense.rs:436: the candidate equals the hypothesis (f32 ==, NaN unequal).
__device__ __forceinline__ bool same_hyp(const float4 &a, const float4 &b) {
    return a.x == b.x && a.y == b.y && a.z == b.z && a.w == b.w;
}

// A cost as an unsigned key with the order of the floats (negative costs
// from rounding included, -0 taken as +0 as == does; NaN above every
// number), and back.
__device__ __forceinline__ uint32_t cost_key(float c) {
    const uint32_t b = __float_as_uint(c + 0.0f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float key_cost(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// PLANE: PL_F16 samples the 8-byte f16 bilinear-coefficient plane (fewer ALU
// operations per sample), PL_U8 the 4-byte 8-bit corner plane (half the bytes per
// gather) — both give the same values, from 8-bit views, in byte units; PL_F32
// samples the 16-byte f32 coefficient plane of f32 grey views (the
// `.view` cache's own grey, dense.rs:104-111), in grey units; PL_F32G reads
// the same views' 4-byte grey plane as two row pairs per sample (the same
// coefficients, formed the same way, so the same costs, from a quarter of
// the plane bytes: the scattered first iterations miss the caches less).
// Waves per SIMD the 7x7 and 5x5 kernels on the 8-bit and f16 planes are
// built for at least: four (128 registers), which the half-sweep's register
// allocation would otherwise miss by a register or two; the others keep the
// compiler's choice.
template <int TAPS, int PLANE>
constexpr int kPmlWavesMin = (TAPS <= 7 && (PLANE == PL_F16 || PLANE == PL_U8)) ? 4 : 1;

template <int TAPS, int MODE, bool PERSIST, int PLANE>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(kPmlWavesMin<TAPS, PLANE>))) void k_pml(PmArgs a) {
    using namespace dev;
    constexpr int N = TAPS * TAPS;
    constexpr int H = TAPS / 2;
    constexpr bool F32 = PLANE == PL_F32 || PLANE == PL_F32G;
    constexpr bool TILE = PLANE == PL_F16T;
    constexpr bool F16Q = PLANE == PL_F16 || TILE;  // samples the f16 coefficient plane
    // each lane's reference window as 8-bit values less an integer near their
    // mean (exact in f16; f32 views: the grey less the window mean, f32),
    // tap-major so that the per-tap reads of a wavefront are free of bank
    // conflicts
    __shared__ _Float16 Ah[F32 ? 1 : N][64];
    __shared__ float Af[F32 ? N : 1][64];
    // half-sweeps: the wavefront's propagation candidates as one list of
    // (pixel lane, candidate) entries, and each pixel's lowest (cost, index)
    __shared__ uint16_t plist[9 * 64];  // up to eight candidates and the own plane per pixel
    __shared__ unsigned long long pmin[64];
    __shared__ int32_t pbin[kDepthBins];
    __shared__ u32x2 tile[TILE ? kTileCap : 1];

    const int lane = threadIdx.x;
    const int chunk = 64 * a.gpc;
    const int32_t nj = a.njobs;
    const int home = blockIdx.x % kNumXcd;
    unsigned long long n_evals = 0;
    uint32_t n_skipped = 0;  // wave-uniform (a scalar register)

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
            const RefJob job = scalar_load(a.jobs + jj);
            const int64_t items = pm_items(job, a, MODE);
            const int64_t base = (int64_t)(c - a.prefix[jj]) * chunk;
            const int32_t w = job.w, h = job.h, K = job.nsrc, stride = job.stride,
                          half = job.half;
            const float rstep = job.ray_step;

            for (int b = 0; b < a.gpc; ++b) {
                const int64_t item0 = base + (int64_t)b * 64;
                if (item0 >= items) break;  // uniform
                const int64_t item = item0 + lane;
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

                // reference window statistics (dense.rs:213, mean_sd :266-271,
                // sums in tap order). The window's grey is the 8-bit value over
                // 255 (to the table's rounding), so its bytes b are rint(255 g);
                // the costs use a = b - mi (mi an integer near the byte mean)
                // and the exact correction delta = mean(b) - mi.
                float mr = 0.0f, sr = 0.0f, delta = 0.0f;
                // (In the sweep the window's loads issue together, the
                // first pass's all at once and the second pass's a row at a
                // time: 28 ms less per 77-view f32 run; the initial-cost
                // pass, one evaluation per pixel, keeps plain row loops,
                // which measured faster there.)
                if (inside && MODE == MODE_SWEEP) {
                    float s = 0.0f, sb = 0.0f, mi = 0.0f, var = 0.0f;
                    auto tap_grey = [&](int ty, int tx) -> float {
                        return as_global(job.gray)[pidx + (int64_t)((ty - H) * stride) * w +
                                                   (tx - H) * stride];
                    };
                    auto second_pass = [&](int ty, int tx, float g) {
                        const float d = g - mr;
                        var = var + d * d;
                        if constexpr (F32)
                            Af[ty * TAPS + tx][lane] = d;
                        else
                            Ah[ty * TAPS + tx][lane] = (_Float16)(rintf(g * 255.0f) - mi);
                    };
                    if constexpr (N <= 49) {
                        // the first pass's loads all issue together
                        float gw[N];
#pragma unroll
                        for (int t = 0; t < N; ++t) gw[t] = tap_grey(t / TAPS, t % TAPS);
#pragma unroll
                        for (int t = 0; t < N; ++t) {
                            s = s + gw[t];
                            sb += rintf(gw[t] * 255.0f);
                        }
                    } else {
#pragma unroll 1
                        for (int ty = 0; ty < TAPS; ++ty) {
                            float gr[TAPS];
#pragma unroll
                            for (int tx = 0; tx < TAPS; ++tx) gr[tx] = tap_grey(ty, tx);
#pragma unroll
                            for (int tx = 0; tx < TAPS; ++tx) {
                                s = s + gr[tx];
                                sb += rintf(gr[tx] * 255.0f);
                            }
                        }
                    }
                    mr = s / (float)N;
                    mi = rintf(sb * (1.0f / (float)N));
                    // the second pass reads the window again, a row's loads
                    // together (holding the first pass's values instead
                    // costs registers the sampling loop needs)
#pragma unroll 1
                    for (int ty = 0; ty < TAPS; ++ty) {
                        float gr[TAPS];
#pragma unroll
                        for (int tx = 0; tx < TAPS; ++tx) gr[tx] = tap_grey(ty, tx);
#pragma unroll
                        for (int tx = 0; tx < TAPS; ++tx) second_pass(ty, tx, gr[tx]);
                    }
                    delta = (sb - (float)N * mi) * (1.0f / (float)N);
                    var = var / (float)N;
                    sr = sqrtf(var);
                } else if (inside) {
                    float s = 0.0f, sb = 0.0f;
#pragma unroll 1
                    for (int ty = 0; ty < TAPS; ++ty) {
                        const float *row = job.gray + pidx + (int64_t)((ty - H) * stride) * w;
#pragma unroll
                        for (int tx = 0; tx < TAPS; ++tx) {
                            const float g = row[(tx - H) * stride];
                            s = s + g;
                            sb += rintf(g * 255.0f);
                        }
                    }
                    mr = s / (float)N;
                    const float mi = rintf(sb * (1.0f / (float)N));
                    delta = (sb - (float)N * mi) * (1.0f / (float)N);
                    float var = 0.0f;
#pragma unroll 1
                    for (int ty = 0; ty < TAPS; ++ty) {
                        const float *row = job.gray + pidx + (int64_t)((ty - H) * stride) * w;
#pragma unroll
                        for (int tx = 0; tx < TAPS; ++tx) {
                            const float g = row[(tx - H) * stride];
                            const float d = g - mr;
                            var = var + d * d;
                            if constexpr (F32)
                                Af[ty * TAPS + tx][lane] = d;
                            else
                                Ah[ty * TAPS + tx][lane] = (_Float16)(rintf(g * 255.0f) - mi);
                        }
                    }
                    var = var / (float)N;
                    sr = sqrtf(var);
                }
                const bool pixel_ok = inside && !(sr < job.texture_min);
                const float2 rc = pixel_ok ? job.rays[pidx] : make_float2(0.0f, 0.0f);
                // N sd of the window in the samples' units; for f32 views the
                // window's taps are centred by mr itself (delta = mr)
                const float nsr = F32 ? (float)N * sr : (float)N * (sr * 255.0f);
                if constexpr (F32) delta = mr;

                // Cost of plane (d, n) at the pixel of lane `col` (centre ray
                // (rcx, rcy), N sd of its bytes nsrv, byte-mean correction dlt,
                // reference window in column col of Ah). Every lane calls it;
                // lanes without a hypothesis to score pass act = false (act
                // includes the pixel's own test, pixel_ok). It leaves the two
                // best sources in epair (first | second << 4; 0xff unless
                // act), which the refinement takes as the plane's pair.
                uint32_t epair = 0xffu;
                auto eval = [&](bool act, float rcx, float rcy, float nsrv, float dlt, int col,
                                float d, float n0, float n1, float n2) -> float {
                    if (act) ++n_evals;
                    // dense.rs:187-190: c = n . (ray d)
                    const float xp0 = rcx * d, xp1 = rcy * d;
                    float cc = n0 * xp0 + n1 * xp1;
                    cc = cc + n2 * d;
                    // |n . r| >= 1e-6 and c / (n . r) > 0 at the window corners
                    const float dc = fmaf(n0, rcx, fmaf(n1, rcy, n2));
                    const float ex = n0 * (rstep * (float)H), ey = n1 * (rstep * (float)H);
                    bool plane_ok = true;
#pragma unroll
                    for (int cr = 0; cr < 4; ++cr) {
                        const float den = dc + ((cr & 1) ? ex : -ex) + ((cr & 2) ? ey : -ey);
                        plane_ok = plane_ok && fabsf(den) >= 1e-6f && cc * den > 0.0f;
                    }
                    act = act && plane_ok;
                    // window depth at the centre ray and the source-space steps
                    // per tap column / row scaled by it (see the file comment)
                    const float rden = act ? __builtin_amdgcn_rcpf(dc) : 0.0f;
                    const float dqc = cc * rden;
                    const float pc0 = rcx * dqc, pc1 = rcy * dqc;
                    const float sx = dqc * rstep;
                    const float tnx = (n0 * rstep) * rden, tny = (n1 * rstep) * rden;
                    float m1 = INFINITY, m2 = INFINITY;
                    uint32_t jj = 0xffu;  // their sources, first | second << 4
                    // (PL_F16T: every lane walks the sources, so that the whole
                    // wavefront can stage each source's tile)
                    if (TILE || act) {
                        for (int k = 0; k < K; ++k) {
                            const SrcCam S = scalar_load(job.srcs + k);
                            const float F = S.f, CXh = S.cx - 0.5f, CYh = S.cy - 0.5f;
                            // centre tap in the source camera (dense.rs:222-226)
                            // and the steps dq_c H e_x, dq_c H e_y, focal
                            // folded into the first two rows
                            float B0[3], B1[3], B2[3];
#pragma unroll
                            for (int i = 0; i < 3; ++i) {
                                const float xc = fmaf(S.R[3 * i], pc0,
                                                      fmaf(S.R[3 * i + 1], pc1,
                                                           fmaf(S.R[3 * i + 2], dqc, S.t[i])));
                                const float c1 = fmaf(S.t[i], tnx, S.R[3 * i] * sx);
                                const float c2 = fmaf(S.t[i], tny, S.R[3 * i + 1] * sx);
                                const float sc = i < 2 ? F : 1.0f;
                                B0[i] = sc * xc;
                                B1[i] = sc * c1;
                                B2[i] = sc * c2;
                            }
                            const float wm1 = (float)(S.w - 1), hm1 = (float)(S.h - 1);
                            // gray_at (dense.rs:69-85) succeeds at every tap iff
                            // it does at the corner taps: the third coordinate is
                            // affine (positive at the corners, positive on the
                            // window), and a projective map that keeps it
                            // positive sends the window's rectangle onto the
                            // quadrilateral of the corner images, so the taps'
                            // bounding box is the corners'. NaN fails.
                            bool ok = true, inner = true;
                            float fxlo = INFINITY, fxhi = -INFINITY, fylo = INFINITY,
                                  fyhi = -INFINITY;
#pragma unroll
                            for (int cr = 0; cr < 4; ++cr) {
                                const float fdx = (cr & 1) ? (float)H : -(float)H;
                                const float fdy = (cr & 2) ? (float)H : -(float)H;
                                const float P0 = fmaf(fdx, B1[0], fmaf(fdy, B2[0], B0[0]));
                                const float P1 = fmaf(fdx, B1[1], fmaf(fdy, B2[1], B0[1]));
                                const float P2 = fmaf(fdx, B1[2], fmaf(fdy, B2[2], B0[2]));
                                const float ip = __builtin_amdgcn_rcpf(P2);
                                const float fx = fmaf(P0, ip, CXh), fy = fmaf(P1, ip, CYh);
                                ok = ok && (P2 > 0.0f) && fx >= 0.0f && fy >= 0.0f && fx < wm1 &&
                                     fy < hm1;
                                inner = inner && fx >= kInnerMargin && fy >= kInnerMargin &&
                                        fx < wm1 - kInnerMargin && fy < hm1 - kInnerMargin;
                                if constexpr (TILE) {
                                    fxlo = fminf(fxlo, fx);
                                    fxhi = fmaxf(fxhi, fx);
                                    fylo = fminf(fylo, fy);
                                    fyhi = fmaxf(fyhi, fy);
                                }
                            }
                            if constexpr (TILE) ok = ok && act;
                            const uint32_t sw = (uint32_t)S.w;
                            // the plane as a structured buffer: 32-bit pixel
                            // indices, and the hardware range check as a
                            // second bound; the resource is wave-uniform
                            // (scalar registers)
                            // (readfirstlane returns a signed int: each half
                            // goes through uint32_t, never sign-extended)
                            constexpr uint32_t QB =
                                PLANE == PL_F32 ? 16u : (F16Q ? 8u : 4u);
                            const uint64_t qa =
                                PLANE == PL_F32    ? (uint64_t)(uintptr_t)S.quadd
                                : F16Q             ? (uint64_t)(uintptr_t)S.quadh
                                : PLANE == PL_F32G ? (uint64_t)(uintptr_t)S.gray
                                                   : (uint64_t)(uintptr_t)S.quad;
                            const uint64_t qlo =
                                (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)qa);
                            const uint64_t qhi =
                                (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(qa >> 32));
                            // records of QB bytes, one per pixel (PL_F32G:
                            // a raw buffer of the grey plane and its zero
                            // padding, in bytes)
                            const __amdgpu_buffer_rsrc_t Qh = __builtin_amdgcn_make_buffer_rsrc(
                                (void *)(uintptr_t)(qlo | (qhi << 32)),
                                (short)(PLANE == PL_F32G ? 0u : QB),
                                __builtin_amdgcn_readfirstlane(
                                    PLANE == PL_F32G
                                        ? (int32_t)(4 * (S.w * S.h + (int32_t)grey_pad(S.w)))
                                        : S.w * S.h),
                                kBufferWord3);
                            // bytes per grey row (PL_F32G)
                            const int32_t row4 = __builtin_amdgcn_readfirstlane(4 * S.w);
                            // PL_F16T: the batch's texels of this source in LDS when
                            // every scoring window lies inside it (1/256 px, so
                            // rounding keeps each tap's texel within the corners'
                            // box) and the box fits the tile (uniform decision)
                            bool tiled = false;
                            uint32_t tw = 1u;
                            int32_t tbase = 0, tcnt = 1;
                            if constexpr (TILE) {
                                typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
                                const float m = kInnerMargin;
                                // (x0, y0) low corner and (0xffff - x1, 0xffff - y1):
                                // both reduce by packed 16-bit minimum
                                u16x2 lo = {0xffff, 0xffff}, hi = {0xffff, 0xffff};
                                if (ok) {
                                    lo = u16x2{(uint16_t)(int)(fxlo - m), (uint16_t)(int)(fylo - m)};
                                    hi = u16x2{(uint16_t)(0xffff - (int)(fxhi + m)),
                                               (uint16_t)(0xffff - (int)(fyhi + m))};
                                }
#pragma unroll
                                for (int off = 32; off >= 1; off >>= 1) {
                                    const u16x2 ol = __builtin_bit_cast(
                                        u16x2, __shfl_xor(__builtin_bit_cast(uint32_t, lo), off));
                                    const u16x2 oh = __builtin_bit_cast(
                                        u16x2, __shfl_xor(__builtin_bit_cast(uint32_t, hi), off));
                                    lo = __builtin_elementwise_min(lo, ol);
                                    hi = __builtin_elementwise_min(hi, oh);
                                }
                                const uint32_t lov = __builtin_amdgcn_readfirstlane(
                                    __builtin_bit_cast(uint32_t, lo));
                                const uint32_t hiv = __builtin_amdgcn_readfirstlane(
                                    __builtin_bit_cast(uint32_t, hi));
                                const int32_t X0 = (int32_t)(lov & 0xffffu), Y0 = (int32_t)(lov >> 16);
                                const int32_t X1 = 0xffff - (int32_t)(hiv & 0xffffu),
                                              Y1 = 0xffff - (int32_t)(hiv >> 16);
                                const bool any = __ballot(ok) != 0ull;
                                const bool allin = __ballot(ok && !inner) == 0ull;
                                const int32_t TW = X1 - X0 + 1, TH = Y1 - Y0 + 1;
                                tiled = any && allin && TW > 0 && TH > 0 && TW <= kTileCap &&
                                        TH <= kTileCap && TW * TH <= kTileCap;
                                if (tiled) {
                                    tw = (uint32_t)TW;
                                    tcnt = TW * TH;
                                    tbase = Y0 * TW + X0;
                                    const float rtw = 1.0f / (float)TW;
                                    for (int i = lane; i < tcnt; i += 64) {
                                        const int r = (int)(((float)i + 0.5f) * rtw);
                                        const int cl = i - r * TW;
                                        tile[i] = amvs_struct_load_b64(
                                            Qh, (int32_t)((uint32_t)(Y0 + r) * sw) + X0 + cl, 0, 0, 0);
                                    }
                                    __syncthreads();
                                }
                            }
                            float cs = 2.0f;  // a failing source scores 2.0
                            if (ok) {
                                float c0 = 0.0f, s1 = 0.0f, s2 = 0.0f, sa = 0.0f;
                                // The window's taps; CLAMP keeps every position
                                // in the sampling range whatever the rounding
                                // (v_med3, NaN to 0), needed only when some lane
                                // scoring this source has a corner near the edge.
                                auto taps = [&](auto clamp_tag, auto tile_tag) {
                                    constexpr bool CLAMP = decltype(clamp_tag)::value;
                                    constexpr bool TL = decltype(tile_tag)::value;
                                    // one tap row per trip: the row's positions,
                                    // then all its gathers, then its samples and
                                    // sums, so that a row's gathers are in flight
                                    // together (a scheduling barrier keeps the
                                    // compiler from sinking each gather to its
                                    // use, which it otherwise does in some
                                    // instantiations); the registers they hold
                                    // leave room for four wavefronts per SIMD
                                    constexpr int QW = PLANE == PL_F32 ? 4 : (PLANE == PL_U8 ? 1 : 2);
                                    constexpr int QN = PLANE == PL_F32G ? 2 : 1;  // loads per tap
#pragma unroll 1
                                    for (int ty = 0; ty < TAPS; ++ty) {
                                        const float fdy = (float)(ty - H);
                                        const float Y0 = fmaf(fdy, B2[0], B0[0]);
                                        const float Y1 = fmaf(fdy, B2[1], B0[1]);
                                        const float Y2 = fmaf(fdy, B2[2], B0[2]);
                                        float axs[TAPS], ays[TAPS];
                                        uint32_t q[TAPS][QN][QW];
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
                                            axs[tx] = __builtin_amdgcn_fractf(fx);
                                            ays[tx] = __builtin_amdgcn_fractf(fy);
                                            if constexpr (TL) {
                                                // the staged texel (index clamped into
                                                // the tile: a safety bound only)
                                                const uint32_t ti = __builtin_elementwise_min(
                                                    (uint32_t)((int32_t)(__umul24(y0, tw) + x0) - tbase),
                                                    (uint32_t)(tcnt - 1));
                                                const u32x2 qv = tile[ti];
                                                q[tx][0][0] = qv.x;
                                                q[tx][0][QW - 1] = qv.y;
                                            } else if constexpr (PLANE == PL_F32G) {
                                                // rows y0 and y0 + 1: g00, g10 and
                                                // g01, g11 (at the last column or row
                                                // the pair's second value, weight 0,
                                                // is the next row's first or the
                                                // zero padding)
                                                const int32_t off = (int32_t)(
                                                    __umul24(y0, (uint32_t)row4) + (x0 << 2));
                                                const amvs_u32x2 r0 = amvs_raw_load_b64(Qh, off, 0, 0);
                                                const amvs_u32x2 r1 = amvs_raw_load_b64(Qh, off, row4, 0);
                                                q[tx][0][0] = r0.x;
                                                q[tx][0][QW - 1] = r0.y;
                                                q[tx][QN - 1][0] = r1.x;
                                                q[tx][QN - 1][QW - 1] = r1.y;
                                            } else if constexpr (PLANE == PL_F32) {
                                                const int32_t idx = (int32_t)(__umul24(y0, sw) + x0);
                                                const amvs_f32x4 e = amvs_struct_load_f4(Qh, idx, 0, 0, 0);
                                                q[tx][0][0] = __float_as_uint(e.x);
                                                q[tx][0][1 % QW] = __float_as_uint(e.y);
                                                q[tx][0][2 % QW] = __float_as_uint(e.z);
                                                q[tx][0][3 % QW] = __float_as_uint(e.w);
                                            } else if constexpr (F16Q) {
                                                const int32_t idx = (int32_t)(__umul24(y0, sw) + x0);
                                                const u32x2 qv = amvs_struct_load_b64(Qh, idx, 0, 0, 0);
                                                q[tx][0][0] = qv.x;
                                                q[tx][0][QW - 1] = qv.y;
                                            } else {
                                                const int32_t idx = (int32_t)(__umul24(y0, sw) + x0);
                                                q[tx][0][0] = amvs_struct_load_b32(Qh, idx, 0, 0, 0);
                                            }
                                        }
                                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                                        for (int tx = 0; tx < TAPS; ++tx) {
                                            const float ax = axs[tx], ay = ays[tx];
                                            // the sample g00 + ax A + ay (B + ax C) as
                                            // t1 + ay u, t1 = g00 + ax A, u = B + ax C
                                            float t1, u;
                                            if constexpr (TL || F16Q) {
                                                // {g00, A} and {B, C}, f16
                                                const h2 r0 = __builtin_bit_cast(h2, q[tx][0][0]);
                                                const h2 r1 = __builtin_bit_cast(h2, q[tx][0][QW - 1]);
                                                t1 = fmaf(ax, (float)r0.y, (float)r0.x);
                                                u = fmaf(ax, (float)r1.y, (float)r1.x);
                                            } else if constexpr (PLANE == PL_F32G) {
                                                // the coefficients as the f32 plane
                                                // forms them
                                                const float g00 = __uint_as_float(q[tx][0][0]);
                                                const float g10 = __uint_as_float(q[tx][0][QW - 1]);
                                                const float g01 = __uint_as_float(q[tx][QN - 1][0]);
                                                const float g11 = __uint_as_float(q[tx][QN - 1][QW - 1]);
                                                const float cA = g10 - g00;
                                                t1 = fmaf(ax, cA, g00);
                                                u = fmaf(ax, (g11 - g01) - cA, g01 - g00);
                                            } else if constexpr (PLANE == PL_F32) {
                                                // g00, A, B, C
                                                t1 = fmaf(ax, __uint_as_float(q[tx][0][1 % QW]),
                                                          __uint_as_float(q[tx][0][0]));
                                                u = fmaf(ax, __uint_as_float(q[tx][0][3 % QW]),
                                                         __uint_as_float(q[tx][0][2 % QW]));
                                            } else {
                                                // the four corners as bytes; the
                                                // coefficients are exact integers, as
                                                // in the f16 plane
                                                const uint32_t qd = q[tx][0][0];
                                                const float g00 = (float)(qd & 0xffu);
                                                const float g10 = (float)((qd >> 8) & 0xffu);
                                                const float g01 = (float)((qd >> 16) & 0xffu);
                                                const float g11 = (float)(qd >> 24);
                                                const float cA = g10 - g00;
                                                t1 = fmaf(ax, cA, g00);
                                                u = fmaf(ax, (g11 - g01) - cA, g01 - g00);
                                            }
                                            const float val = fmaf(ay, u, t1);
                                            if (tx == 0 && ty == 0) c0 = val;
                                            const float dd = val - c0;
                                            s1 += dd;
                                            s2 = fmaf(dd, dd, s2);
                                            if constexpr (F32)
                                                sa = fmaf(Af[ty * TAPS + tx][col], dd, sa);
                                            else
                                                sa = fmaf((float)Ah[ty * TAPS + tx][col], dd, sa);
                                        }
                                    }
                                };
                                if constexpr (TILE) {
                                    if (tiled)
                                        taps(std::false_type{}, std::true_type{});
                                    else if (__ballot(!inner) == 0ull)
                                        taps(std::false_type{}, std::false_type{});
                                    else
                                        taps(std::true_type{}, std::false_type{});
                                } else if (__ballot(!inner) == 0ull) {  // uniform over the scoring lanes
                                    taps(std::false_type{}, std::false_type{});
                                } else {
                                    taps(std::true_type{}, std::false_type{});
                                }
                                // NCC from the shifted sums, in byte units:
                                // N var = s2 - s1^2 / N and
                                // cov = sum (b - mean b) dd = sa - delta s1
                                // (f32 views: grey units, and sa already sums
                                // the centred window, cov = sa)
                                const float moff = s1 * (1.0f / (float)N);
                                const float var =
                                    fmaxf(fmaf(-s1, moff, s2), 0.0f) * (1.0f / (float)N);
                                const float cov = F32 ? sa : fmaf(-dlt, s1, sa);
                                const float ss = __builtin_amdgcn_sqrtf(var);
                                // dense.rs:246-248's 1e-3 grey units are 0.255
                                // byte units
                                if (!(ss < (F32 ? 1e-3f : 1e-3f * 255.0f)))
                                    cs = 1.0f - cov * __builtin_amdgcn_rcpf(nsrv * ss);
                            }
                            // best two (dense.rs:255-262), and their sources
                            // (ties to the lower index, as the sort's)
                            const bool b1 = cs < m1, b2 = !b1 && cs < m2;
                            jj = b1   ? ((jj << 4) & 0xf0u) | (uint32_t)k
                                 : b2 ? (jj & 0x0fu) | ((uint32_t)k << 4)
                                      : jj;
                            const float lo = fminf(m1, cs);
                            m2 = fminf(fmaxf(m1, cs), m2);
                            m1 = lo;
                        }
                    }
                    float cst = K >= 2 ? (m1 + m2) * 0.5f : m1;
                    epair = act && K >= 2 ? jj : 0xffu;
                    if (!act) cst = 2.0f;
                    return cst;
                };

                if (MODE == MODE_COST_LIST) {
                    const float4 hp = have ? a.hyp_in[item] : make_float4(1.0f, 0.0f, 0.0f, -1.0f);
                    const float cst = eval(have && pixel_ok, rc.x, rc.y, nsr, delta, lane,
                                           hp.x, hp.y, hp.z, hp.w);
                    if (have) a.out[item] = cst;
                } else if (MODE == MODE_EVAL_CURRENT) {
                    const float4 hp = have ? job.hyp[pidx] : make_float4(1.0f, 0.0f, 0.0f, -1.0f);
                    const float cst = eval(have && pixel_ok, rc.x, rc.y, nsr, delta, lane,
                                           hp.x, hp.y, hp.z, hp.w);
                    if (have) {
                        job.cost[pidx] = cst;
                        job.pair[pidx] = (uint8_t)epair;
                    }
                } else if constexpr (PLANE == PL_U8) {
                    // the ten proposals of dense.rs:427-462 in the reference's
                    // order (propagation in CAND order, then two random
                    // refinements), every lane in step. The 8-bit plane serves
                    // the scattered first iterations, whose candidates are all
                    // distinct (nothing to compact), and a row of gathers in
                    // flight matters most there: this loop keeps the code the
                    // compiler schedules that way.
                    float4 best = have ? job.hyp[pidx] : make_float4(1.0f, 0.0f, 0.0f, -1.0f);
                    float bcost = have ? job.cost[pidx] : 2.0f;
                    // a candidate is taken only at a strictly lower cost, so
                    // the hypothesis changed exactly when one was taken
                    bool taken = false;
                    uint32_t bpair = 0xffu;
                    bool pair_set = false;
                    if (a.eval_own) {
                        // the pixel's own plane has no cost yet: scored first
                        const float c0 = eval(have && pixel_ok, rc.x, rc.y, nsr, delta, lane,
                                              best.x, best.y, best.z, best.w);
                        if (have && pixel_ok) {
                            bcost = c0;
                            bpair = epair;
                            pair_set = true;
                        }
                    }
                    uint64_t s = seed_refine((uint64_t)pidx, (uint64_t)a.it, (uint64_t)a.parity);
                    const float span = job.imax - job.imin;
#pragma unroll 1
                    for (int k = 0; k < 10; ++k) {
                        float4 hp;
                        bool act;
                        if (k < 8) {
                            int nx, ny;
                            cand_pos(k, x, y, nx, ny);
                            const bool inb = have && nx >= 0 && ny >= 0 && nx < w && ny < h;
                            hp = make_float4(1.0f, 0.0f, 0.0f, -1.0f);
                            if (inb) hp = job.hyp[(int64_t)ny * w + nx];
                            act = inb && !same_hyp(hp, best);
                            // scored and not taken two half-sweeps before
                            const bool skip =
                                act && a.skip_unchanged && !job.changed[(int64_t)ny * w + nx];
                            n_skipped += (uint32_t)__popcll(__ballot(skip));
                            act = act && !skip;
                        } else {
                            float stp = (rnd_next(s) - 0.5f) * a.range;
                            stp = stp * 0.25f;
                            stp = stp * span;
                            const float id = 1.0f / best.x + stp;
                            hp.x = (id > job.imin * 0.5f && id < job.imax * 1.5f) ? 1.0f / id
                                                                                  : best.x;
                            random_normal(s, rc.x, rc.y, best.y, best.z, best.w, 0.25f * a.range,
                                          hp.y, hp.z, hp.w);
                            act = have;
                        }
                        const float cst = eval(act && pixel_ok, rc.x, rc.y, nsr, delta, lane,
                                               hp.x, hp.y, hp.z, hp.w);
                        if (act && cst < bcost) {
                            best = hp;
                            bcost = cst;
                            taken = true;
                            bpair = epair;
                        }
                    }
                    if (have) {
                        job.hyp[pidx] = best;
                        job.cost[pidx] = bcost;
                        job.changed[pidx] = taken;
                        if (taken || pair_set) job.pair[pidx] = (uint8_t)bpair;
                    }
                } else {
                    // the ten proposals of dense.rs:427-462: the eight
                    // propagation candidates in CAND order, then two random
                    // refinements
                    float4 best = have ? job.hyp[pidx] : make_float4(1.0f, 0.0f, 0.0f, -1.0f);
                    float bcost = have ? job.cost[pidx] : 2.0f;
                    bool taken = false;  // (at a strictly lower cost: a change)
                    uint32_t bpair = 0xffu;
                    // eval_own: the pixel's own plane has no cost yet and is
                    // scored first, as list entry 8 ordered ahead of the
                    // candidates (its cost wins ties, so a candidate is taken
                    // only at a strictly lower cost); its pair is then known
                    const bool own = a.eval_own && pixel_ok;
                    bool pair_set = false;
                    // A candidate equal to the current best is not scored
                    // (dense.rs:436-438); one equal to an earlier candidate
                    // cannot be accepted either (its cost is that candidate's,
                    // which was not lower, or it is the best by now), so only
                    // the distinct ones are scored. Accepting each strictly
                    // lower cost in order ends on the lowest (cost, index), so
                    // the wavefront scores the candidates of all its pixels as
                    // one compacted list, no lane idle while others score,
                    // and keeps each pixel's lowest (cost, index) in LDS.
                    uint32_t amask = 0;
                    uint64_t bins = 0;  // kBinBits per candidate
                    const bool by_depth = a.it < kSortIters;  // uniform
                    // A candidate from a neighbour that its last half-sweep
                    // left unchanged was scored at this pixel two half-sweeps
                    // before and not taken, or taken (and is the best or was
                    // beaten since): its cost is what it was, and the best
                    // cost can only have fallen, so it cannot be taken now.
                    // (The neighbours' flags and planes are loaded before the
                    // candidate list, at in-bounds positions — the pixel's
                    // own where the neighbour is outside — so that their
                    // loads issue together; their values are used only
                    // where inside: 23 ms less per 77-view run. Read among
                    // the candidates' loads, the flags made the compiler
                    // serialise the taps' gathers.)
                    int64_t nidx[8];
                    uint32_t inbm0 = 0;
                    uint32_t unchanged = 0;
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        int nx, ny;
                        cand_pos(k, x, y, nx, ny);
                        const bool inb = pixel_ok && nx >= 0 && ny >= 0 && nx < w && ny < h;
                        nidx[k] = inb ? (int64_t)ny * w + nx : pidx;
                        inbm0 |= (inb ? 1u : 0u) << k;
                    }
                    if (a.skip_unchanged && pixel_ok) {
                        uint8_t fl[8];
#pragma unroll
                        for (int k = 0; k < 8; ++k) fl[k] = as_global(job.changed)[nidx[k]];
#pragma unroll
                        for (int k = 0; k < 8; ++k)
                            if (((inbm0 >> k) & 1u) && !fl[k]) unchanged |= 1u << k;
                    }
                    const float bscale = (float)kDepthBins / (job.imax - job.imin);
                    {
                        float4 cand[8];
                        uint32_t inbm = 0;
#pragma unroll
                        for (int k = 0; k < 8; ++k) {
                            const f32x4 c = as_global(job.hyp)[nidx[k]];
                            cand[k] = make_float4(c.x, c.y, c.z, c.w);
                        }
#pragma unroll
                        for (int k = 0; k < 8; ++k) {
                            const bool inb = (inbm0 >> k) & 1u;
                            if (!inb) cand[k] = make_float4(1.0f, 0.0f, 0.0f, -1.0f);
                            bool act = inb && !same_hyp(cand[k], best);
#pragma unroll
                            for (int j = 0; j < k; ++j)
                                act = act && !(((inbm >> j) & 1u) && same_hyp(cand[j], cand[k]));
                            const bool skip = act && ((unchanged >> k) & 1u);
                            n_skipped += (uint32_t)__popcll(__ballot(skip));
                            act = act && !skip;
                            inbm |= (inb ? 1u : 0u) << k;
                            amask |= (act ? 1u : 0u) << k;
                            if (by_depth) {
                                const float iv = __builtin_amdgcn_rcpf(cand[k].x);
                                const int b = (int)__builtin_amdgcn_fmed3f(
                                    (iv - job.imin) * bscale, 0.0f, (float)(kDepthBins - 1));
                                bins |= (uint64_t)b << (kBinBits * k);
                            }
                        }
                    }
                    int32_t total = 0;
                    if (by_depth) {
                        // the list by depth bin: counts, their scan, then
                        // each candidate at its bin's next slot (the order
                        // within a bin does not matter: the minimum is)
                        for (int i = lane; i < kDepthBins; i += 64) pbin[i] = 0;
                        __syncthreads();
                        uint32_t bin_own = 0u;
                        if (own)
                            bin_own = (uint32_t)__builtin_amdgcn_fmed3f(
                                (__builtin_amdgcn_rcpf(best.x) - job.imin) * bscale, 0.0f,
                                (float)(kDepthBins - 1));
                        auto bin_of = [&](int k) {
                            return (uint32_t)(bins >> (kBinBits * k)) & (uint32_t)(kDepthBins - 1);
                        };
#pragma unroll
                        for (int k = 0; k < 8; ++k)
                            if ((amask >> k) & 1u) atomicAdd(&pbin[bin_of(k)], 1);
                        if (own) atomicAdd(&pbin[bin_own], 1);
                        __syncthreads();
                        // each lane holds kBinsPerLane consecutive bins
                        int loc[kBinsPerLane], cnt = 0;
#pragma unroll
                        for (int j = 0; j < kBinsPerLane; ++j) {
                            const int bi = lane * kBinsPerLane + j;
                            loc[j] = bi < kDepthBins ? pbin[bi] : 0;
                            cnt += loc[j];
                        }
                        int incl = cnt;
#pragma unroll
                        for (int off = 1; off < 64; off <<= 1) {
                            const int t = __shfl_up(incl, off);
                            if (lane >= off) incl += t;
                        }
                        total = __shfl(incl, 63);
                        __syncthreads();
                        int run = incl - cnt;
#pragma unroll
                        for (int j = 0; j < kBinsPerLane; ++j) {
                            const int bi = lane * kBinsPerLane + j;
                            if (bi < kDepthBins) pbin[bi] = run;
                            run += loc[j];
                        }
                        __syncthreads();
#pragma unroll
                        for (int k = 0; k < 8; ++k)
                            if ((amask >> k) & 1u)
                                plist[atomicAdd(&pbin[bin_of(k)], 1)] =
                                    (uint16_t)(lane | (k << 6));
                        if (own) plist[atomicAdd(&pbin[bin_own], 1)] = (uint16_t)(lane | (8 << 6));
                    } else {
                        // later iterations: the list pixel-major, each pixel's
                        // candidates (nearly one plane once the first
                        // iterations are done) in neighbouring lanes, own
                        // plane first, the pixels in walk order (77-view
                        // half-sweeps: 1711 ms against 1745 with the pixels'
                        // k-th candidates together)
                        const int cnt = __popc(amask) + (own ? 1 : 0);
                        int incl = cnt;
#pragma unroll
                        for (int off = 1; off < 64; off <<= 1) {
                            const int t = __shfl_up(incl, off);
                            if (lane >= off) incl += t;
                        }
                        total = __shfl(incl, 63);
                        int at = incl - cnt;
                        if (own) plist[at++] = (uint16_t)(lane | (8 << 6));
#pragma unroll
                        for (int k = 0; k < 8; ++k)
                            if ((amask >> k) & 1u) plist[at++] = (uint16_t)(lane | (k << 6));
                    }
                    pmin[lane] = ~0ull;
                    __syncthreads();  // the list, pmin and every lane's window in Ah
#pragma unroll 1
                    for (int e0 = 0; e0 < total; e0 += 64) {
                        const int e = e0 + lane;
                        const bool act = e < total;
                        const int ent = act ? (int)plist[e] : 0;
                        const int o = ent & 63, k = ent >> 6;
                        // the owner's pixel
                        const int ox = __shfl(x, o), oy = __shfl(y, o);
                        const float orx = __shfl(rc.x, o), ory = __shfl(rc.y, o);
                        const float onsr = __shfl(nsr, o), odl = __shfl(delta, o);
                        int nx = ox, ny = oy;  // (entry 8: the pixel's own plane)
                        if (k < 8) cand_pos(k, ox, oy, nx, ny);
                        const float4 hp = act ? job.hyp[(int64_t)ny * w + nx]
                                              : make_float4(1.0f, 0.0f, 0.0f, -1.0f);
                        const float cst = eval(act, orx, ory, onsr, odl, o, hp.x,
                                               hp.y, hp.z, hp.w);
                        // (cost, order) orders the keys, the own plane first
                        // and then the candidates in CAND order; the pair
                        // rides below them
                        const uint32_t ord = k == 8 ? 0u : (uint32_t)k + 1u;
                        if (act)
                            atomicMin(&pmin[o], ((unsigned long long)cost_key(cst) << 32) |
                                                    ((unsigned long long)ord << 8) |
                                                    (unsigned long long)epair);
                    }
                    __syncthreads();
                    const unsigned long long pm = pmin[lane];
                    if (pm != ~0ull) {
                        const float cst = key_cost((uint32_t)(pm >> 32));
                        const int ord = (int)((pm >> 8) & 15ull);
                        if (ord == 0) {
                            // the own plane scored lowest: its cost, no change
                            bcost = cst;
                            bpair = (uint32_t)(pm & 0xffull);
                            pair_set = true;
                        } else if (cst < bcost) {
                            int nx, ny;
                            cand_pos(ord - 1, x, y, nx, ny);
                            best = job.hyp[(int64_t)ny * w + nx];
                            bcost = cst;
                            taken = true;
                            bpair = (uint32_t)(pm & 0xffull);
                        }
                    }
                    // the two random refinements from the best after
                    // propagation (dense.rs:443-462)
                    uint64_t s = seed_refine((uint64_t)pidx, (uint64_t)a.it, (uint64_t)a.parity);
                    const float span = job.imax - job.imin;
#pragma unroll 1
                    for (int k = 8; k < 10; ++k) {
                        float4 hp;
                        float stp = (rnd_next(s) - 0.5f) * a.range;
                        stp = stp * 0.25f;
                        stp = stp * span;
                        const float id = 1.0f / best.x + stp;
                        hp.x = (id > job.imin * 0.5f && id < job.imax * 1.5f) ? 1.0f / id : best.x;
                        random_normal(s, rc.x, rc.y, best.y, best.z, best.w, 0.25f * a.range, hp.y,
                                      hp.z, hp.w);
                        const bool act = have && pixel_ok;
                        const float cst = eval(act, rc.x, rc.y, nsr, delta, lane, 
                                               hp.x, hp.y, hp.z, hp.w);
                        if (act && cst < bcost) {
                            best = hp;
                            bcost = cst;
                            taken = true;
                            bpair = epair;
                        }
                    }
                    if (have) {
                        job.hyp[pidx] = best;
                        job.cost[pidx] = bcost;
                        job.changed[pidx] = taken;
                        if (taken || pair_set) job.pair[pidx] = (uint8_t)bpair;
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
    if (MODE == MODE_SWEEP && a.skipped && lane == 0 && n_skipped)
        atomicAdd(a.skipped, (unsigned long long)n_skipped);
}

template <int TAPS, int MODE, bool PERSIST, int PLANE>
void launch_pml_class(const PmArgs &a, int32_t grid, hipStream_t st) {
    hipLaunchKernelGGL(k_pm_prep, dim3(1), dim3(64), 0, st, a, (int32_t)MODE, 64 * a.gpc);
    hipLaunchKernelGGL((k_pml<TAPS, MODE, PERSIST, PLANE>), dim3((unsigned)grid), dim3(64), 0, st,
                       a);
}

template <int TAPS, int PLANE>
void pml_taps(int mode, bool persist, const PmArgs &a, int32_t grid, hipStream_t st) {
    if (mode == MODE_EVAL_CURRENT)
        launch_pml_class<TAPS, MODE_EVAL_CURRENT, false, PLANE>(a, grid, st);
    else if (mode == MODE_SWEEP && persist)
        launch_pml_class<TAPS, MODE_SWEEP, true, PLANE>(a, grid, st);
    else if (mode == MODE_SWEEP)
        launch_pml_class<TAPS, MODE_SWEEP, false, PLANE>(a, grid, st);
    else
        launch_pml_class<TAPS, MODE_COST_LIST, false, PLANE>(a, grid, st);
}

template <int TAPS>
void pml_planes(int mode, int plane, bool persist, const PmArgs &a, int32_t grid,
                hipStream_t st) {
    if (plane == PL_F16)
        pml_taps<TAPS, PL_F16>(mode, persist, a, grid, st);
    else if (plane == PL_U8)
        pml_taps<TAPS, PL_U8>(mode, persist, a, grid, st);
    else if (plane == PL_F32)
        pml_taps<TAPS, PL_F32>(mode, persist, a, grid, st);
    else if (plane == PL_F32G)
        pml_taps<TAPS, PL_F32G>(mode, persist, a, grid, st);
    else
        pml_taps<TAPS, PL_F16T>(mode, persist, a, grid, st);
}

}  // namespace

int launch_pml(int mode, int32_t taps, int plane, bool persist, const PmArgs &a, int32_t grid,
               hipStream_t st) {
    if (plane < PL_U8 || plane > PL_F16T) return -1;
    if (taps == 5)
        pml_planes<5>(mode, plane, persist, a, grid, st);
    else if (taps == 7)
        pml_planes<7>(mode, plane, persist, a, grid, st);
    else if (taps == 11)
        pml_planes<11>(mode, plane, persist, a, grid, st);
    else
        return -1;
    return 0;
}

}  // namespace amvs
