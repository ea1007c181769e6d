// Device helpers shared by the gfx950 kernels: the reference's scalar pieces
// (RNG, seeds, random normals), the exact division and projection helpers and
// small utilities. Included by mvs_kernels.hip and the PatchMatch units.
#pragma once

#include <math.h>
#include <stdint.h>

#include "mvs_internal.h"

namespace amvs {
namespace dev {

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
__device__ __forceinline__ float div1(const float a, const float b, const float r) {
    float q = a * r;
    float res = fmaf(-b, q, a);
    q = fmaf(res, r, q);
    res = fmaf(-b, q, a);
    return fmaf(res, r, q);
}
// refined reciprocal of b (the shared first steps)
__device__ __forceinline__ float recip_refined(const float b) {
    float r = __builtin_amdgcn_rcpf(b);
    const float e = fmaf(-b, r, 1.0f);
    return fmaf(e, r, r);
}
__device__ __forceinline__ void div2_shared(const float a0, const float a1, const float b,
                                            float &q0, float &q1) {
    const float r = recip_refined(b);
    q0 = div1(a0, b, r);
    q1 = div1(a1, b, r);
}

// Global-address-space view of a device pointer (global_load rather than
// flat_load).
typedef const __attribute__((address_space(1))) float gfloat;
__device__ __forceinline__ gfloat *as_global(const float *p) { return (gfloat *)p; }
typedef const __attribute__((address_space(1))) uint32_t gu32;
__device__ __forceinline__ gu32 *as_global(const uint32_t *p) { return (gu32 *)p; }
typedef const __attribute__((address_space(1))) uint8_t gu8;
__device__ __forceinline__ gu8 *as_global(const uint8_t *p) { return (gu8 *)p; }
// A copy of *p read through the constant address space, for tables that no
// lane writes while the kernel runs: at a wave-uniform address the reads
// become scalar loads shared by the wave.
template <class T>
__device__ __forceinline__ T scalar_load(const T *p) {
    typedef const __attribute__((address_space(4))) uint32_t cu32;
    static_assert(sizeof(T) % 4 == 0, "word-sized tables only");
    T v;
    uint32_t *d = reinterpret_cast<uint32_t *>(&v);
    const cu32 *s = (const cu32 *)p;
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) d[i] = s[i];
    return v;
}
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) f32x4 gf32x4;
__device__ __forceinline__ gf32x4 *as_global(const float4 *p) { return (gf32x4 *)p; }

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
// refinement geometry (no reference counterpart)
// ---------------------------------------------------------------------------

// Tangent basis of a unit normal n: two unit directions orthogonal to n and
// to each other, the axes of the refinement's two normal angles.
__device__ __forceinline__ void tangent_dirs(const float n[3], float dirs[2][3]) {
    float e1[3] = {1.0f - n[0] * n[0], -n[0] * n[1], -n[0] * n[2]};
    float l1 = sqrtf(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    if (!(l1 > 0.5f)) {
        e1[0] = -n[1] * n[0];
        e1[1] = 1.0f - n[1] * n[1];
        e1[2] = -n[1] * n[2];
        l1 = sqrtf(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) e1[c] /= l1;
    dirs[0][0] = e1[0];
    dirs[0][1] = e1[1];
    dirs[0][2] = e1[2];
    dirs[1][0] = n[1] * e1[2] - n[2] * e1[1];
    dirs[1][1] = n[2] * e1[0] - n[0] * e1[2];
    dirs[1][2] = n[0] * e1[1] - n[1] * e1[0];
}

// The refinement's Gauss-Newton terms in (d, n0, n1, n2). H (packed upper:
// 00 01 02 11 12 22) and g are J^T J and J^T r of the pair's stacked
// residuals w.r.t. theta = (depth, angle along dirs[0], angle along dirs[1]);
// sum_s (1 - NCC_s) = |r|^2 / 2 is twice the pair cost. With B = [e_d, (0,
// dirs[0]), (0, dirs[1])] (orthonormal columns) and the residuals invariant
// under scaling n, J_(d,n) = J_theta B^T, so the cost gradient is B g / 2 and
// J^T J is B H B^T (out: packed upper of the 4 x 4, row-major order 00 01 02
// 03 11 12 13 22 23 33).
__device__ __forceinline__ void terms_to_dn(const float H[6], const float g[3],
                                            const float dirs[2][3], float grad[4],
                                            float jtj[10]) {
    const float M[3][3] = {{H[0], H[1], H[2]}, {H[1], H[3], H[4]}, {H[2], H[4], H[5]}};
    float B[4][3];  // rows: d, n0, n1, n2; columns: theta
    B[0][0] = 1.0f;
    B[0][1] = 0.0f;
    B[0][2] = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        B[1 + c][0] = 0.0f;
        B[1 + c][1] = dirs[0][c];
        B[1 + c][2] = dirs[1][c];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) grad[a] = 0.5f * (B[a][0] * g[0] + B[a][1] * g[1] + B[a][2] * g[2]);
    int m = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = a; b < 4; ++b, ++m) {
            float v = 0.0f;
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int q = 0; q < 3; ++q) v += B[a][p] * M[p][q] * B[b][q];
            jtj[m] = v;
        }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float unit24(uint64_t h) { return (float)(uint32_t)(h >> 40) * 0x1p-24f; }

}  // namespace dev
}  // namespace amvs
