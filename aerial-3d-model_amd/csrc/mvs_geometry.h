// Per-pixel geometry shared by the consistency and fusion kernels (f64, as
// the reference's fusion, crates/stitch-core/src/recon/dense.rs:586-612).
#pragma once

#include "mvs_internal.h"

namespace amvs {

// Agreement of pixel i of a reference depth map with neighbour maps: -1 where
// depth <= 0, else the number of neighbours that agree. rule 0 (dense.rs:
// 598-612): relative depth difference < 1% at the nearest neighbour pixel;
// rule 1: the neighbour's own point at that pixel, back-projected from the
// pixel centre, reprojects within 1 px of the reference pixel centre. xw
// receives the pixel's world point (dense.rs:594-596).
__device__ inline int32_t agree_pixel(const RefCam64 &rf, const float *__restrict__ depth,
                                      int32_t w, const NbrCam *__restrict__ nbrs, int32_t nn,
                                      int32_t rule, int64_t i, double xw[3]) {
    const float di = depth[i];
    if (di <= 0.0f) return -1;
    const int64_t y = i / w, x = i - y * w;
    const double qx = ((double)x + 0.5 - rf.cx) / rf.f;
    const double qy = ((double)y + 0.5 - rf.cy) / rf.f;
    // (undistorted camera: every fixed-point step divides by exactly 1.0 and
    // returns q itself, so the steps are skipped without changing a bit)
    double px = qx, py = qy;
    if (rf.k1 != 0.0 || rf.k2 != 0.0) {
        for (int k = 0; k < 8; ++k) {
            const double r2 = px * px + py * py;
            double d = 1.0 + rf.k1 * r2;
            d = d + (rf.k2 * r2) * r2;
            px = qx / d;
            py = qy / d;
        }
    }
    const double dd = (double)di;
    const double xc0 = px * dd, xc1 = py * dd, xc2 = dd;
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
        if (rule == 0) {
            if (dn > 0.0 && fabs((dn - xs[2]) / xs[2]) < 0.01) ++a;
            continue;
        }
        if (!(dn > 0.0)) continue;
        // back-project the neighbour pixel centre at its own depth
        const double nqx = (fx + 0.5 - v.cx) / v.f, nqy = (fy + 0.5 - v.cy) / v.f;
        double npx = nqx, npy = nqy;
        if (v.k1 != 0.0 || v.k2 != 0.0) {
            for (int k2 = 0; k2 < 8; ++k2) {
                const double r2n = npx * npx + npy * npy;
                double dd = 1.0 + v.k1 * r2n;
                dd = dd + (v.k2 * r2n) * r2n;
                npx = nqx / dd;
                npy = nqy / dd;
            }
        }
        const double xn0 = npx * dn, xn1 = npy * dn, xn2 = dn;
        double er[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            double acc = v.R[0 * 3 + r] * xn0;
            acc = acc + v.R[1 * 3 + r] * xn1;
            acc = acc + v.R[2 * 3 + r] * xn2;
            er[r] = acc + v.C[r] - rf.C[r];
        }
        double xr[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            double acc = rf.R[r * 3 + 0] * er[0];
            acc = acc + rf.R[r * 3 + 1] * er[1];
            acc = acc + rf.R[r * 3 + 2] * er[2];
            xr[r] = acc;
        }
        if (!(xr[2] > 0.0)) continue;
        const double rx = xr[0] / xr[2], ry = xr[1] / xr[2];
        const double r2r = rx * rx + ry * ry;
        double dr = 1.0 + rf.k1 * r2r;
        dr = dr + (rf.k2 * r2r) * r2r;
        const double fr = rf.f * dr;
        const double ex = (fr * rx + rf.cx) - ((double)x + 0.5);
        const double ey = (fr * ry + rf.cy) - ((double)y + 0.5);
        if (ex * ex + ey * ey <= 1.0) ++a;
    }
    return a;
}

}  // namespace amvs
