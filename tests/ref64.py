"""Independent float64 statement of the NCC cost's mathematics (numpy).

Not a transcription of the reference's operation order: it evaluates the same
definition (plane-induced warp of a window into each source, bilinear sampling
at pixel-centre coordinates, two-pass normalised cross-correlation, mean of the
best two sources) vectorised in float64, so a dropped term, a transposed
operand or a wrong index in the f32 oracle or the GPU kernel shows up as a
difference far above f32 rounding.
"""
import numpy as np


def window_offsets(taps, stride):
    r = taps // 2
    a = (np.arange(taps) - r) * stride
    dy, dx = np.meshgrid(a, a, indexing="ij")
    return dx.ravel(), dy.ravel()


def bilinear64(g, x, y):
    """Grey at pixel-centre coordinates (x, y); NaN where the reference's
    gray_at returns None."""
    h, w = g.shape
    fx, fy = x - 0.5, y - 0.5
    ok = (fx >= 0) & (fy >= 0) & (fx < w - 1) & (fy < h - 1)
    x0 = np.where(ok, np.floor(np.where(ok, fx, 0)), 0).astype(np.int64)
    y0 = np.where(ok, np.floor(np.where(ok, fy, 0)), 0).astype(np.int64)
    ax, ay = fx - x0, fy - y0
    g = g.astype(np.float64)
    v = (g[y0, x0] * (1 - ax) * (1 - ay) + g[y0, x0 + 1] * ax * (1 - ay)
         + g[y0 + 1, x0] * (1 - ax) * ay + g[y0 + 1, x0 + 1] * ax * ay)
    return np.where(ok, v, np.nan)


def cost64(ref_gray, rays, srcs, x, y, hyp, taps, stride, texture_min=0.008):
    """srcs: list of (gray f32 plane, R_rel 3x3, t_rel 3, f, cx, cy) with R_rel,
    t_rel mapping reference-camera points to the source camera."""
    h, w = ref_gray.shape
    dx, dy = window_offsets(taps, stride)
    qx, qy = x + dx, y + dy
    if qx.min() < 0 or qy.min() < 0 or qx.max() >= w or qy.max() >= h:
        return 2.0
    d, n = float(hyp[0]), np.asarray(hyp[1:4], np.float64)
    r = rays[qy, qx].astype(np.float64)                     # (N, 2)
    rc = rays[y, x].astype(np.float64)
    X = np.array([rc[0] * d, rc[1] * d, d])
    den = r @ n[:2] + n[2]
    if np.any(np.abs(den) < 1e-6):
        return 2.0
    dq = (n @ X) / den
    if np.any(dq <= 0):
        return 2.0
    P = np.stack([r[:, 0] * dq, r[:, 1] * dq, dq], axis=1)  # (N, 3)
    refv = ref_gray[qy, qx].astype(np.float64)
    sr = refv.std()
    if sr < texture_min:
        return 2.0
    costs = []
    for (g, R, t, f, cx, cy) in srcs:
        Xs = P @ np.asarray(R, np.float64).T + np.asarray(t, np.float64)
        if np.any(Xs[:, 2] <= 0):
            costs.append(2.0)
            continue
        u = f * Xs[:, 0] / Xs[:, 2] + cx
        v = f * Xs[:, 1] / Xs[:, 2] + cy
        sv = bilinear64(g, u, v)
        if np.any(np.isnan(sv)):
            costs.append(2.0)
            continue
        ss = sv.std()
        if ss < 1e-3:
            costs.append(2.0)
            continue
        ncc = np.mean((refv - refv.mean()) * (sv - sv.mean())) / (sr * ss)
        costs.append(1.0 - ncc)
    costs = sorted(costs)
    return costs[0] if len(costs) == 1 else 0.5 * (costs[0] + costs[1])
