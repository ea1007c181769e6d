"""GPU NCC backward pass and Gauss-Newton refinement against the f64 oracle.

Bars (stated in each test): the f32 pair cost within 3e-5 of the oracle's f64
cost; the GPU's choice of the best source pair costs at most 3e-5 more (in
f64) than the f64 best pair; the f32 gradient within 0.3% of the f64 analytic
gradient (max over components, in units where depth and normal steps move
samples alike) on 99% of items and within 0.02% at the median, and the same
against f64 central finite differences; refinement never raises the f64 pair
cost of a pixel by more than 3e-5 and lowers the mean cost and the depth
error against the FP64 ground truth.
"""
import sys

import numpy as np
import pytest

from aerial_mvs import capi

import oracle as orc
from conftest import survey_cached

pytestmark = pytest.mark.gpu

COST_TOL = 3e-5


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _setup(name, k, taps, stride):
    s = survey_cached(name)
    ref = s.reference_index()
    nb = s.neighbours[ref][:k]
    v = s.views[ref]
    srcs = [s.views[j] for j in nb]
    o = orc.OracleDepth(v, srcs, *s.spec.depth_range, taps=taps, stride=stride)
    return s, v, srcs, o


def _items(v, n, rng, margin=12, spread=0.01, tilt=0.05):
    xs = rng.integers(margin, v.w - margin, n).astype(np.int32)
    ys = rng.integers(margin, v.h - margin, n).astype(np.int32)
    hy = np.zeros((n, 4), np.float32)
    hy[:, 0] = v.depth_gt[ys, xs] * rng.uniform(1 - spread, 1 + spread, n)
    nrm = rng.normal([0, 0, -1], tilt, (n, 3))
    hy[:, 1:] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    return xs, ys, hy


def _grad_rel_err(g, want, depth):
    scale = np.stack([np.ones_like(depth), 1 / depth, 1 / depth, 1 / depth], axis=1)
    e = np.abs(g - want) / scale
    r = np.abs(want) / scale
    return e.max(axis=1) / np.maximum(r.max(axis=1), 1e-12)


CASES = [("tiny", 4, 5, 2), ("small", 4, 7, 1), ("small", 8, 7, 1), ("small", 4, 11, 1)]


@pytest.mark.parametrize("name,k,taps,stride", CASES)
def test_pair_cost_and_gradient_match_oracle(gpu_ctx, name, k, taps, stride):
    s, v, srcs, o = _setup(name, k, taps, stride)
    prm = capi.params(taps=taps, stride=stride)
    rng = np.random.default_rng(21 + taps + k)
    xs, ys, hy = _items(v, 1500, rng)
    pairs = np.stack([rng.integers(0, k, len(xs)), rng.integers(0, k, len(xs))], 1)
    pairs[:, 1] = np.where(pairs[:, 0] == pairs[:, 1], (pairs[:, 1] + 1) % k, pairs[:, 1])
    _say(f"pair cost {name} k={k} taps={taps}: {len(xs)} items")
    cost, pq, grad = gpu_ctx.ncc_pair(v, srcs, *s.spec.depth_range, prm, xs, ys, hy,
                                      pairs=pairs)
    assert np.array_equal(pq, pairs)
    want, wgrad = o.pair_costs64(xs, ys, hy.astype(np.float64), pairs)
    undefined = np.isnan(want)
    # undefined in f64 <=> undefined on the GPU, except at the boundary of a
    # defined region (a sample within rounding of an image edge)
    mismatch = np.mean(undefined != (cost >= 2.0))
    _say(f"  defined {np.mean(~undefined):.3f}, definedness mismatch {mismatch:.4f}")
    assert mismatch <= 0.002
    both = ~undefined & (cost < 2.0)
    assert both.sum() > 300
    d = np.abs(cost[both] - want[both])
    _say(f"  cost max|diff| {d.max():.3g}")
    assert d.max() <= COST_TOL
    assert np.all(grad[cost >= 2.0] == 0)
    rel = _grad_rel_err(grad[both].astype(np.float64), wgrad[both], hy[both, 0].astype(np.float64))
    _say(f"  grad rel err median {np.median(rel):.3g} p99 {np.quantile(rel, 0.99):.3g}")
    assert np.median(rel) < 2e-4
    assert np.quantile(rel, 0.99) < 3e-3


def test_pair_gradient_matches_f64_finite_differences(gpu_ctx):
    s, v, srcs, o = _setup("small", 4, 7, 1)
    prm = capi.params(taps=7, stride=1)
    rng = np.random.default_rng(33)
    xs, ys, hy = _items(v, 800, rng, spread=0.02, tilt=0.1)
    pairs = np.tile(np.array([[0, 1]], np.int32), (len(xs), 1))
    cost, _, grad = gpu_ctx.ncc_pair(v, srcs, *s.spec.depth_range, prm, xs, ys, hy, pairs=pairs)
    hs = (1e-4, 1e-5, 1e-5, 1e-5)
    rels = []
    for i in range(len(xs)):
        if cost[i] >= 2.0:
            continue
        h64 = hy[i].astype(np.float64)
        base = o.pair_cost64(xs[i], ys[i], h64, 0, 1)
        if base is None:
            continue
        fd = np.zeros(4)
        ok = True
        for a in range(4):
            hp, hm = h64.copy(), h64.copy()
            hp[a] += hs[a]
            hm[a] -= hs[a]
            p, m = o.pair_cost64(xs[i], ys[i], hp, 0, 1), o.pair_cost64(xs[i], ys[i], hm, 0, 1)
            if p is None or m is None or p[2] != base[2] or m[2] != base[2]:
                ok = False
                break
            fd[a] = (p[0] - m[0]) / (2 * hs[a])
        if ok:
            rels.append(_grad_rel_err(grad[i:i + 1].astype(np.float64), fd[None],
                                      np.array([h64[0]]))[0])
    rels = np.array(rels)
    _say(f"FD check: {len(rels)} items, median {np.median(rels):.3g} "
         f"p99 {np.quantile(rels, 0.99):.3g}")
    assert len(rels) > 200
    assert np.median(rels) < 2e-4
    assert np.quantile(rels, 0.99) < 3e-3


def test_best_pair_selection(gpu_ctx):
    s, v, srcs, o = _setup("small", 8, 7, 1)
    k = 8
    prm = capi.params(taps=7, stride=1)
    rng = np.random.default_rng(41)
    xs, ys, hy = _items(v, 600, rng)
    cost, pq, _ = gpu_ctx.ncc_pair(v, srcs, *s.spec.depth_range, prm, xs, ys, hy, backward=False)
    same = n = 0
    for i in range(len(xs)):
        single = []
        for sidx in range(k):
            r = o.pair_cost64(xs[i], ys[i], hy[i].astype(np.float64), sidx, sidx)
            if r is not None:
                single.append((r[0], sidx))
        if len(single) < 2:
            assert tuple(pq[i]) == (-1, -1) or cost[i] >= 2.0
            continue
        single.sort()
        best = 0.5 * (single[0][0] + single[1][0])
        assert pq[i][0] >= 0 and pq[i][0] != pq[i][1]
        chosen = o.pair_cost64(xs[i], ys[i], hy[i].astype(np.float64), pq[i][0], pq[i][1])
        assert chosen is not None
        assert chosen[0] <= best + COST_TOL
        assert abs(cost[i] - best) <= COST_TOL
        same += set(pq[i]) == {single[0][1], single[1][1]}
        n += 1
    _say(f"best pair: {n} items, identical pair {same / n:.4f}")
    assert n > 300
    assert same / n > 0.99


def test_grad_out_scales_the_gradient(gpu_ctx):
    s, v, srcs, o = _setup("tiny", 4, 5, 2)
    prm = capi.params(taps=5, stride=2)
    rng = np.random.default_rng(43)
    xs, ys, hy = _items(v, 300, rng)
    c1, pq, g1 = gpu_ctx.ncc_pair(v, srcs, *s.spec.depth_range, prm, xs, ys, hy)
    go = rng.uniform(-2, 2, len(xs)).astype(np.float32)
    c2, pq2, g2 = gpu_ctx.ncc_pair(v, srcs, *s.spec.depth_range, prm, xs, ys, hy, pairs=pq,
                                   grad_out=go)
    assert np.array_equal(c1, c2)
    np.testing.assert_allclose(g2, g1 * go[:, None], rtol=1e-6, atol=0)


def test_pair_argument_errors(gpu_ctx):
    s, v, srcs, o = _setup("tiny", 4, 5, 2)
    prm = capi.params(taps=5, stride=2)
    xs = np.array([80], np.int32)
    ys = np.array([60], np.int32)
    hy = np.array([[50.0, 0, 0, -1]], np.float32)
    for bad in ([[0, 0]], [[0, 4]], [[-1, 2]]):
        with pytest.raises(capi.MvsError):
            gpu_ctx.ncc_pair(v, srcs, *s.spec.depth_range, prm, xs, ys, hy, pairs=bad)
    with pytest.raises(capi.MvsError):
        gpu_ctx.ncc_pair(v, srcs[:1], *s.spec.depth_range, prm, xs, ys, hy)
    c, pq, g = gpu_ctx.ncc_pair(v, srcs, *s.spec.depth_range, prm, xs[:0], ys[:0], hy[:0])
    assert len(c) == 0


def _jtj_rel_err(H, want, depth):
    """Max-abs error of J^T J relative to its largest entry, depth and normal
    entries weighed alike (normal rows and columns scaled by 1/depth)."""
    sc = np.stack([np.ones_like(depth), 1 / depth, 1 / depth, 1 / depth], axis=1)
    S = sc[:, :, None] * sc[:, None, :]
    e = np.abs((H - want) * S).reshape(len(H), -1).max(axis=1)
    r = np.abs(want * S).reshape(len(H), -1).max(axis=1)
    return e / np.maximum(r, 1e-30)


# view kinds: "u8" and "f32" undistorted views run the one-lane-per-pixel
# kernel (f16 corner plane / f32 corner-difference plane), "dist" (8-bit views
# of the distorted survey) the two-lanes-per-pixel kernel
TERMS_CASES = [("tiny", 4, 5, 2, "u8"), ("small", 4, 7, 1, "u8"), ("small", 8, 7, 1, "u8"),
               ("small", 4, 11, 1, "u8"), ("small", 4, 7, 1, "f32"), ("small", 8, 11, 1, "f32"),
               ("small_dist", 4, 7, 1, "dist"), ("small_dist", 8, 11, 1, "dist")]


@pytest.mark.parametrize("name,k,taps,stride,kind", TERMS_CASES)
def test_refinement_terms_match_oracle(gpu_ctx, name, k, taps, stride, kind):
    """The Gauss-Newton terms the refinement kernels form — one lane per pixel
    on undistorted views (k_refine_lane, 8-bit or f32), two lanes per pixel
    on distorted ones (k_refine) — against the f64 oracle: pair cost within
    3e-5, cost gradient and J^T J within 0.02% at the median and 0.3% on 99%
    of items (max-abs relative, depth and normal weighed alike), pairs chosen
    at the plane costing no more than the f64 best pair."""
    s, v, srcs, o = _setup(name, k, taps, stride)
    f32 = kind == "f32"
    if f32:
        v, srcs = _as_f32([v])[0], _as_f32(srcs)
    prm = capi.params(taps=taps, stride=stride)
    rng = np.random.default_rng(61 + taps + k + 7 * f32 + 13 * (kind == "dist"))
    xs, ys, hy = _items(v, 1500, rng)
    pairs = np.stack([rng.integers(0, k, len(xs)), rng.integers(0, k, len(xs))], 1)
    pairs[:, 1] = np.where(pairs[:, 0] == pairs[:, 1], (pairs[:, 1] + 1) % k, pairs[:, 1])
    pairs[: len(xs) // 4] = -1  # a quarter of the items choose their pair
    gpu_ctx.reset_stats()
    _say(f"refine terms {name} k={k} taps={taps} {kind}: {len(xs)} items")
    cost, pq, grad, jtj = gpu_ctx.refine_terms(v, srcs, *s.spec.depth_range, prm, xs, ys, hy,
                                               pairs=pairs)
    assert gpu_ctx.stats()["split"] == (1 if kind == "dist" else -1)  # the kernel under test
    given = pairs[:, 0] >= 0
    assert np.array_equal(pq[given], pairs[given])
    chose = ~given & (pq[:, 0] >= 0)
    want = np.full(len(xs), np.nan)
    wgrad = np.full((len(xs), 4), np.nan)
    wjtj = np.full((len(xs), 4, 4), np.nan)
    for i in range(len(xs)):
        if pq[i, 0] < 0:
            continue
        t = o.pair_terms64(xs[i], ys[i], hy[i].astype(np.float64), pq[i, 0], pq[i, 1])
        if t is not None:
            want[i], wgrad[i], wjtj[i] = t[0], t[1], t[2]
    undefined = np.isnan(want)
    mismatch = np.mean(undefined[given] != (cost[given] >= 2.0))
    _say(f"  defined {np.mean(~undefined):.3f}, definedness mismatch {mismatch:.4f}")
    assert mismatch <= 0.002
    both = ~undefined & (cost < 2.0)
    assert both.sum() > 300
    d = np.abs(cost[both] - want[both])
    _say(f"  cost max|diff| {d.max():.3g}")
    assert d.max() <= COST_TOL
    assert np.all(grad[cost >= 2.0] == 0) and np.all(jtj[cost >= 2.0] == 0)
    dep = hy[both, 0].astype(np.float64)
    rel = _grad_rel_err(grad[both].astype(np.float64), wgrad[both], dep)
    relj = _jtj_rel_err(jtj[both].astype(np.float64), wjtj[both], dep)
    _say(f"  grad rel err median {np.median(rel):.3g} p99 {np.quantile(rel, 0.99):.3g}; "
         f"J^T J rel err median {np.median(relj):.3g} p99 {np.quantile(relj, 0.99):.3g}")
    assert np.median(rel) < 2e-4 and np.quantile(rel, 0.99) < 3e-3
    assert np.median(relj) < 2e-4 and np.quantile(relj, 0.99) < 3e-3
    # chosen pairs: no worse (in f64) than the f64 best pair at the plane
    n_chose = 0
    for i in np.nonzero(chose)[0][:150]:
        single = []
        for sidx in range(k):
            r = o.pair_cost64(xs[i], ys[i], hy[i].astype(np.float64), sidx, sidx)
            if r is not None:
                single.append(r[0])
        if len(single) < 2 or np.isnan(want[i]):
            continue
        single.sort()
        assert want[i] <= 0.5 * (single[0] + single[1]) + COST_TOL
        n_chose += 1
    assert n_chose > 50


def test_lane_refinement_terms_match_finite_differences(gpu_ctx):
    """The one-lane-per-pixel refinement kernel's gradient and J^T J against
    f64 central differences (1e-4 m in depth, 1e-5 per normal component) of
    the oracle's pair cost and of its residual vectors, skipping items whose
    bilinear cells change within a step."""
    from test_oracle_gradient import _jtj_fd
    s, v, srcs, o = _setup("small", 4, 7, 1)
    prm = capi.params(taps=7, stride=1)
    rng = np.random.default_rng(67)
    xs, ys, hy = _items(v, 600, rng, spread=0.02, tilt=0.1)
    pairs = np.tile(np.array([[0, 1]], np.int32), (len(xs), 1))
    gpu_ctx.reset_stats()
    cost, _, grad, jtj = gpu_ctx.refine_terms(v, srcs, *s.spec.depth_range, prm, xs, ys, hy,
                                              pairs=pairs)
    assert gpu_ctx.stats()["split"] == -1
    hs = (1e-4, 1e-5, 1e-5, 1e-5)
    rg, rj = [], []
    for i in range(len(xs)):
        if cost[i] >= 2.0:
            continue
        h64 = hy[i].astype(np.float64)
        base = o.pair_cost64(xs[i], ys[i], h64, 0, 1)
        if base is None:
            continue
        fd = np.zeros(4)
        ok = True
        for a in range(4):
            hp, hm = h64.copy(), h64.copy()
            hp[a] += hs[a]
            hm[a] -= hs[a]
            p, m = o.pair_cost64(xs[i], ys[i], hp, 0, 1), o.pair_cost64(xs[i], ys[i], hm, 0, 1)
            if p is None or m is None or p[2] != base[2] or m[2] != base[2]:
                ok = False
                break
            fd[a] = (p[0] - m[0]) / (2 * hs[a])
        hfd = _jtj_fd(o, xs[i], ys[i], h64, 0, 1) if ok else None
        if hfd is None:
            continue
        dep = np.array([h64[0]])
        rg.append(_grad_rel_err(grad[i:i + 1].astype(np.float64), fd[None], dep)[0])
        rj.append(_jtj_rel_err(jtj[i:i + 1].astype(np.float64), hfd[None], dep)[0])
    rg, rj = np.array(rg), np.array(rj)
    _say(f"lane FD check: {len(rg)} items; grad median {np.median(rg):.3g} "
         f"p99 {np.quantile(rg, 0.99):.3g}; J^T J median {np.median(rj):.3g} "
         f"p99 {np.quantile(rj, 0.99):.3g}")
    assert len(rg) > 150
    assert np.median(rg) < 2e-4 and np.quantile(rg, 0.99) < 3e-3
    assert np.median(rj) < 2e-4 and np.quantile(rj, 0.99) < 3e-3


def _as_f32(views):
    """The same views given as f32 grey (the table values of their bytes): the
    refinement then samples the f32 corner-difference plane instead of the f16
    corner plane of 8-bit views."""
    import copy
    lut = orc.gray_lut()
    out = []
    for v in views:
        v2 = copy.copy(v)
        v2.gray_f32 = lut[v.gray].astype(np.float32)
        out.append(v2)
    return out


@pytest.mark.parametrize("name,k,taps,stride,iters,f32", [("tiny", 4, 5, 2, 3, False),
                                                          ("small", 4, 7, 1, 4, False),
                                                          ("small", 4, 7, 1, 4, True),
                                                          ("small_dist", 4, 7, 1, 4, False)])
def test_refinement_lowers_cost_and_depth_error(gpu_ctx, name, k, taps, stride, iters, f32):
    """One lane per pixel on undistorted 8-bit and f32 views, two lanes per
    pixel on the distorted survey."""
    s, v, srcs, o = _setup(name, k, taps, stride)
    if f32:
        v, srcs = _as_f32([v])[0], _as_f32(srcs)
    rng0, rng1 = s.spec.depth_range
    base = capi.params(taps=taps, stride=stride, iterations=iters)
    refp = capi.params(taps=taps, stride=stride, iterations=iters, refine_steps=4)
    _say(f"refine {name} ({'f32' if f32 else '8-bit'} views): PatchMatch then 4 Gauss-Newton "
         "steps")
    d0, n0, c0 = gpu_ctx.depth_map(v, srcs, rng0, rng1, base)
    gpu_ctx.reset_stats()
    d1, n1, c1 = gpu_ctx.depth_map(v, srcs, rng0, rng1, refp)
    st = gpu_ctx.stats()
    _say(f"  stats: refined {st['refine_pixels']} px, accepted {st['refine_accepted']} steps")
    assert np.array_equal(c0, c1)  # the PatchMatch cost (the output mask) is kept
    assert st["refine_pixels"] > 0
    valid = (d0 > 0) & (d1 > 0)
    assert np.mean((d0 > 0) == (d1 > 0)) > 0.999
    ys, xs = np.nonzero(valid)
    h0 = np.concatenate([d0[valid][:, None], n0.reshape(v.h, v.w, 3)[valid]], 1)
    h1 = np.concatenate([d1[valid][:, None], n1.reshape(v.h, v.w, 3)[valid]], 1)
    # the pair the refinement used: the best pair at the PatchMatch plane
    rng = np.random.default_rng(51)
    pick = rng.choice(len(xs), size=min(1500, len(xs)), replace=False)
    xs, ys, h0, h1 = xs[pick].astype(np.int32), ys[pick].astype(np.int32), h0[pick], h1[pick]
    _, pq, _ = gpu_ctx.ncc_pair(v, srcs, rng0, rng1, base, xs, ys, h0, backward=False)
    ok = pq[:, 0] >= 0
    c_pre, _ = o.pair_costs64(xs[ok], ys[ok], h0[ok].astype(np.float64), pq[ok])
    c_post, _ = o.pair_costs64(xs[ok], ys[ok], h1[ok].astype(np.float64), pq[ok])
    fin = np.isfinite(c_pre) & np.isfinite(c_post)
    worse = c_post[fin] - c_pre[fin]
    _say(f"  f64 pair cost mean {c_pre[fin].mean():.5f} -> {c_post[fin].mean():.5f}, "
         f"max rise {worse.max():.3g}, moved {np.mean(np.any(h0 != h1, axis=1)):.3f}")
    assert fin.sum() > 500
    assert worse.max() <= COST_TOL
    assert c_post[fin].mean() < c_pre[fin].mean()
    gt = v.depth_gt
    e0 = np.abs(d0[valid] - gt[valid]) / gt[valid]
    e1 = np.abs(d1[valid] - gt[valid]) / gt[valid]
    _say(f"  median rel depth error {np.median(e0):.3g} -> {np.median(e1):.3g}; "
         f"within 0.2%: {np.mean(e0 <= 0.002):.4f} -> {np.mean(e1 <= 0.002):.4f}")
    assert np.median(e1) < np.median(e0)
    assert np.mean(e1 <= 0.002) > np.mean(e0 <= 0.002)


def test_refinement_batch_matches_single_view(gpu_ctx, tiny_survey):
    s = tiny_survey
    prm = capi.params(taps=5, stride=2, iterations=3, refine_steps=3)
    refs = [0, 3, 5]
    jobs = [(r, list(s.neighbours[r][:4]), s.spec.depth_range) for r in refs]
    sv = capi.Survey(gpu_ctx, s.views)
    try:
        sv.depth_maps(jobs, prm)
        for (r, nb, rng) in jobs:
            db, nbm, cb = sv.download(r)
            d1, n1, c1 = gpu_ctx.depth_map(s.views[r], [s.views[j] for j in nb], rng[0], rng[1],
                                           prm)
            assert np.array_equal(db, d1)
            assert np.array_equal(nbm.reshape(-1), n1.reshape(-1))
    finally:
        sv.close()
