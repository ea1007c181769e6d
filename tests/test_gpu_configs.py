"""Every benchmark workload at its stated shape on the GPU (the 77-view
2000x1500 survey has its own file, test_gpu_fullsize.py):

- 16 views 1024x768, one reference view with 4 neighbours, 7x7 window, 4
  iterations, range [40, 80] m, run in full: the exact kernels' map equals the
  oracle's bit for bit, the fast kernels' map is within 0.5% of it on >= 99%
  of the pixels valid in both.
- 2000x1500 views with an 11x11 window, 8 neighbours, 6 iterations and 5
  Gauss-Newton steps: a 320x240 crop of the reference view against the
  full-size sources (exact identical, fast within 0.5%); the whole view refined
  (deterministic, >= 95% of valid pixels within 0.5% of the FP64 ray-cast
  depth); and on 10^4 random pixels of the refined map the refinement kernel's
  own cost, gradient and J^T J against the f64 oracle and f64 central
  differences (1e-4 m in depth, 1e-5 per normal component).
- The uneven workload at full size (1600x1200 nadir and 1920x1080 oblique
  views; per-view K up to 12 and windows 5x5 / 7x7 / 11x11), six reference
  views of both kinds and every window in one survey call: each map equals a
  one-view call (exact kernels), crops of three of them equal the oracle's, and
  the fast batch stays within 0.5% of the exact one on >= 99%.
- The large survey's 2400x1800 views with 10 neighbours over terrain with box
  extrusions, kept as lean views with depth-only maps: the maps (fast, 5
  refinement steps) equal a full survey's under the same schedule, crops of a
  nadir and an oblique reference view are within 0.5% of the oracle (fast) and
  identical to it (exact), and the refined maps are within 0.5% of the FP64
  depth on >= 80% of their valid pixels (occlusions at the boxes included).

Only the views a case reads are rendered.
"""
import copy
import sys

import numpy as np
import pytest

from aerial_mvs import capi, survey

import oracle as orc

pytestmark = pytest.mark.gpu

_CACHE = {}
COST_TOL = 3e-5


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _within(a, b, tol=0.005):
    both = (a > 0) & (b > 0)
    return float(np.mean(np.abs(a[both] - b[both]) / b[both] <= tol)), float(both.mean())


def _crop(v, cw, ch):
    x0, y0 = (v.w - cw) // 2, (v.h - ch) // 2
    cv = copy.copy(v)
    cv.gray = np.ascontiguousarray(v.gray[y0:y0 + ch, x0:x0 + cw])
    cv.cx, cv.cy = v.cx - x0, v.cy - y0
    cv.depth_gt = None
    return cv


def _crop_parity(ctx, v, srcs, rng, taps, stride, iters, cw=320, ch=240, label=""):
    """Exact crop map == oracle; fast crop map within 0.5% on >= 99%."""
    cv = _crop(v, cw, ch)
    o = orc.OracleDepth(cv, srcs, rng[0], rng[1], taps=taps, stride=stride)
    want, _, _, _ = o.depth_map(iterations=iters)
    geom = dict(taps=taps, stride=stride, iterations=iters)
    ge, _, _ = ctx.depth_map(cv, srcs, rng[0], rng[1], capi.params(**geom))
    gf, _, _ = ctx.depth_map(cv, srcs, rng[0], rng[1], capi.params(**geom, fast=True))
    w, both = _within(gf, want)
    _say(f"  {label} crop {cw}x{ch} {taps}x{taps} K={len(srcs)}: oracle valid "
         f"{np.mean(want > 0):.3f}, exact identical {np.mean(ge == want):.6f}, fast within 0.5% "
         f"{w:.5f}, valid agree {np.mean((gf > 0) == (want > 0)):.5f}")
    assert np.mean(want > 0) > 0.3
    assert np.array_equal(ge, want)
    assert w >= 0.99 and np.mean((gf > 0) == (want > 0)) >= 0.98


# ---- 16 views 1024x768, run in full -------------------------------------------

def test_c1_full_view_exact_identical_and_fast_within_tolerance(gpu_ctx):
    s = survey.make_survey("c1")
    spec = s.spec
    ref = s.reference_index()
    nb = capi.select_neighbours(s.centres(), ref, s.neighbours[ref], spec.depth_range[0],
                                spec.neighbours)
    v, srcs = s.views[ref], [s.views[j] for j in nb]
    assert (v.w, v.h, len(srcs), spec.taps, spec.iterations) == (1024, 768, 4, 7, 4)
    _say(f"c1: view {ref} {v.w}x{v.h}, neighbours {nb}, 7x7, 4 iterations")
    o = orc.OracleDepth(v, srcs, *spec.depth_range, taps=spec.taps, stride=spec.stride)
    want, _, _, ev = o.depth_map(iterations=spec.iterations, nthreads=16)
    geom = dict(taps=spec.taps, stride=spec.stride, iterations=spec.iterations)
    gpu_ctx.reset_stats()
    ge, _, _ = gpu_ctx.depth_map(v, srcs, *spec.depth_range, capi.params(**geom))
    gev = gpu_ctx.stats()["sweep_evals"]  # the half-sweeps scored the same hypotheses
    assert abs(gev - ev) <= max(2, ev // 10000), (gev, ev)
    gf, _, _ = gpu_ctx.depth_map(v, srcs, *spec.depth_range, capi.params(**geom, fast=True))
    w, both = _within(gf, want)
    _say(f"  oracle valid {np.mean(want > 0):.3f} ({ev} evaluations); exact identical "
         f"{np.mean(ge == want):.6f}; fast within 0.5% {w:.5f} of {both:.3f} valid in both")
    assert np.mean(want > 0) > 0.8
    assert np.array_equal(ge, want)
    assert w >= 0.99 and np.mean((gf > 0) == (want > 0)) >= 0.98
    gt = v.depth_gt
    wg, _ = _within(want, gt)
    _say(f"  oracle map within 0.5% of the FP64 depth on {wg:.4f} of its valid pixels")
    assert wg > 0.9


# ---- 2000x1500, 11x11 window, 8 neighbours, refined ------------------------------

def _c4():
    if "c4" not in _CACHE:
        spec = survey.SPECS["c4"]
        lay = survey._layout(spec, np.random.default_rng(spec.seed + 1))
        cent = np.array([[c[0], c[1], 0.0] for (c, _, _, _) in lay])
        m = cent[:, :2].mean(axis=0)
        ref = int(np.argmin(np.hypot(cent[:, 0] - m[0], cent[:, 1] - m[1])))
        near = survey.nearest_neighbours(cent, 8)[ref]
        _say(f"rendering c4 views {ref} + {near}")
        s = survey.make_survey(spec, views=set([ref] + near))
        assert s.reference_index() == ref
        nb = capi.select_neighbours(s.centres(), ref, s.neighbours[ref], spec.depth_range[0],
                                    spec.neighbours)
        assert set(nb) <= set(near)
        _CACHE["c4"] = (s, ref, nb)
    return _CACHE["c4"]


def test_c4_full_size_crop_parity(gpu_ctx):
    s, ref, nb = _c4()
    spec = s.spec
    assert (spec.width, spec.height, spec.taps, len(nb), spec.iterations) == (2000, 1500, 11, 8, 6)
    _say("c4 crop against the oracle")
    _crop_parity(gpu_ctx, s.views[ref], [s.views[j] for j in nb], spec.depth_range, spec.taps,
                 spec.stride, spec.iterations, label="c4")


def test_c4_refined_view_and_gradient_check(gpu_ctx):
    s, ref, nb = _c4()
    spec = s.spec
    views = [s.views[ref]] + [s.views[j] for j in nb]
    prm = capi.params(taps=spec.taps, stride=spec.stride, iterations=spec.iterations,
                      refine_steps=5, fast=True)
    _say("c4 whole view: fast, 5 Gauss-Newton steps, twice")
    sv = capi.Survey(gpu_ctx, views)
    try:
        maps = []
        for _ in range(2):
            gpu_ctx.reset_stats()
            sv.depth_maps([(0, list(range(1, len(views))), spec.depth_range)], prm)
            maps.append(sv.download(0))
        st = gpu_ctx.stats()
    finally:
        sv.close()
    d, n, _ = maps[0]
    assert np.array_equal(maps[0][0], maps[1][0]) and np.array_equal(maps[0][1], maps[1][1])
    assert st["refine_pixels"] > 0 and st["refine_accepted"] > 0
    gt = s.views[ref].depth_gt
    w, both = _within(d, gt)
    _say(f"  refined {st['refine_pixels']} px; valid {both:.3f}, within 0.5% of FP64 {w:.5f}")
    assert both > 0.8 and w >= 0.95

    # 10^4 pixels of the refined map, planes perturbed off the Gauss-Newton
    # minimum (where the gradient vanishes and a relative error says nothing)
    v, srcs = views[0], views[1:]
    rng = np.random.default_rng(0xC4)
    m = 12
    yy, xx = np.nonzero(d[m:-m, m:-m] > 0)
    pick = rng.choice(len(xx), 10_000, replace=False)
    xs, ys = (xx[pick] + m).astype(np.int32), (yy[pick] + m).astype(np.int32)
    nrm = n[ys, xs] + rng.normal(0, 0.05, (len(xs), 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    hy = np.concatenate([(d[ys, xs] * rng.uniform(0.99, 1.01, len(xs)))[:, None], nrm],
                        axis=1).astype(np.float32)
    prm0 = capi.params(taps=spec.taps, stride=spec.stride)
    gpu_ctx.reset_stats()
    cost, pq, grad, jtj = gpu_ctx.refine_terms(v, srcs, *spec.depth_range, prm0, xs, ys, hy)
    assert gpu_ctx.stats()["split"] == -1  # the one-lane-per-pixel refinement kernel
    o = orc.OracleDepth(v, srcs, *spec.depth_range, taps=spec.taps, stride=spec.stride)
    hs = (1e-4, 1e-5, 1e-5, 1e-5)
    rels, jrels, dc = [], [], []
    _say(f"  gradient check on {len(xs)} pixels")
    for k in range(len(xs)):
        if cost[k] >= 2.0:
            continue
        h64 = hy[k].astype(np.float64)
        s0, s1 = int(pq[k][0]), int(pq[k][1])
        base = o.pair_terms64(xs[k], ys[k], h64, s0, s1)
        if base is None:
            continue
        dc.append(abs(float(cost[k]) - base[0]))
        sc = np.array([1.0, 1 / h64[0], 1 / h64[0], 1 / h64[0]])
        S = np.outer(sc, sc)
        jrels.append(np.abs((jtj[k] - base[2]) * S).max() / np.abs(base[2] * S).max())
        fd = np.zeros(4)
        ok = True
        for a in range(4):
            hp, hm = h64.copy(), h64.copy()
            hp[a] += hs[a]
            hm[a] -= hs[a]
            p_, m_ = o.pair_cost64(xs[k], ys[k], hp, s0, s1), o.pair_cost64(xs[k], ys[k], hm,
                                                                           s0, s1)
            if p_ is None or m_ is None or p_[2] != base[3] or m_[2] != base[3]:
                ok = False
                break
            fd[a] = (p_[0] - m_[0]) / (2 * hs[a])
        if ok:
            e = np.abs(grad[k].astype(np.float64) - fd) / sc
            rels.append(e.max() / max((np.abs(fd) / sc).max(), 1e-12))
    rels, jrels, dc = np.array(rels), np.array(jrels), np.array(dc)
    _say(f"  {len(dc)} defined, {len(rels)} differentiable: gradient median "
         f"{np.median(rels):.3g} p99 {np.quantile(rels, 0.99):.3g}; J^T J median "
         f"{np.median(jrels):.3g} p99 {np.quantile(jrels, 0.99):.3g}; cost p99 "
         f"{np.quantile(dc, 0.99):.3g}")
    assert len(rels) > 3000
    assert np.median(rels) < 2e-4 and np.quantile(rels, 0.99) < 3e-3
    assert np.median(jrels) < 2e-4 and np.quantile(jrels, 0.99) < 3e-3
    assert np.quantile(dc, 0.99) < COST_TOL


# ---- the uneven workload at full size ----------------------------------------------

C3_REFS = (6, 9, 10, 223, 225, 228)  # nadir and oblique, every window, K 5..12


def _c3():
    if "c3" not in _CACHE:
        spec = survey.UNEVEN["c3"]
        s0, per = survey.make_uneven_survey(spec, views=set())
        need = set(C3_REFS)
        for r in C3_REFS:
            need |= set(s0.neighbours[r][:per[r]["k"]])
        _say(f"rendering {len(need)} views of the uneven survey at full size")
        s, per = survey.make_uneven_survey(spec, views=need)
        _CACHE["c3"] = (s, per)
    return _CACHE["c3"]


def _c3_jobs(s, per, fast):
    jobs, prms = [], []
    for r in C3_REFS:
        p = per[r]
        jobs.append((r, list(s.neighbours[r][:p["k"]]), p["range"]))
        prms.append(capi.params(taps=p["taps"], stride=p["stride"], iterations=4, fast=fast))
    return jobs, prms


def test_c3_full_size_batch_equals_single_views(gpu_ctx):
    s, per = _c3()
    shapes = {(s.views[r].w, s.views[r].h) for r in C3_REFS}
    assert shapes == {(1600, 1200), (1920, 1080)}
    assert {per[r]["taps"] for r in C3_REFS} == {5, 7, 11}
    assert max(per[r]["k"] for r in C3_REFS) == 12
    jobs, prms = _c3_jobs(s, per, False)
    _say(f"c3: {len(jobs)} full-size views in one call, exact: "
         + ", ".join(f"{r} {per[r]['kind']} {per[r]['taps']}x{per[r]['taps']} K={per[r]['k']}"
                     for r in C3_REFS))
    sv = capi.Survey(gpu_ctx, s.views)
    try:
        sv.depth_maps(jobs, prms)
        exact = {}
        for (r, nb, rg), prm in zip(jobs, prms):
            db, _, cb = sv.download(r)
            d1, _, c1 = gpu_ctx.depth_map(s.views[r], [s.views[j] for j in nb], rg[0], rg[1], prm)
            _say(f"  view {r}: valid {np.mean(db > 0):.3f}, batch == one-view call "
                 f"{np.array_equal(db, d1)}")
            assert np.array_equal(db, d1) and np.array_equal(cb, c1), r
            assert np.mean(db > 0) > 0.1, r
            exact[r] = db
        jobs_f, prms_f = _c3_jobs(s, per, True)
        _say("c3: the same views, fast arithmetic")
        sv.depth_maps(jobs_f, prms_f)
        for (r, _, _) in jobs_f:
            df, _, _ = sv.download(r)
            w, _ = _within(df, exact[r])
            _say(f"  view {r}: fast within 0.5% of exact {w:.5f}")
            assert w >= 0.99, r
    finally:
        sv.close()


@pytest.mark.parametrize("r", [9, 223, 225])
def test_c3_full_size_crop_parity(gpu_ctx, r):
    s, per = _c3()
    p = per[r]
    srcs = [s.views[j] for j in s.neighbours[r][:p["k"]]]
    _say(f"c3 crop of view {r} ({p['kind']} {s.views[r].w}x{s.views[r].h})")
    _crop_parity(gpu_ctx, s.views[r], srcs, p["range"], p["taps"], p["stride"], 4,
                 cw=256, ch=192, label=f"c3 view {r}")


# ---- the large survey's views, lean and depth-only -----------------------------------

C5_REFS = (700, 725, 1425, 1450)  # two nadir, two oblique (views with FP64 depth)


def _c5():
    if "c5" not in _CACHE:
        spec = survey.LARGE["c5"]
        s0, per = survey.make_large_survey(spec, views=set())
        jobs = []
        for r in C5_REFS:
            rg = per[r]["range"]
            nb = capi.select_neighbours(s0.centres(), r, s0.neighbours[r], rg[0],
                                        spec.neighbours)
            jobs.append((r, nb, rg))
        need = sorted(set(C5_REFS) | {j for (_, nb, _) in jobs for j in nb})
        _say(f"rendering {len(need)} views of the large survey on the GPU")
        s, per = survey.make_large_survey(spec, views=set(need))
        # a survey of just these views, re-indexed
        idx = {g: i for i, g in enumerate(need)}
        views = [s.views[g] for g in need]
        ljobs = [(idx[r], [idx[j] for j in nb], rg) for (r, nb, rg) in jobs]
        _CACHE["c5"] = (spec, views, ljobs, jobs, per)
    return _CACHE["c5"]


def test_c5_lean_depth_only_views_match_full_survey(gpu_ctx):
    spec, views, ljobs, jobs, per = _c5()
    assert all((views[j[0]].w, views[j[0]].h) == (2400, 1800) for j in ljobs)
    ks = [len(nb) for (_, nb, _) in ljobs]
    assert max(ks) == 10 and min(ks) >= 8, ks
    prm = capi.params(taps=spec.taps, stride=spec.stride, iterations=spec.iterations,
                      refine_steps=5, fast=True)
    _say(f"c5: {len(ljobs)} reference views 2400x1800 K=10 of {len(views)} resident, lean + "
         "depth-only vs full, split -1")
    gpu_ctx.set_schedule(-1, 1, 16, False)
    lean = capi.Survey(gpu_ctx, views,
                       options=capi.MVS_SURVEY_LEAN_VIEWS | capi.MVS_SURVEY_DEPTH_ONLY)
    full = capi.Survey(gpu_ctx, views)
    try:
        lean.depth_maps(ljobs, prm)
        full.depth_maps(ljobs, prm)
        for (i, _, _), (g, _, _) in zip(ljobs, jobs):
            dl, nl, _ = lean.download(i)
            df, _, _ = full.download(i)
            assert nl is None
            assert np.array_equal(dl, df), g
            w, both = _within(dl, views[i].depth_gt.astype(np.float64))
            _say(f"  view {g} ({per[g]['kind']}): valid {np.mean(dl > 0):.3f}, within 0.5% of "
                 f"FP64 {w:.4f}")
            assert both > 0.5 and w >= 0.8, g
    finally:
        gpu_ctx.set_schedule(0, 0, 16, True)
        lean.close()
        full.close()


@pytest.mark.parametrize("which", [0, 2])
def test_c5_crop_parity(gpu_ctx, which):
    spec, views, ljobs, jobs, per = _c5()
    i, nb, rg = ljobs[which]
    _say(f"c5 crop of view {jobs[which][0]} ({per[jobs[which][0]]['kind']})")
    _crop_parity(gpu_ctx, views[i], [views[j] for j in nb], rg, spec.taps, spec.stride,
                 spec.iterations, label=f"c5 view {jobs[which][0]}")
