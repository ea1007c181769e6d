"""The fast-arithmetic PatchMatch kernels (MVS_MODE_FAST) against the oracle.

Bars (the float tolerance of the fast path's contract): NCC costs within
2e-5 absolute of the oracle wherever both are defined (the median difference
is ~1e-6; single-source costs, which no best-two average smooths, reach
1.9e-5 on these surveys), definedness equal
except within rounding of an image edge (<= 0.2% of items); full depth maps
within 0.5% relative depth of the oracle's on >= 99% of the pixels valid in
both; the same inputs and schedule give the same maps (determinism), and a
survey batch equals one-view calls.
"""
import sys

import numpy as np
import pytest

from aerial_mvs import capi, survey

import oracle as orc
from conftest import survey_cached

pytestmark = pytest.mark.gpu

COST_TOL = 2e-5


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _hyps(v, n, rng, near=0.6):
    xs = rng.integers(0, v.w, n).astype(np.int32)
    ys = rng.integers(0, v.h, n).astype(np.int32)
    hy = np.zeros((n, 4), np.float32)
    close = rng.uniform(size=n) < near
    hy[:, 0] = np.where(close, v.depth_gt[ys, xs] * rng.uniform(0.98, 1.02, n),
                        rng.uniform(30, 90, n))
    nrm = rng.normal([0, 0, -1], np.where(close, 0.05, 0.4)[:, None], (n, 3))
    hy[:, 1:] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    return xs, ys, hy


# split -1: the one-lane-per-pixel kernels (undistorted cameras)
@pytest.mark.parametrize("name,taps,stride,k,split", [
    ("tiny", 5, 2, 4, 1), ("small", 7, 1, 8, 1), ("small", 7, 1, 8, 2), ("small", 7, 1, 4, 4),
    ("small", 11, 1, 8, 2), ("small", 11, 1, 4, 4), ("tiny", 5, 2, 4, -1),
    ("small", 7, 1, 8, -1), ("small", 7, 1, 3, -1), ("small", 11, 1, 8, -1),
    ("small", 5, 2, 1, -1), ("small", 7, 1, 1, 1), ("small", 7, 1, 1, -1),
    ("tiny", 5, 2, 4, -2), ("small", 7, 1, 8, -2), ("small", 11, 1, 4, -2),
    ("small", 7, 1, 8, -5), ("tiny", 5, 2, 4, -5)])
def test_fast_costs_within_tolerance(gpu_ctx, name, taps, stride, k, split):
    s = survey_cached(name)
    ref = s.reference_index()
    nb = s.neighbours[ref][:k]
    v = s.views[ref]
    srcs = [s.views[j] for j in nb]
    rng = np.random.default_rng(300 + taps + k + split)
    xs, ys, hy = _hyps(v, 4000, rng)
    o = orc.OracleDepth(v, srcs, *s.spec.depth_range, taps=taps, stride=stride)
    want = o.cost(xs, ys, hy)
    gpu_ctx.set_schedule(split, 1, 16, False)
    try:
        got = gpu_ctx.ncc_cost(v, srcs, *s.spec.depth_range,
                               capi.params(taps=taps, stride=stride, fast=True), xs, ys, hy)
    finally:
        gpu_ctx.set_schedule(0, 0, 16, True)
    dw, dg = want < 2.0, got < 2.0
    mismatch = np.mean(dw != dg)
    both = dw & dg
    diff = np.abs(got[both].astype(np.float64) - want[both])
    _say(f"fast {name} {taps}x{taps} K={k} split={split}: defined {dw.mean():.3f} "
         f"mismatch {mismatch:.5f} max|diff| {diff.max():.3g} median {np.median(diff):.3g}")
    assert mismatch <= 0.002
    assert diff.max() <= COST_TOL


@pytest.mark.parametrize("name,k,iters", [("tiny", 4, 3), ("small", 4, 4), ("small", 8, 4)])
def test_fast_depth_maps_within_tolerance(gpu_ctx, name, k, iters):
    s = survey_cached(name)
    spec = s.spec
    ref = s.reference_index()
    nb = s.neighbours[ref][:k]
    v = s.views[ref]
    srcs = [s.views[j] for j in nb]
    taps, stride = spec.taps, spec.stride
    o = orc.OracleDepth(v, srcs, *spec.depth_range, taps=taps, stride=stride)
    od, _, _, _ = o.depth_map(iterations=iters)
    prm = capi.params(taps=taps, stride=stride, iterations=iters, fast=True)
    _say(f"fast depth map {name} K={k} I={iters}")
    d1, n1, c1 = gpu_ctx.depth_map(v, srcs, *spec.depth_range, prm)
    d2, n2, c2 = gpu_ctx.depth_map(v, srcs, *spec.depth_range, prm)
    assert np.array_equal(d1, d2) and np.array_equal(c1, c2)  # deterministic
    both = (d1 > 0) & (od > 0)
    rel = np.abs(d1[both] - od[both]) / od[both]
    within = np.mean(rel <= 0.005)
    valid_agree = np.mean((d1 > 0) == (od > 0))
    _say(f"  valid both {both.mean():.3f}, valid agree {valid_agree:.4f}, "
         f"within 0.5% {within:.5f}, identical {np.mean(d1[both] == od[both]):.3f}")
    assert both.sum() > 1000
    assert within >= 0.99
    assert valid_agree >= 0.98


@pytest.mark.parametrize("name,k,iters,split", [("tiny", 4, 3, -1), ("small", 8, 4, -1),
                                                ("small", 8, 4, -2)])
def test_pixel_lane_depth_maps_within_tolerance(gpu_ctx, name, k, iters, split):
    """Depth maps with every half-sweep on the one-lane-per-pixel kernels
    (split -1: f16 corner plane, -2: 8-bit corner plane)."""
    s = survey_cached(name)
    spec = s.spec
    ref = s.reference_index()
    v = s.views[ref]
    srcs = [s.views[j] for j in s.neighbours[ref][:k]]
    o = orc.OracleDepth(v, srcs, *spec.depth_range, taps=spec.taps, stride=spec.stride)
    od, _, _, _ = o.depth_map(iterations=iters)
    prm = capi.params(taps=spec.taps, stride=spec.stride, iterations=iters, fast=True)
    _say(f"pixel-lane depth map {name} K={k} I={iters} split={split}")
    gpu_ctx.set_schedule(split, 1, 16, False)
    try:
        gpu_ctx.reset_stats()
        d1, _, c1 = gpu_ctx.depth_map(v, srcs, *spec.depth_range, prm)
        st = gpu_ctx.stats()
        d2, _, c2 = gpu_ctx.depth_map(v, srcs, *spec.depth_range, prm)
    finally:
        gpu_ctx.set_schedule(0, 0, 16, True)
    assert st["split"] == split
    assert np.array_equal(d1, d2) and np.array_equal(c1, c2)  # deterministic
    both = (d1 > 0) & (od > 0)
    rel = np.abs(d1[both] - od[both]) / od[both]
    within = np.mean(rel <= 0.005)
    valid_agree = np.mean((d1 > 0) == (od > 0))
    _say(f"  valid both {both.mean():.3f}, valid agree {valid_agree:.4f}, "
         f"within 0.5% {within:.5f}")
    assert both.sum() > 1000
    assert within >= 0.99
    assert valid_agree >= 0.98


def _oblique():
    """An oblique 45-degree view of the small uneven survey and its nearest
    views (nadir and oblique): window scales far from one in the sources."""
    if "oblique" not in _OBL:
        s, per = survey.make_uneven_survey("uneven_small")
        i = max(j for j, p in enumerate(per) if p["kind"] == "oblique")
        _OBL["oblique"] = (s, i, per[i]["range"])
    return _OBL["oblique"]


_OBL = {}


@pytest.mark.parametrize("grey", ["u8", "f32"])
@pytest.mark.parametrize("name,taps,stride,k", [("tiny", 5, 2, 4), ("small", 7, 1, 8),
                                                ("small", 11, 1, 8), ("oblique", 7, 1, 6),
                                                ("oblique", 11, 1, 6)])
def test_pixel_lane_planes_give_identical_costs(gpu_ctx, name, taps, stride, k, grey):
    """The planes of one view hold the same corner values, so the
    one-lane-per-pixel kernels give the same costs, bit for bit, on either.
    8-bit views: split -2 (8-bit plane, bytes unpacked and differenced per
    sample), split -1 (f16 plane with the column differences stored) and split
    -5 (the f16 plane staged in LDS tiles where a batch's windows fit). f32
    views: split -3 (f32 corner-difference plane) and split -4 (the grey plane,
    two row pairs per sample, differenced per sample; windows at the last
    column or row read the next row or the zero padding at weight 0).
    Hypotheses near the surface and scattered, on nadir views and on an
    oblique view whose sources see its windows at scales far from one."""
    if name == "oblique":
        s, ref, rng_d = _oblique()
        nb = list(s.neighbours[ref][1:k + 1]) if s.neighbours[ref][0] == ref else \
            list(s.neighbours[ref][:k])
    else:
        s = survey_cached(name)
        ref = s.reference_index()
        nb = s.neighbours[ref][:k]
        rng_d = s.spec.depth_range
    views = s.views if grey == "u8" else _f32_views(s, 31 + taps + k)
    v = views[ref]
    srcs = [views[j] for j in nb]
    splits = (-1, -2, -5) if grey == "u8" else (-3, -4)
    rng = np.random.default_rng(700 + taps + k)
    n = 30000  # the oblique view's costs are defined for ~3% of random items
    xs = rng.integers(0, v.w, n).astype(np.int32)
    ys = rng.integers(0, v.h, n).astype(np.int32)
    hy = np.zeros((n, 4), np.float32)
    close = rng.uniform(size=n) < 0.5
    gt = v.depth_gt[ys, xs]
    hy[:, 0] = np.where(close & (gt > 0), gt * rng.uniform(0.98, 1.02, n),
                        rng.uniform(rng_d[0], rng_d[1], n))
    up = np.asarray(v.R, np.float64).reshape(3, 3) @ np.array([0.0, 0.0, 1.0])  # camera frame
    nrm = rng.normal(up, np.where(close, 0.05, 0.5)[:, None], (n, 3))
    hy[:, 1:] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    # and first, blocks of 64 neighbouring pixels (32 x 2) with planes near
    # the surface, one 64-item batch each: windows that fit an LDS tile
    # (split -5)
    nblk, m = 60, 2 * taps * stride
    bx = rng.integers(m, v.w - 32 - m, nblk)
    by = rng.integers(m, v.h - 2 - m, nblk)
    gx = (bx[:, None] + np.tile(np.arange(32), 2)[None, :]).ravel().astype(np.int32)
    gy = (by[:, None] + np.repeat(np.arange(2), 32)[None, :]).ravel().astype(np.int32)
    gt = v.depth_gt[gy, gx]
    bh = np.zeros((len(gx), 4), np.float32)
    bh[:, 0] = np.where(gt > 0, gt * rng.uniform(0.999, 1.001, len(gx)), rng_d[0])
    nb_ = rng.normal(up, 0.02, (len(gx), 3))
    bh[:, 1:] = nb_ / np.linalg.norm(nb_, axis=1, keepdims=True)
    xs, ys, hy = np.concatenate([gx, xs]), np.concatenate([gy, ys]), np.concatenate([bh, hy])
    prm = capi.params(taps=taps, stride=stride, fast=True)
    got = {}
    _say(f"planes {grey} {name} {taps}x{taps} K={k}: view {ref} {v.w}x{v.h}, sources {nb}")
    try:
        for split in splits:
            gpu_ctx.set_schedule(split, 1, 16, False)
            gpu_ctx.reset_stats()
            got[split] = gpu_ctx.ncc_cost(v, srcs, *rng_d, prm, xs, ys, hy)
            assert gpu_ctx.stats()["split"] == split
    finally:
        gpu_ctx.set_schedule(0, 0, 16, True)
    a = got[splits[0]]
    defined = a < 2.0
    _say(f"  defined {defined.mean():.3f}, identical "
         f"{[float(np.mean(a == got[sp])) for sp in splits[1:]]}")
    assert defined.sum() >= 800
    for sp in splits[1:]:
        assert np.array_equal(a, got[sp]), sp


@pytest.mark.parametrize("grey", ["u8", "f32"])
def test_pixel_lane_planes_give_identical_depth_maps(gpu_ctx, grey):
    """Whole depth maps (initial costs, every half-sweep) on one plane equal
    those on the others: 8-bit, f16 and LDS-staged f16 planes of 8-bit views,
    corner-difference and grey planes of f32 views."""
    s = survey_cached("small")
    spec = s.spec
    ref = s.reference_index()
    views = s.views if grey == "u8" else _f32_views(s, 41)
    v = views[ref]
    srcs = [views[j] for j in s.neighbours[ref][:8]]
    prm = capi.params(taps=spec.taps, stride=spec.stride, iterations=4, fast=True)
    splits = (-1, -2, -5) if grey == "u8" else (-3, -4)
    maps = {}
    try:
        for split in splits:
            _say(f"planes: depth map small {grey} split={split}")
            gpu_ctx.set_schedule(split, 1, 16, False)
            gpu_ctx.reset_stats()
            maps[split] = gpu_ctx.depth_map(v, srcs, *spec.depth_range, prm)
            assert gpu_ctx.stats()["split"] == split
    finally:
        gpu_ctx.set_schedule(0, 0, 16, True)
    for sp in splits[1:]:
        for a, b in zip(maps[splits[0]], maps[sp]):
            assert np.array_equal(a, b), sp
    assert (maps[splits[0]][0] > 0).mean() > 0.5


@pytest.mark.parametrize("grey,split,iters", [("u8", -1, 4), ("u8", -2, 4), ("f32", -3, 4),
                                              ("u8", -1, 1), ("u8", -5, 1), ("f32", -4, 1)])
def test_skipped_candidates_change_nothing(gpu_ctx, grey, split, iters):
    """A propagation candidate from a neighbour that its last half-sweep left
    unchanged is not scored again by the one-lane-per-pixel kernels (it was
    scored two half-sweeps before and could not be taken now), and a call
    that makes a half-sweep of each colour scores the initial planes inside
    those half-sweeps rather than in a pass of its own. Half-sweeps run in
    one call (initial costs in the first two, skipping from the third on)
    must land on the states of the initialisation and the same half-sweeps
    run one call each (each from a state the caller sets, where neither
    applies), bit for bit, and the candidates skipped plus those scored must
    be the ones scored there."""
    s = survey_cached("small")
    spec = s.spec
    ref = s.reference_index()
    views = s.views if grey == "u8" else _f32_views(s, 43)
    v = views[ref]
    srcs = [views[j] for j in s.neighbours[ref][:8]]
    prm = capi.params(taps=spec.taps, stride=spec.stride, iterations=iters, fast=True)
    gpu_ctx.set_schedule(split, 1, 16, False)
    try:
        _say(f"skip: small {grey} split={split}, {2 * iters} half-sweeps in one call")
        gpu_ctx.reset_stats()
        st_a, _, ev_a = gpu_ctx.patchmatch_steps(v, srcs, *spec.depth_range, prm, -1, 2 * iters)
        skipped = gpu_ctx.stats()["sweep_skipped"]
        _say("skip: the same half-sweeps one call each")
        st_b, _, ev_b = gpu_ctx.patchmatch_steps(v, srcs, *spec.depth_range, prm, -1, 0)
        for step in range(2 * iters):
            st_b, _, e = gpu_ctx.patchmatch_steps(v, srcs, *spec.depth_range, prm, step, 1,
                                                  state=st_b)
            ev_b += e
    finally:
        gpu_ctx.set_schedule(0, 0, 16, True)
    _say(f"  scored {ev_a} + skipped {skipped}, one call each {ev_b}")
    assert skipped > 0 or iters == 1  # (one iteration: nothing to skip yet)
    assert np.array_equal(st_a, st_b)
    assert ev_a + skipped == ev_b


def test_pixel_lane_schedule_leaves_exact_mode_exact(gpu_ctx, tiny_survey):
    """Forcing split -1 applies to fast classes only; exact-mode costs stay
    bit-identical to the oracle's."""
    s = tiny_survey
    ref = s.reference_index()
    v = s.views[ref]
    srcs = [s.views[j] for j in s.neighbours[ref][:4]]
    rng = np.random.default_rng(77)
    xs, ys, hy = _hyps(v, 2000, rng)
    want = orc.OracleDepth(v, srcs, *s.spec.depth_range, taps=5, stride=2).cost(xs, ys, hy)
    gpu_ctx.set_schedule(-1, 1, 16, False)
    try:
        got = gpu_ctx.ncc_cost(v, srcs, *s.spec.depth_range, capi.params(taps=5, stride=2),
                               xs, ys, hy)
    finally:
        gpu_ctx.set_schedule(0, 0, 16, True)
    assert np.array_equal(got, want)


def test_fast_survey_batch_matches_single_view(gpu_ctx, tiny_survey):
    s = tiny_survey
    prm = capi.params(taps=5, stride=2, iterations=3, fast=True, refine_steps=2)
    refs = [0, 4, 7]
    jobs = [(r, list(s.neighbours[r][:4]), s.spec.depth_range) for r in refs]
    sv = capi.Survey(gpu_ctx, s.views)
    try:
        sv.depth_maps(jobs, prm)
        for (r, nb, rng) in jobs:
            db, _, _ = sv.download(r)
            d1, _, _ = gpu_ctx.depth_map(s.views[r], [s.views[j] for j in nb], rng[0], rng[1],
                                         prm)
            assert np.array_equal(db, d1)
    finally:
        sv.close()


def _f32_views(s, seed):
    """The survey's views as the reference's `.view` grey: an RGB photo made
    from each 8-bit image (per-channel offsets and noise), through the grey
    formula of dense.rs:108 in f32 — values off the byte lattice."""
    import copy
    rng = np.random.default_rng(seed)
    out = []
    for v in s.views:
        g = v.gray.astype(np.int32)
        rgb = [np.clip(g + rng.integers(-12, 13, g.shape) + off, 0, 255).astype(np.float32)
               for off in (6, -4, 9)]
        grey = ((np.float32(0.299) * rgb[0] + np.float32(0.587) * rgb[1])
                + np.float32(0.114) * rgb[2]) / np.float32(255.0)
        v2 = copy.copy(v)
        v2.gray_f32 = grey.astype(np.float32)
        out.append(v2)
    return out


@pytest.mark.parametrize("name,taps,stride,k", [("tiny", 5, 2, 4), ("small", 7, 1, 8),
                                                ("small", 11, 1, 4), ("small", 7, 1, 1)])
def test_fast_f32_costs_within_tolerance(gpu_ctx, name, taps, stride, k):
    """f32 grey views on undistorted cameras run the one-lane-per-pixel
    kernels on the f32 corner-difference plane (split -3)."""
    s = survey_cached(name)
    views = _f32_views(s, 5 + taps + k)
    ref = s.reference_index()
    v, srcs = views[ref], [views[j] for j in s.neighbours[ref][:k]]
    rng = np.random.default_rng(500 + taps + k)
    xs, ys, hy = _hyps(v, 4000, rng)
    want = orc.OracleDepth(v, srcs, *s.spec.depth_range, taps=taps, stride=stride).cost(xs, ys, hy)
    gpu_ctx.reset_stats()
    got = gpu_ctx.ncc_cost(v, srcs, *s.spec.depth_range,
                           capi.params(taps=taps, stride=stride, fast=True), xs, ys, hy)
    assert gpu_ctx.stats()["split"] == -3
    dw, dg = want < 2.0, got < 2.0
    mismatch = np.mean(dw != dg)
    both = dw & dg
    diff = np.abs(got[both].astype(np.float64) - want[both])
    _say(f"fast f32 {name} {taps}x{taps} K={k}: defined {dw.mean():.3f} mismatch "
         f"{mismatch:.5f} max|diff| {diff.max():.3g} median {np.median(diff):.3g}")
    assert mismatch <= 0.002
    assert diff.max() <= COST_TOL


@pytest.mark.parametrize("name,k,iters", [("tiny", 4, 3), ("small", 8, 4)])
def test_fast_f32_depth_maps_within_tolerance(gpu_ctx, name, k, iters):
    s = survey_cached(name)
    spec = s.spec
    views = _f32_views(s, 17 + k)
    ref = s.reference_index()
    v, srcs = views[ref], [views[j] for j in s.neighbours[ref][:k]]
    od, _, _, _ = orc.OracleDepth(v, srcs, *spec.depth_range, taps=spec.taps,
                                  stride=spec.stride).depth_map(iterations=iters)
    prm = capi.params(taps=spec.taps, stride=spec.stride, iterations=iters, fast=True)
    _say(f"fast f32 depth map {name} K={k} I={iters}")
    gpu_ctx.reset_stats()
    d1, _, c1 = gpu_ctx.depth_map(v, srcs, *spec.depth_range, prm)
    st = gpu_ctx.stats()
    d2, _, c2 = gpu_ctx.depth_map(v, srcs, *spec.depth_range, prm)
    assert st["split"] in (-3, -4)
    assert np.array_equal(d1, d2) and np.array_equal(c1, c2)  # deterministic
    both = (d1 > 0) & (od > 0)
    rel = np.abs(d1[both] - od[both]) / od[both]
    within = np.mean(rel <= 0.005)
    valid_agree = np.mean((d1 > 0) == (od > 0))
    _say(f"  valid both {both.mean():.3f}, valid agree {valid_agree:.4f}, within 0.5% "
         f"{within:.5f}")
    assert both.sum() > 1000
    assert within >= 0.99 and valid_agree >= 0.98
    # a survey batch of f32 views equals one-view calls
    refs = [ref, (ref + 1) % len(views)]
    jobs = [(r, list(s.neighbours[r][:k]), spec.depth_range) for r in refs]
    sv = capi.Survey(gpu_ctx, views)
    try:
        sv.depth_maps(jobs, prm)
        for (r, nb, rg) in jobs:
            db, _, _ = sv.download(r)
            d1r, _, _ = gpu_ctx.depth_map(views[r], [views[j] for j in nb], rg[0], rg[1], prm)
            assert np.array_equal(db, d1r), r
    finally:
        sv.close()


def test_fast_mode_with_distorted_f32_input_runs_exact_kernels(gpu_ctx, tiny_survey):
    """The f32 pixel-lane kernels need undistorted cameras; distorted f32
    views run the exact kernels, so their costs equal the oracle's."""
    s = tiny_survey
    ref = s.reference_index()
    views = _f32_views(s, 9)
    for v in views:
        v.k1, v.k2 = -0.02, 0.005
    rng = np.random.default_rng(9)
    nb = s.neighbours[ref][:4]
    xs, ys, hy = _hyps(views[ref], 2000, rng)
    got = gpu_ctx.ncc_cost(views[ref], [views[j] for j in nb], 40, 80, capi.params(fast=True),
                           xs, ys, hy)
    want = orc.OracleDepth(views[ref], [views[j] for j in nb], 40, 80).cost(xs, ys, hy)
    assert np.array_equal(got, want)
