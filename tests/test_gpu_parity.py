"""Parity of the gfx950 engine (through the C ABI) with the CPU oracle.

Bars (stated in each test): NCC costs within 1e-5 absolute of the oracle and
bit-exact in practice (the kernels round operation by operation like the
reference); PatchMatch state after each half-sweep identical to the oracle's
when both start from the same state; full depth maps within 0.5% relative depth
on >= 99% of the pixels valid in both; integer work (textured mask, evaluation
counts, consistency agree counts) bit-exact.
"""
import copy
import sys

import numpy as np
import pytest

from aerial_mvs import capi

import oracle as orc

pytestmark = pytest.mark.gpu

COST_TOL = 1e-5


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _random_hyps(survey, ref, n, rng, near_truth=0.5):
    v = survey.views[ref]
    xs = rng.integers(0, v.w, n).astype(np.int32)
    ys = rng.integers(0, v.h, n).astype(np.int32)
    hy = np.zeros((n, 4), np.float32)
    near = rng.uniform(size=n) < near_truth
    hy[:, 0] = np.where(near, v.depth_gt[ys, xs] * rng.uniform(0.98, 1.02, n),
                        rng.uniform(30, 90, n))
    nrm = rng.normal([0, 0, -1], np.where(near, 0.05, 0.4)[:, None], (n, 3))
    hy[:, 1:] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    return xs, ys, hy


def _compare_costs(got, want, label):
    diff = np.abs(got.astype(np.float64) - want.astype(np.float64))
    exact = float(np.mean(got == want))
    _say(f"{label}: n={len(got)} exact={exact:.6f} max|diff|={diff.max():.3g} "
         f"defined={np.mean(want < 2):.3f}")
    assert diff.max() <= COST_TOL, label
    return exact


@pytest.mark.parametrize("name,taps,stride,k", [
    ("tiny", 5, 2, 4), ("tiny", 5, 2, 1), ("tiny", 5, 2, 2), ("small", 7, 1, 4),
    ("small", 7, 1, 8), ("small", 5, 2, 3), ("small", 7, 1, 6)])
def test_ncc_cost_parity(request, gpu_ctx, name, taps, stride, k):
    from conftest import survey_cached
    s = survey_cached(name)
    ref = s.reference_index()
    nb = s.neighbours[ref][:k]
    rng = np.random.default_rng(100 + k + taps)
    xs, ys, hy = _random_hyps(s, ref, 4000, rng)
    r0, r1 = s.spec.depth_range
    prm = capi.params(taps=taps, stride=stride)
    _say(f"case ncc_cost {name} taps={taps} stride={stride} K={k}")
    got = gpu_ctx.ncc_cost(s.views[ref], [s.views[j] for j in nb], r0, r1, prm, xs, ys, hy)
    want = orc.OracleDepth(s.views[ref], [s.views[j] for j in nb], r0, r1, taps=taps,
                           stride=stride).cost(xs, ys, hy)
    exact = _compare_costs(got, want, f"{name}/{taps}x{stride}/K{k}")
    assert exact > 0.999


def test_ncc_cost_general_camera_path(gpu_ctx, small_survey):
    """Distortion and non-f32 intrinsics take the float-float projection; the
    oracle evaluates the reference's f64 distortion."""
    s = small_survey
    ref = s.reference_index()
    views = []
    for v in s.views:
        v2 = copy.copy(v)
        v2.k1, v2.k2 = -0.02, 0.005
        v2.f = v.f * 1.0000001
        v2.cx = v.cx + 0.1234567
        views.append(v2)
    nb = s.neighbours[ref][:4]
    rng = np.random.default_rng(5)
    xs, ys, hy = _random_hyps(s, ref, 4000, rng)
    prm = capi.params(taps=7, stride=1)
    _say("case ncc_cost general camera")
    got = gpu_ctx.ncc_cost(views[ref], [views[j] for j in nb], 40, 80, prm, xs, ys, hy)
    want = orc.OracleDepth(views[ref], [views[j] for j in nb], 40, 80, taps=7,
                           stride=1).cost(xs, ys, hy)
    same_def = np.mean((got < 2) == (want < 2))
    both = (got < 2) & (want < 2)
    diff = np.abs(got[both] - want[both])
    _say(f"general: same-defined={same_def:.5f} max|diff|={diff.max():.3g} "
         f"exact={np.mean(got == want):.4f}")
    assert same_def > 0.999
    assert np.quantile(diff, 0.999) <= COST_TOL


def test_ncc_cost_f32_grey_input(gpu_ctx, tiny_survey):
    """`.view`-style f32 grey input (MVS_GRAY_F32) is used unchanged."""
    s = tiny_survey
    ref = s.reference_index()
    lut = orc.gray_lut()
    rng = np.random.default_rng(6)
    views = []
    for v in s.views:
        v2 = copy.copy(v)
        v2.gray_f32 = (lut[v.gray] * 0.9 + rng.uniform(0, 0.05, v.gray.shape)).astype(np.float32)
        views.append(v2)
    nb = s.neighbours[ref][:4]
    xs, ys, hy = _random_hyps(s, ref, 3000, rng)
    prm = capi.params()
    _say("case ncc_cost f32 grey")
    got = gpu_ctx.ncc_cost(views[ref], [views[j] for j in nb], 40, 80, prm, xs, ys, hy)
    want = orc.OracleDepth(views[ref], [views[j] for j in nb], 40, 80).cost(xs, ys, hy)
    assert _compare_costs(got, want, "f32 grey") > 0.999


@pytest.mark.parametrize("name,taps,stride,k,iters", [("tiny", 5, 2, 4, 3), ("small", 7, 1, 4, 2)])
def test_patchmatch_lock_step(gpu_ctx, name, taps, stride, k, iters):
    """Init, then every half-sweep started from the oracle's own state: the GPU
    must land on the oracle's next state, with the same evaluation count."""
    from conftest import survey_cached
    s = survey_cached(name)
    ref = s.reference_index()
    rv, sv = s.views[ref], [s.views[j] for j in s.neighbours[ref][:k]]
    r0, r1 = s.spec.depth_range
    prm = capi.params(taps=taps, stride=stride, iterations=iters)
    octx = orc.OracleDepth(rv, sv, r0, r1, taps=taps, stride=stride)
    _say(f"case lock-step init {name}")
    ost, otex = octx.init()
    gst, gtex, gev = gpu_ctx.patchmatch_steps(rv, sv, r0, r1, prm, -1, 0)
    assert np.array_equal(gtex, otex)
    assert gev == int(otex.sum())
    same = np.all(gst == ost, axis=-1)
    _say(f"init: identical pixels {same.mean():.6f}")
    assert np.array_equal(gst[..., :4], ost[..., :4])          # RNG draws bit-exact
    assert np.abs(gst[..., 4] - ost[..., 4]).max() <= COST_TOL
    assert same.mean() > 0.999
    for step in range(2 * iters):
        it, par = step // 2, step % 2
        _say(f"case lock-step half-sweep {name} it={it} parity={par}")
        start = ost.copy()
        oev = octx.half_sweep(ost, otex, it, par)
        gst, _, gev = gpu_ctx.patchmatch_steps(rv, sv, r0, r1, prm, step, 1, state=start)
        same = np.all(gst == ost, axis=-1)
        _say(f"  identical pixels {same.mean():.6f} evals gpu={gev} oracle={oev}")
        assert same.mean() > 0.999
        assert abs(gev - oev) <= max(2, oev // 10000)


@pytest.mark.parametrize("name,k", [("tiny", 4), ("small", 4)])
def test_full_depth_map_parity(gpu_ctx, name, k):
    from conftest import survey_cached
    s = survey_cached(name)
    ref = s.reference_index()
    rv, sv = s.views[ref], [s.views[j] for j in s.neighbours[ref][:k]]
    spec = s.spec
    r0, r1 = spec.depth_range
    prm = capi.params(taps=spec.taps, stride=spec.stride, iterations=spec.iterations)
    _say(f"case full depth map {name}")
    gd, gn, gc = gpu_ctx.depth_map(rv, sv, r0, r1, prm)
    od, ost, _, _ = orc.OracleDepth(rv, sv, r0, r1, taps=spec.taps, stride=spec.stride).depth_map(
        iterations=spec.iterations)
    both = (gd > 0) & (od > 0)
    rel = np.abs(gd[both] - od[both]) / od[both]
    within = float(np.mean(rel <= 0.005))
    _say(f"{name}: valid gpu={np.mean(gd > 0):.4f} oracle={np.mean(od > 0):.4f} "
         f"within0.5%={within:.6f} identical={np.mean(gd == od):.6f}")
    assert within >= 0.99
    assert abs(np.mean(gd > 0) - np.mean(od > 0)) < 0.01
    assert np.array_equal(gn, ost[..., 1:4]) or np.mean(np.all(gn == ost[..., 1:4], -1)) > 0.99


def test_survey_batch_and_consistency(gpu_ctx, tiny_survey):
    s = tiny_survey
    spec = s.spec
    prm = capi.params(taps=spec.taps, stride=spec.stride, iterations=spec.iterations)
    centres = s.centres()
    jobs = []
    for i in range(len(s.views)):
        nb = capi.select_neighbours(centres, i, s.neighbours[i], spec.depth_range[0],
                                    spec.neighbours)
        jobs.append((i, nb, spec.depth_range))
    _say("case survey batch")
    sv = capi.Survey(gpu_ctx, s.views)
    seen = []
    sv.depth_maps(jobs, prm, progress=lambda u, st, d, t: seen.append((st, d, t)) or 0)
    assert [x[1] for x in seen] == list(range(1, len(jobs) + 1))
    maps = {i: sv.download(i)[0] for i in range(len(s.views))}
    # the batch equals one-view calls
    i0 = s.reference_index()
    d0, _, _ = gpu_ctx.depth_map(s.views[i0], [s.views[j] for j in jobs[i0][1]],
                                 *spec.depth_range, prm)
    assert np.array_equal(d0, maps[i0])
    # consistency (dense.rs:575-615): agree counts equal the oracle's
    _say("case consistency")
    for ref in (i0, 0):
        nbrs = s.neighbours[ref][:spec.neighbours + 1]
        agree, n_ok = sv.consistency(ref, nbrs)
        oag, on = orc.consistency(s.views[ref], maps[ref], [(s.views[j], maps[j]) for j in nbrs])
        _say(f"ref {ref}: consistent gpu={n_ok} oracle={on}")
        assert np.array_equal(agree, oag) and n_ok == on
        # reprojection rule (<= 1 px in >= 3 views): both sides compute in f64
        # with the same operation order, so the counts are identical
        agree, n_ok = sv.consistency(ref, nbrs, rule=capi.MVS_CONS_REPROJ)
        oag, on = orc.consistency(s.views[ref], maps[ref], [(s.views[j], maps[j]) for j in nbrs],
                                  rule=1, need=3)
        _say(f"ref {ref}: reprojection-consistent gpu={n_ok} oracle={on}")
        assert np.array_equal(agree, oag) and n_ok == on
        assert 0 < n_ok
    with pytest.raises(capi.MvsError):
        sv.consistency(i0, s.neighbours[i0][:2], rule=7)
    sv.close()


def test_edge_cases(gpu_ctx, tiny_survey):
    s = tiny_survey
    ref = s.reference_index()
    rv = s.views[ref]
    prm = capi.params()
    _say("case edge: no neighbours")
    assert gpu_ctx.depth_map(rv, [], 40, 80, prm) is None
    _say("case edge: bad window")
    with pytest.raises(capi.MvsError):
        gpu_ctx.depth_map(rv, [s.views[0]], 40, 80, capi.params(taps=9))
    _say("case edge: image smaller than window")
    small = copy.copy(rv)
    small.gray = np.ascontiguousarray(rv.gray[:8, :8])
    with pytest.raises(capi.MvsError):
        gpu_ctx.depth_map(small, [s.views[0]], 40, 80, prm)
    _say("case edge: cancel")
    import ctypes
    flag = ctypes.c_int32(1)
    with pytest.raises(capi.MvsError) as e:
        gpu_ctx.depth_map(rv, [s.views[0]], 40, 80, prm, cancel=flag)
    assert e.value.code == -3
    _say("case edge: progress reports")
    nb = [s.views[j] for j in s.neighbours[ref][:4]]
    seen = []
    d0, _, _ = gpu_ctx.depth_map(rv, nb, 40, 80, prm)
    d1, _, _ = gpu_ctx.depth_map(rv, nb, 40, 80, prm,
                                 progress=lambda st, k, n: seen.append((st, k, n)) and False)
    total = 2 * prm.iterations
    assert seen == [("patchmatch", k, total) for k in range(total + 1)]
    assert np.array_equal(d0, d1)
    _say("case edge: cancel from the progress callback")
    with pytest.raises(capi.MvsError) as e:
        gpu_ctx.depth_map(rv, nb, 40, 80, prm, progress=lambda st, k, n: k >= 3)
    assert e.value.code == -3
    # the context stays usable after a cancelled call
    d2, _, _ = gpu_ctx.depth_map(rv, nb, 40, 80, prm)
    assert np.array_equal(d0, d2)
    _say("case edge: untextured reference")
    flat = copy.copy(rv)
    flat.gray = np.full_like(rv.gray, 90)
    d, n, c = gpu_ctx.depth_map(flat, [s.views[j] for j in s.neighbours[ref][:4]], 40, 80, prm)
    assert np.all(d == 0) and np.all(c == 2.0)


def test_shared_reciprocal_division_is_ieee(gpu_ctx):
    """The kernels' two-quotient division equals IEEE f32 `/` bit for bit."""
    _say("case selftest division")
    bad = gpu_ctx.selftest_division(1 << 30, seed=12345)
    _say(f"division mismatches: {bad} of {2 << 30}")
    assert bad == 0


@pytest.mark.parametrize("taps,stride,k", [(7, 1, 4), (7, 1, 8), (11, 1, 4), (11, 1, 8)])
def test_ncc_cost_parity_every_schedule(gpu_ctx, small_survey, taps, stride, k):
    """Each tap split (lanes per source) and chunk size gives the oracle's
    costs bit for bit: the slice-to-slice sums keep the reference's order."""
    s = small_survey
    ref = s.reference_index()
    nb = s.neighbours[ref][:k]
    rng = np.random.default_rng(200 + taps + k)
    xs, ys, hy = _random_hyps(s, ref, 3000, rng)
    prm = capi.params(taps=taps, stride=stride)
    want = orc.OracleDepth(s.views[ref], [s.views[j] for j in nb], 40, 80, taps=taps,
                           stride=stride).cost(xs, ys, hy)
    splits = {5: [1], 7: [1, 2], 11: [2, 4]}[taps]
    try:
        for split in splits:
            for gpc, per_cu in ((1, 16), (8, 16), (1, 0)):
                _say(f"case ncc_cost schedule taps={taps} K={k} split={split} gpc={gpc}")
                gpu_ctx.set_schedule(split=split, gpc=gpc, per_cu=per_cu, autotune=False)
                got = gpu_ctx.ncc_cost(s.views[ref], [s.views[j] for j in nb], 40, 80, prm,
                                       xs, ys, hy)
                assert _compare_costs(got, want, f"{taps}x{stride}/K{k}/split{split}") == 1.0
    finally:
        gpu_ctx.set_schedule(0, 0, 16, True)


def test_depth_map_identical_across_schedules(gpu_ctx, small_survey):
    s = small_survey
    ref = s.reference_index()
    rv, sv = s.views[ref], [s.views[j] for j in s.neighbours[ref][:8]]
    prm = capi.params(taps=7, stride=1, iterations=3)
    maps = []
    try:
        for split, gpc, per_cu in ((1, 4, 16), (2, 16, 16), (2, 1, 0), (4, 1, 16)):
            _say(f"case depth map schedule split={split} gpc={gpc} per_cu={per_cu}")
            gpu_ctx.set_schedule(split=split, gpc=gpc, per_cu=per_cu, autotune=False)
            maps.append(gpu_ctx.depth_map(rv, sv, 40, 80, prm))
    finally:
        gpu_ctx.set_schedule(0, 0, 16, True)
    for m in maps[1:]:
        for a, b in zip(maps[0], m):
            assert np.array_equal(a, b)
    od, _, _, _ = orc.OracleDepth(rv, sv, 40, 80, taps=7, stride=1).depth_map(iterations=3)
    assert np.array_equal(maps[0][0], od)


def test_full_depth_map_11x11(gpu_ctx, small_survey):
    s = small_survey
    ref = s.reference_index()
    rv, sv = s.views[ref], [s.views[j] for j in s.neighbours[ref][:4]]
    prm = capi.params(taps=11, stride=1, iterations=2)
    _say("case full depth map 11x11")
    gd, _, _ = gpu_ctx.depth_map(rv, sv, 40, 80, prm)
    od, _, _, _ = orc.OracleDepth(rv, sv, 40, 80, taps=11, stride=1).depth_map(iterations=2)
    both = (gd > 0) & (od > 0)
    within = float(np.mean(np.abs(gd[both] - od[both]) / od[both] <= 0.005))
    _say(f"11x11: within0.5%={within:.6f} identical={np.mean(gd == od):.6f}")
    assert within >= 0.99


def test_half_sweep_graph_replays_match_direct_launches(gpu_ctx, small_survey):
    """Half-sweeps run as cached hipGraphs give the same maps as direct
    launches, on every replay, with other work (a cost list, a depth map of
    another window) and host synchronisations between the replays."""
    s = small_survey
    ref = s.reference_index()
    srcs = [s.views[j] for j in s.neighbours[ref][:4]]
    prm = capi.params(taps=7, stride=1, iterations=3)
    other = capi.params(taps=5, stride=2, iterations=2)
    rng = np.random.default_rng(7)
    xs, ys, hy = _random_hyps(s, ref, 500, rng)
    _say("case graphs off")
    gpu_ctx.set_graphs(False)
    want, wn, wc = gpu_ctx.depth_map(s.views[ref], srcs, *s.spec.depth_range, prm)
    gpu_ctx.set_graphs(True)
    gpu_ctx.reset_stats()
    for rep in range(3):
        _say(f"case graph replay {rep}")
        d, n, c = gpu_ctx.depth_map(s.views[ref], srcs, *s.spec.depth_range, prm)
        assert np.array_equal(d, want) and np.array_equal(n, wn) and np.array_equal(c, wc)
        gpu_ctx.ncc_cost(s.views[ref], srcs, *s.spec.depth_range, prm, xs, ys, hy)
        gpu_ctx.depth_map(s.views[ref], srcs[:3], *s.spec.depth_range, other)
        gpu_ctx.synchronize()
    st = gpu_ctx.stats()
    _say(f"graph captures {st['graph_captures']}, replays {st['graph_replays']}")
    assert st["graph_replays"] >= 3 * 6
    assert st["graph_captures"] < st["graph_replays"]
