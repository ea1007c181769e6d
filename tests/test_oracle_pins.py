"""Pins of the CPU oracle (oracle/dense_oracle.c) against what the mathematics and
the reference's own tests fix. The reference has no tests for recon/dense.rs, so
each check here is chosen so that a plausible transcription mistake (a wrong
shift or constant, a swapped operand, a dropped term, a wrong index) fails it.
"""
import copy
from decimal import Decimal, getcontext

import numpy as np
import pytest

import oracle as orc
from ref64 import cost64

MASK64 = (1 << 64) - 1


# --------------------------------------------------------------------------
# dense.rs:273-279 — xorshift64 (13, 7, 17)
# --------------------------------------------------------------------------

def _step(state):
    _, s = orc.rnd_stream(state, 1)
    return s


def _apply(cols, v):
    out = 0
    k = 0
    while v:
        if v & 1:
            out ^= cols[k]
        v >>= 1
        k += 1
    return out


def _compose(a, b):  # a after b
    return [_apply(a, c) for c in b]


def _mat_pow(cols, e):
    result = [1 << k for k in range(64)]
    base = cols
    while e:
        if e & 1:
            result = _compose(base, result)
        base = _compose(base, base)
        e >>= 1
    return result


def test_xorshift_is_a_full_period_gf2_map():
    """The oracle's state update, read back as a 64x64 matrix over GF(2), must
    be linear and have order exactly 2^64 - 1 (full period), which holds for
    Marsaglia's (13, 7, 17) triple and fails for most mistyped shifts."""
    cols = [_step(1 << k) for k in range(64)]
    rng = np.random.default_rng(5)
    for _ in range(20):
        a = int(rng.integers(1, 2**63)) * 2 + 1
        b = int(rng.integers(1, 2**63))
        assert _step(a ^ b) == _step(a) ^ _step(b)
        assert _step(a) == _apply(cols, a)
    order = (1 << 64) - 1
    ident = [1 << k for k in range(64)]
    assert _mat_pow(cols, order) == ident
    for p in (3, 5, 17, 257, 641, 65537, 6700417):
        assert order % p == 0
        assert _mat_pow(cols, order // p) != ident


def test_rnd_output_is_top_24_bits():
    draws, _ = orc.rnd_stream(0x1234567, 50)
    s = 0x1234567
    for d in draws:
        s = _step(s)
        assert d == np.float32((s >> 40) / float(1 << 24))
        assert 0.0 <= d < 1.0


def test_seed_constants():
    # dense.rs:374 uses the Fibonacci-hashing constant floor(2^64 / phi)
    getcontext().prec = 60
    phi = (1 + Decimal(5).sqrt()) / 2
    golden = int(Decimal(2**64) / phi)
    assert orc.seed_init(1) == golden | 1
    assert orc.seed_init(0) == 1
    assert orc.seed_init(3) == (3 * golden & MASK64) | 1
    # dense.rs:446-448 uses the xorshift64* multiplier 2685821657736338717
    m = 2685821657736338717
    assert orc.seed_refine(0, 0, 0) == 1
    assert orc.seed_refine(1, 0, 0) == (256 * m & MASK64) | 1
    assert orc.seed_refine(5, 2, 1) == (((5 << 8) ^ 5) * m & MASK64) | 1
    # (i, it, parity) -> seed is injective on the sweep's domain (before |1)
    seen = set()
    for i in range(0, 2000, 7):
        for it in range(6):
            for par in range(2):
                seen.add(orc.seed_refine(i, it, par) >> 1)
    assert len(seen) == len(range(0, 2000, 7)) * 12


# --------------------------------------------------------------------------
# dense.rs:281-297 — random_normal
# --------------------------------------------------------------------------

def test_random_normal_invariants():
    rng = np.random.default_rng(2)
    for _ in range(200):
        ray = rng.uniform(-0.6, 0.6, 2).astype(np.float32)
        around = rng.normal(size=3).astype(np.float32)
        seed = int(rng.integers(1, 2**62)) | 1
        n, s_after = orc.random_normal(seed, ray, around, 0.3)
        assert abs(np.linalg.norm(n.astype(np.float64)) - 1.0) < 1e-6
        assert n[0] * ray[0] + n[1] * ray[1] + n[2] <= 1e-7   # faces the camera
        _, s3 = orc.rnd_stream(seed, 3)                        # exactly three draws
        assert s_after == s3
    # spread 0: the (possibly flipped) normalised base vector
    n, _ = orc.random_normal(7, [0.1, -0.2], [0.0, 0.0, -2.0], 0.0)
    assert np.allclose(n, [0, 0, -1], atol=1e-7)
    n, _ = orc.random_normal(7, [0.1, -0.2], [0.0, 0.0, 2.0], 0.0)
    assert np.allclose(n, [0, 0, -1], atol=1e-7)


# --------------------------------------------------------------------------
# geometry.rs:26-45 — camera model
# --------------------------------------------------------------------------

def test_distortion_round_trip_reference_test():
    # geometry.rs:306-316 and :335-340
    k = (3000.0, 2000.0, 1500.0, -0.02, 0.005)
    u = orc.distort(*k, 0.31, -0.22)
    p = orc.undistort(*k, u[0], u[1])
    assert np.hypot(p[0] - 0.31, p[1] + 0.22) < 1e-9


def test_pinhole_closed_form():
    for x, y in [(0.0, 0.0), (0.31, -0.22), (-0.5, 0.4)]:
        u = orc.distort(900.0, 512.0, 384.0, 0.0, 0.0, x, y)
        assert u[0] == 900.0 * x + 512.0 and u[1] == 900.0 * y + 384.0
        p = orc.undistort(900.0, 512.0, 384.0, 0.0, 0.0, u[0], u[1])
        assert abs(p[0] - x) < 1e-15 and abs(p[1] - y) < 1e-15


def test_distort_matches_textbook_radial_model():
    k1, k2 = -0.03, 0.004
    rng = np.random.default_rng(1)
    for x, y in rng.uniform(-0.5, 0.5, (50, 2)):
        r2 = x * x + y * y
        fac = 1 + k1 * r2 + k2 * r2 ** 2
        u = orc.distort(1234.5, 600.25, 401.75, k1, k2, x, y)
        assert abs(u[0] - (1234.5 * fac * x + 600.25)) < 1e-9
        assert abs(u[1] - (1234.5 * fac * y + 401.75)) < 1e-9


# --------------------------------------------------------------------------
# dense.rs:108 — grey
# --------------------------------------------------------------------------

def test_gray_formula():
    lut = orc.gray_lut()
    assert lut[0] == 0.0
    assert abs(lut[255] - 1.0) <= 2 * np.spacing(np.float32(1.0))
    assert np.all(np.diff(lut) > 0)
    f32 = np.float32
    for g in (1, 17, 128, 200, 254):
        v = (f32(0.299) * f32(g) + f32(0.587) * f32(g)) + f32(0.114) * f32(g)
        assert lut[g] == v / f32(255.0)
    assert abs(orc.gray_rgb(10, 200, 30) - (0.299 * 10 + 0.587 * 200 + 0.114 * 30) / 255) < 1e-6


# --------------------------------------------------------------------------
# dense.rs:299-309 — relative pose
# --------------------------------------------------------------------------

def _pose(rng):
    from aerial_mvs.survey import from_angles
    return from_angles(rng.uniform(-50, 50, 3), rng.uniform(0, 360), rng.uniform(-90, -40),
                       rng.uniform(-5, 5))


def test_relative_identity_and_composition():
    rng = np.random.default_rng(3)
    R, Cc = _pose(rng)
    Rr, t = orc.relative(R, Cc, R, Cc)
    assert np.allclose(Rr, np.eye(3), atol=1e-7) and np.all(t == 0)
    for _ in range(10):
        Ra, Ca = _pose(rng)
        Rb, Cb = _pose(rng)
        Rr, t = orc.relative(Ra, Ca, Rb, Cb)
        X = rng.uniform(-30, 30, 3)
        xa = Ra @ (X - Ca)
        xb = Rb @ (X - Cb)
        assert np.allclose(Rr.astype(np.float64) @ xa + t, xb, atol=1e-4)


# --------------------------------------------------------------------------
# dense.rs:69-85 — bilinear grey; dense.rs:266-271 — mean_sd
# --------------------------------------------------------------------------

def test_gray_at_bilinear():
    from scipy.ndimage import map_coordinates
    rng = np.random.default_rng(4)
    g = rng.uniform(0, 1, (12, 17)).astype(np.float32)
    for y in range(12):
        for x in range(17):
            v = orc.gray_at(g, x + 0.5, y + 0.5)
            if x + 1 < 17 and y + 1 < 12:
                assert v == g[y, x]
            else:
                assert v is None  # x0 + 1 >= w / y0 + 1 >= h
    for _ in range(300):
        x, y = rng.uniform(0.5, 16.49), rng.uniform(0.5, 11.49)
        v = orc.gray_at(g, x, y)
        ref = map_coordinates(g.astype(np.float64), [[y - 0.5], [x - 0.5]], order=1)[0]
        assert abs(v - ref) < 1e-6
    assert orc.gray_at(g, 0.49, 3.0) is None
    assert orc.gray_at(g, 3.0, 0.4) is None
    assert orc.gray_at(g, 16.5, 3.0) is None
    assert orc.gray_at(g, float("nan"), 3.0) is None


def test_mean_sd():
    rng = np.random.default_rng(6)
    for n in (25, 49, 121):
        v = rng.uniform(0, 1, n).astype(np.float32)
        m, s = orc.mean_sd(v)
        assert abs(m - v.astype(np.float64).mean()) < 1e-6
        assert abs(s - v.astype(np.float64).std()) < 1e-6
    m, s = orc.mean_sd(np.full(25, 0.5, np.float32))   # exact sums
    assert m == 0.5 and s == 0.0
    m, s = orc.mean_sd(np.full(25, 0.37, np.float32))  # rounding only
    assert s < 1e-6


# --------------------------------------------------------------------------
# dense.rs:324-334 — neighbour choice
# --------------------------------------------------------------------------

def test_select_neighbours():
    centres = np.array([[0, 0, 60], [0.5, 0, 60], [10, 0, 60], [0, 0.79, 60], [0, 0.81, 60],
                        [30, 0, 60]], np.float64)
    # 0.02 * max(40, 1) = 0.8 m: views 1 (0.5 m) and 3 (0.79 m) are dropped
    assert orc.select_neighbours(centres, 0, [1, 2, 3, 4, 5], 40.0, 8) == [2, 4, 5]
    assert orc.select_neighbours(centres, 0, [1, 2, 3, 4, 5], 40.0, 2) == [2, 4]
    # range below 1 uses 1: limit 0.02 m
    assert orc.select_neighbours(centres, 0, [1, 3], 0.2, 8) == [1, 3]
    assert orc.select_neighbours(centres, 0, [], 40.0, 8) == []


# --------------------------------------------------------------------------
# dense.rs:181-263 — the cost
# --------------------------------------------------------------------------

def _ctx(survey, ref, srcs, taps=5, stride=2, views=None):
    views = views or survey.views
    r = survey.spec.depth_range
    return orc.OracleDepth(views[ref], [views[s] for s in srcs], r[0], r[1], taps=taps,
                           stride=stride)


def _true_hyp(survey, ref, x, y):
    v = survey.views[ref]
    return np.array([v.depth_gt[y, x], 0.0, 0.0, -1.0], np.float32)


def test_cost_is_zero_against_itself(tiny_survey):
    s = tiny_survey
    ref = s.reference_index()
    for taps, stride in ((5, 2), (7, 1)):
        ctx = _ctx(s, ref, [ref, ref], taps, stride)
        rng = np.random.default_rng(7)
        xs = rng.integers(10, s.views[ref].w - 10, 300)
        ys = rng.integers(10, s.views[ref].h - 10, 300)
        hy = np.zeros((300, 4), np.float32)
        hy[:, 0] = rng.uniform(40, 80, 300)
        nrm = rng.normal([0, 0, -1], 0.2, (300, 3))
        hy[:, 1:] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
        c = ctx.cost(xs, ys, hy)
        ok = c < 2.0
        assert ok.mean() > 0.9
        assert np.abs(c[ok]).max() < 2e-5


def test_cost_invariant_to_affine_intensity(tiny_survey):
    s = tiny_survey
    ref = s.reference_index()
    nb = s.neighbours[ref][:2]
    views = list(s.views)
    lut = orc.gray_lut()
    for j in nb:
        v = copy.copy(views[j])
        v.gray_f32 = (0.6 * lut[v.gray] + 0.2).astype(np.float32)
        views[j] = v
    base = _ctx(s, ref, nb)
    scaled = _ctx(s, ref, nb, views=views)
    rng = np.random.default_rng(8)
    xs = rng.integers(8, s.views[ref].w - 8, 400)
    ys = rng.integers(8, s.views[ref].h - 8, 400)
    hy = np.stack([_true_hyp(s, ref, x, y) for x, y in zip(xs, ys)])
    hy[:, 0] *= rng.uniform(0.99, 1.01, 400).astype(np.float32)
    a, b = base.cost(xs, ys, hy), scaled.cost(xs, ys, hy)
    assert np.all((a < 2) == (b < 2))
    assert np.abs(a - b).max() < 2e-5


def test_cost_undefined_cases(tiny_survey):
    s = tiny_survey
    ref = s.reference_index()
    nb = s.neighbours[ref][:4]
    ctx = _ctx(s, ref, nb)
    v = s.views[ref]
    hy = _true_hyp(s, ref, 50, 40)
    # window off the image (half footprint 4)
    assert ctx.cost([3], [40], [hy])[0] == 2.0
    assert ctx.cost([v.w - 4], [40], [hy])[0] == 2.0
    # negative depth: every window depth <= 0
    bad = hy.copy()
    bad[0] = -10.0
    assert ctx.cost([50], [40], [bad])[0] == 2.0
    # plane containing the centre ray: n . (ray, 1) = 0 at the centre tap
    rays = ctx.rays()
    r = rays[40, 50].astype(np.float64)
    n = np.array([1.0, 0.0, -r[0]])
    n /= np.linalg.norm(n)
    assert ctx.cost([50], [40], [[60.0, *n]])[0] == 2.0
    # constant reference image
    flat = copy.copy(v)
    flat.gray = np.full_like(v.gray, 120)
    views = list(s.views)
    views[ref] = flat
    fctx = _ctx(s, ref, nb, views=views)
    assert np.all(fctx.cost([50, 60], [40, 50], [hy, hy]) == 2.0)


def test_cost_best_two_by_brute_force(tiny_survey):
    s = tiny_survey
    ref = s.reference_index()
    nb = s.neighbours[ref][:5]
    rng = np.random.default_rng(9)
    xs = rng.integers(8, s.views[ref].w - 8, 200)
    ys = rng.integers(8, s.views[ref].h - 8, 200)
    hy = np.stack([_true_hyp(s, ref, x, y) for x, y in zip(xs, ys)])
    hy[:, 0] *= rng.uniform(0.97, 1.03, 200).astype(np.float32)
    singles = np.stack([_ctx(s, ref, [j]).cost(xs, ys, hy) for j in nb], axis=1)
    for k in (1, 2, 3, 5):
        got = _ctx(s, ref, nb[:k]).cost(xs, ys, hy)
        srt = np.sort(singles[:, :k], axis=1)
        want = srt[:, 0] if k == 1 else (srt[:, 0] + srt[:, 1]) * np.float32(0.5)
        assert np.array_equal(got, want.astype(np.float32))


@pytest.mark.parametrize("taps,stride", [(5, 2), (7, 1)])
def test_cost_matches_float64_statement(small_survey, taps, stride):
    s = small_survey
    ref = s.reference_index()
    nb = s.neighbours[ref][:4]
    ctx = _ctx(s, ref, nb, taps, stride)
    lut = orc.gray_lut()
    rays = ctx.rays()
    srcs = []
    for i, j in enumerate(nb):
        R, t = ctx.relpose(i)
        v = s.views[j]
        srcs.append((lut[v.gray], R, t, v.f, v.cx, v.cy))
    rng = np.random.default_rng(10)
    n = 150
    xs = rng.integers(0, s.views[ref].w, n)
    ys = rng.integers(0, s.views[ref].h, n)
    hy = np.zeros((n, 4), np.float32)
    hy[:, 0] = rng.uniform(45, 75, n)
    nrm = rng.normal([0, 0, -1], 0.3, (n, 3))
    hy[:, 1:] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    got = ctx.cost(xs, ys, hy)
    g = lut[s.views[ref].gray]
    want = np.array([cost64(g, rays, srcs, x, y, h, taps, stride) for x, y, h in zip(xs, ys, hy)])
    both = (got < 2) & (want < 2)
    assert both.mean() > 0.5
    assert np.abs(got[both] - want[both]).max() < 2e-4
    # undefined exactly where the f64 statement is undefined, except for
    # borderline samples (within f32 rounding of an image edge)
    assert np.mean((got == 2) == (want == 2)) > 0.97


# --------------------------------------------------------------------------
# dense.rs:365-477 — PatchMatch invariants
# --------------------------------------------------------------------------

def test_patchmatch_invariants(tiny_survey):
    s = tiny_survey
    ref = s.reference_index()
    nb = s.neighbours[ref][:4]
    ctx = _ctx(s, ref, nb)
    st, tex = ctx.init(nthreads=3)
    dmin, dmax, imin, imax = ctx.limits()
    inv = 1.0 / st[..., 0].astype(np.float64)
    assert inv.min() >= imin * (1 - 1e-6) and inv.max() <= imax * (1 + 1e-6)
    nrm = st[..., 1:4].astype(np.float64)
    assert np.allclose(np.linalg.norm(nrm, axis=-1), 1.0, atol=1e-6)
    assert np.all(st[..., 4][tex == 0] == 2.0)
    # border of half footprint 4 is never textured
    assert not tex[:4].any() and not tex[-4:].any() and not tex[:, :4].any()
    yy, xx = np.mgrid[0:s.views[ref].h, 0:s.views[ref].w]
    prev = st.copy()
    for it in range(2):
        for par in range(2):
            other = ((xx + yy) % 2) != par
            ev = ctx.half_sweep(st, tex, it, par, nthreads=4)
            assert 0 < ev <= 10 * int(((~other) & (tex == 1)).sum())
            assert np.all(st[..., 4] <= prev[..., 4])            # never worse
            assert np.array_equal(st[other], prev[other])        # other colour frozen
            assert np.array_equal(st[tex == 0], prev[tex == 0])  # untextured frozen
            prev = st.copy()


def test_patchmatch_thread_count_independent(tiny_survey):
    s = tiny_survey
    ref = s.reference_index()
    nb = s.neighbours[ref][:4]
    d1, st1, _, e1 = _ctx(s, ref, nb).depth_map(iterations=2, nthreads=1)
    d7, st7, _, e7 = _ctx(s, ref, nb).depth_map(iterations=2, nthreads=7)
    assert np.array_equal(st1, st7) and e1 == e7 and np.array_equal(d1, d7)


def test_patchmatch_recovers_ground_truth(small_survey):
    s = small_survey
    ref = s.reference_index()
    nb = s.neighbours[ref][:4]
    ctx = _ctx(s, ref, nb, taps=7, stride=1)
    d, st, tex, _ = ctx.depth_map(iterations=4, nthreads=8)
    gt = s.views[ref].depth_gt
    valid = d > 0
    assert valid.mean() > 0.6
    rel = np.abs(d[valid] - gt[valid]) / gt[valid]
    assert np.mean(rel < 0.005) > 0.9


# --------------------------------------------------------------------------
# dense.rs:585-615 — consistency
# --------------------------------------------------------------------------

def _consistency64(ref, depth, nbrs):
    h, w = depth.shape
    agree = np.full((h, w), -1, np.int64)
    ys, xs = np.nonzero(depth > 0)
    d = depth[ys, xs].astype(np.float64)
    q = np.stack([(xs + 0.5 - ref.cx) / ref.f, (ys + 0.5 - ref.cy) / ref.f], 1)
    Xc = np.stack([q[:, 0] * d, q[:, 1] * d, d], 1)
    Xw = Xc @ np.asarray(ref.R) + ref.C
    cnt = np.zeros(len(d), np.int64)
    for v, dn in nbrs:
        xs_ = (Xw - v.C) @ np.asarray(v.R).T
        z = xs_[:, 2]
        with np.errstate(divide="ignore", invalid="ignore"):
            u = np.floor(v.f * xs_[:, 0] / z + v.cx)
            vv = np.floor(v.f * xs_[:, 1] / z + v.cy)
        ok = (z > 0) & (u >= 0) & (vv >= 0) & (u < v.w) & (vv < v.h)
        dd = np.zeros(len(d))
        dd[ok] = dn[vv[ok].astype(int), u[ok].astype(int)]
        cnt += ok & (dd > 0) & (np.abs((dd - z) / np.where(ok, z, 1)) < 0.01)
    agree[ys, xs] = cnt
    return agree


def test_consistency_brute_force(tiny_survey):
    s = tiny_survey
    views = s.views
    depth = {i: views[i].depth_gt.astype(np.float32) for i in range(len(views))}
    rng = np.random.default_rng(11)
    ref = s.reference_index()
    dref = depth[ref] * rng.uniform(0.995, 1.005, depth[ref].shape).astype(np.float32)
    dref[rng.uniform(size=dref.shape) < 0.1] = 0.0
    nbrs = [(views[j], depth[j]) for j in s.neighbours[ref][:5]]
    agree, n = orc.consistency(views[ref], dref, nbrs)
    want = _consistency64(views[ref], dref, nbrs)
    # identical except for samples within f64 rounding of a pixel edge or the
    # 1% threshold
    assert np.mean(agree == want) > 0.999
    assert n == int(((agree >= 2) & (dref > 0)).sum())
    # a neighbour identical to the reference always agrees
    a2, _ = orc.consistency(views[ref], dref, [(views[ref], dref)])
    assert np.all(a2[dref > 0] == 1) and np.all(a2[dref <= 0] == -1)


def _reproj64(ref, depth, nbrs):
    """float64 numpy statement of the reprojection rule (pinhole cameras):
    ref pixel -> world -> neighbour pixel (floor) -> the neighbour's depth
    there, back-projected from that pixel's centre -> reference image; agree
    when within 1 px of the ref pixel centre."""
    h, w = depth.shape
    agree = np.full((h, w), -1, np.int64)
    ys, xs = np.nonzero(depth > 0)
    d = depth[ys, xs].astype(np.float64)
    Rr = np.asarray(ref.R)
    Xc = np.stack([(xs + 0.5 - ref.cx) / ref.f * d, (ys + 0.5 - ref.cy) / ref.f * d, d], 1)
    Xw = Xc @ Rr + ref.C
    cnt = np.zeros(len(d), np.int64)
    for v, dn in nbrs:
        Rn = np.asarray(v.R)
        xs_ = (Xw - v.C) @ Rn.T
        z = xs_[:, 2]
        with np.errstate(divide="ignore", invalid="ignore"):
            u = np.floor(v.f * xs_[:, 0] / z + v.cx)
            vv = np.floor(v.f * xs_[:, 1] / z + v.cy)
        ok = (z > 0) & (u >= 0) & (vv >= 0) & (u < v.w) & (vv < v.h)
        dd = np.zeros(len(d))
        dd[ok] = dn[vv[ok].astype(int), u[ok].astype(int)]
        ok &= dd > 0
        Xn = np.stack([(u + 0.5 - v.cx) / v.f * dd, (vv + 0.5 - v.cy) / v.f * dd, dd], 1) @ Rn + v.C
        xr = (Xn - ref.C) @ Rr.T
        with np.errstate(divide="ignore", invalid="ignore"):
            ur = ref.f * xr[:, 0] / xr[:, 2] + ref.cx
            vr = ref.f * xr[:, 1] / xr[:, 2] + ref.cy
        err2 = (ur - (xs + 0.5)) ** 2 + (vr - (ys + 0.5)) ** 2
        cnt += ok & (xr[:, 2] > 0) & (err2 <= 1.0)
    agree[ys, xs] = cnt
    return agree


def test_consistency_reprojection_rule(tiny_survey):
    s = tiny_survey
    views = s.views
    ref = s.reference_index()
    rng = np.random.default_rng(12)
    dref = views[ref].depth_gt.astype(np.float32) * rng.uniform(
        0.99, 1.01, views[ref].depth_gt.shape).astype(np.float32)
    dref[rng.uniform(size=dref.shape) < 0.1] = 0.0
    nbrs = []
    for j in s.neighbours[ref][:5]:
        dj = views[j].depth_gt.astype(np.float32) * rng.uniform(
            0.99, 1.01, views[j].depth_gt.shape).astype(np.float32)
        nbrs.append((views[j], dj))
    agree, n = orc.consistency(views[ref], dref, nbrs, rule=1, need=3)
    want = _reproj64(views[ref], dref, nbrs)
    # identical except within f64 rounding of a pixel edge or the 1 px circle
    assert np.mean(agree == want) > 0.999
    assert n == int(((agree >= 3) & (dref > 0)).sum())
    # depth noise of 1% moves reprojections: some but not all agree
    frac = np.mean(agree[dref > 0] >= 3)
    assert 0.05 < frac < 0.999
    # a neighbour identical to the reference reprojects exactly
    a2, _ = orc.consistency(views[ref], dref, [(views[ref], dref)], rule=1, need=1)
    assert np.all(a2[dref > 0] == 1) and np.all(a2[dref <= 0] == -1)
    # exact ground-truth maps agree almost everywhere they overlap
    gt = [(views[j], views[j].depth_gt.astype(np.float32)) for j in s.neighbours[ref][:5]]
    a3, _ = orc.consistency(views[ref], views[ref].depth_gt.astype(np.float32), gt, rule=1,
                            need=1)
    w3 = _reproj64(views[ref], views[ref].depth_gt.astype(np.float32), gt)
    assert np.mean(a3 == w3) > 0.999
    assert np.mean(a3 >= 1) > 0.9
