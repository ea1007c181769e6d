"""Pins of the oracle's f64 fixed-pair cost and its analytic gradient
(orc_pair_cost64, orc_pair_terms64), the checker of the GPU backward pass and
refinement.

Each check is chosen so that a plausible mistake fails it:
  - forward: with the pair of the two best sources, the f64 pair cost equals
    the reference restatement's f32 best-two cost (dense.rs:181-263) — a wrong
    pose, window or mean fails it;
  - the pair cost is the mean of the two single-source costs, and symmetric;
  - gradient vs central finite differences of the f64 cost (steps 1e-4 m in
    depth, 1e-5 in each normal component), skipping items whose bilinear cell
    signature changes within the step — a dropped or mis-signed term fails it;
  - the plane is invariant under scaling the normal, so n . dcost/dn = 0 — a
    missing d(dq)/dn term fails it;
  - along the true surface the depth gradient points toward the truth;
  - the Gauss-Newton matrix J^T J against J^T J formed from f64 central
    differences of the residual vectors themselves, symmetric, positive
    semi-definite, with the normal as a null direction.
"""
import numpy as np
import pytest

import oracle as orc
from conftest import survey_cached

H_DEPTH = 1e-4
H_NORMAL = 1e-5


def _ctx(name, k, taps, stride):
    s = survey_cached(name)
    ref = s.reference_index()
    nb = s.neighbours[ref][:k]
    v = s.views[ref]
    o = orc.OracleDepth(v, [s.views[j] for j in nb], *s.spec.depth_range, taps=taps,
                        stride=stride)
    return s, v, o


def _items(v, n, rng, margin=8, spread=0.01, tilt=0.05):
    xs = rng.integers(margin, v.w - margin, n).astype(np.int32)
    ys = rng.integers(margin, v.h - margin, n).astype(np.int32)
    hy = np.zeros((n, 4), np.float64)
    hy[:, 0] = v.depth_gt[ys, xs] * rng.uniform(1 - spread, 1 + spread, n)
    nrm = rng.normal([0, 0, -1], tilt, (n, 3))
    hy[:, 1:] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    return xs, ys, hy


@pytest.mark.parametrize("name,k,taps,stride", [("tiny", 4, 5, 2), ("small", 4, 7, 1)])
def test_pair_cost_with_best_pair_equals_reference_cost(name, k, taps, stride):
    s, v, o = _ctx(name, k, taps, stride)
    rng = np.random.default_rng(3)
    xs, ys, hy = _items(v, 200, rng)
    ref_cost = o.cost(xs, ys, hy.astype(np.float32))
    checked = 0
    for i in range(len(xs)):
        single = []
        for sidx in range(k):
            r = o.pair_cost64(xs[i], ys[i], hy[i].astype(np.float32).astype(np.float64), sidx,
                              sidx)
            if r is not None:
                single.append((1.0 - r[0], sidx))
        if len(single) < 2 or ref_cost[i] >= 2.0:
            continue
        single.sort(key=lambda t: (-t[0], t[1]))
        (n0, s0), (n1, s1) = single[0], single[1]
        c = o.pair_cost64(xs[i], ys[i], hy[i].astype(np.float32).astype(np.float64), s0, s1)
        assert c is not None
        # the top-two average does not depend on how near-ties are broken
        assert abs(c[0] - ref_cost[i]) < 3e-4, (i, c[0], ref_cost[i])
        assert abs(c[0] - (1.0 - 0.5 * (n0 + n1))) < 1e-12
        checked += 1
    assert checked > 150


def test_pair_cost_is_mean_of_singles_and_symmetric():
    s, v, o = _ctx("tiny", 4, 5, 2)
    rng = np.random.default_rng(5)
    xs, ys, hy = _items(v, 150, rng)
    n = 0
    for i in range(len(xs)):
        a = o.pair_cost64(xs[i], ys[i], hy[i], 0, 2)
        b = o.pair_cost64(xs[i], ys[i], hy[i], 2, 0)
        a0 = o.pair_cost64(xs[i], ys[i], hy[i], 0, 0)
        a2 = o.pair_cost64(xs[i], ys[i], hy[i], 2, 2)
        if a is None:
            assert b is None
            continue
        n += 1
        assert abs(a[0] - b[0]) < 1e-14
        assert abs(a[0] - 0.5 * (a0[0] + a2[0])) < 1e-14
        np.testing.assert_allclose(a[1], 0.5 * (a0[1] + a2[1]), rtol=1e-12, atol=1e-18)
    assert n > 50


def _fd_grad(o, x, y, hyp, s0, s1):
    """Central differences of the f64 cost; None where a step crosses a
    bilinear cell edge or leaves the defined region."""
    base = o.pair_cost64(x, y, hyp, s0, s1)
    if base is None:
        return None
    g = np.zeros(4)
    for a, h in enumerate((H_DEPTH, H_NORMAL, H_NORMAL, H_NORMAL)):
        hp, hm = hyp.copy(), hyp.copy()
        hp[a] += h
        hm[a] -= h
        p, m = o.pair_cost64(x, y, hp, s0, s1), o.pair_cost64(x, y, hm, s0, s1)
        if p is None or m is None or p[2] != base[2] or m[2] != base[2]:
            return None
        g[a] = (p[0] - m[0]) / (2 * h)
    return base, g


@pytest.mark.parametrize("name,k,taps,stride", [("tiny", 4, 5, 2), ("small", 4, 7, 1)])
def test_analytic_gradient_matches_finite_differences(name, k, taps, stride):
    s, v, o = _ctx(name, k, taps, stride)
    rng = np.random.default_rng(11)
    xs, ys, hy = _items(v, 400, rng, spread=0.02, tilt=0.1)
    checked = 0
    errs = []
    for i in range(len(xs)):
        r = _fd_grad(o, xs[i], ys[i], hy[i], 0, 1)
        if r is None:
            continue
        (c, g, _), gfd = r
        scale = np.array([1.0, 1.0 / hy[i, 0], 1.0 / hy[i, 0], 1.0 / hy[i, 0]])
        # compare in units where depth and normal steps move the samples alike
        err = np.abs(g - gfd) / scale
        ref = np.abs(g) / scale
        errs.append(err.max() / max(ref.max(), 1e-9))
        assert np.all(err <= 1e-5 * ref.max() + 1e-9), (i, g, gfd)
        checked += 1
    assert checked > 100, checked
    assert np.median(errs) < 1e-6


def test_gradient_orthogonal_to_normal():
    s, v, o = _ctx("small", 4, 7, 1)
    rng = np.random.default_rng(13)
    xs, ys, hy = _items(v, 400, rng, tilt=0.2)
    n = 0
    for i in range(len(xs)):
        r = o.pair_cost64(xs[i], ys[i], hy[i], 0, 1)
        if r is None:
            continue
        g = r[1]
        assert abs(np.dot(g[1:], hy[i, 1:])) <= 1e-9 * np.abs(g[1:]).max() + 1e-14
        # and the cost itself is unchanged by scaling the normal
        h2 = hy[i].copy()
        h2[1:] *= 1.7
        r2 = o.pair_cost64(xs[i], ys[i], h2, 0, 1)
        assert abs(r2[0] - r[0]) < 1e-12
        n += 1
    assert n > 100


def test_depth_gradient_points_to_truth():
    """A fronto-parallel plane 0.4% in front of or behind the true surface
    (the terrain is smooth at window scale) has a depth gradient whose sign
    points to the truth for the large majority of matched pixels."""
    s, v, o = _ctx("small", 4, 7, 1)
    rng = np.random.default_rng(17)
    n_ok = n_all = 0
    for sign in (-1.0, 1.0):
        xs, ys, hy = _items(v, 300, rng, spread=0.0, tilt=0.0)
        hy[:, 0] *= 1.0 + sign * 0.004
        for i in range(len(xs)):
            r = o.pair_cost64(xs[i], ys[i], hy[i], 0, 1)
            if r is None or r[0] > 0.3:
                continue
            n_all += 1
            n_ok += (r[1][0] * sign) > 0  # cost rises further from the truth
    assert n_all > 200
    assert n_ok / n_all > 0.8, (n_ok, n_all)


def test_undefined_cases():
    s, v, o = _ctx("tiny", 4, 5, 2)
    h = np.array([v.depth_gt[60, 80], 0.0, 0.0, -1.0])
    assert o.pair_cost64(80, 60, h, 0, 4) is None      # source out of range
    assert o.pair_cost64(80, 60, h, -1, 0) is None
    assert o.pair_cost64(1, 1, h, 0, 1) is None        # window off the image
    assert o.pair_cost64(80, 60, np.array([h[0], 1.0, 0.0, 0.0]), 0, 1) is None  # edge-on plane
    assert o.pair_cost64(80, 60, np.array([-h[0], 0.0, 0.0, -1.0]), 0, 1) is None  # behind


def _jtj_fd(o, x, y, h, s0, s1):
    """J^T J of the pair's residuals by f64 central differences of the unit
    centred sample windows (r_s = a/|a| - u_s, so J_r = -J_u); None where a
    step leaves the definedness region or crosses a bilinear cell edge."""
    steps = (H_DEPTH, H_NORMAL, H_NORMAL, H_NORMAL)
    H = np.zeros((4, 4))
    for s in (s0, s1):
        base = o.pair_unit64(x, y, h, s)
        if base is None:
            return None
        J = np.zeros((len(base[0]), 4))
        for a in range(4):
            hp, hm = h.copy(), h.copy()
            hp[a] += steps[a]
            hm[a] -= steps[a]
            p, m = o.pair_unit64(x, y, hp, s), o.pair_unit64(x, y, hm, s)
            if p is None or m is None or p[1] != base[1] or m[1] != base[1]:
                return None
            J[:, a] = (p[0] - m[0]) / (2 * steps[a])
        H += J.T @ J
    return H


def test_pair_jtj_matches_finite_differences_of_the_residuals():
    """The oracle's analytic Gauss-Newton matrix (the checker of both GPU
    refinement kernels' J^T J) against J^T J formed from f64 central
    differences of the residual vectors themselves: a dropped centring or
    projection term, or a wrong source pose, fails it. The matrix is also
    symmetric positive semi-definite with the normal as a null direction (the
    plane, hence every residual, is invariant under scaling n)."""
    s, v, o = _ctx("small", 4, 7, 1)
    rng = np.random.default_rng(17)
    xs, ys, hy = _items(v, 300, rng, spread=0.02, tilt=0.1)
    rels, nulls = [], []
    for i in range(len(xs)):
        t = o.pair_terms64(xs[i], ys[i], hy[i], 0, 1)
        if t is None:
            continue
        H = t[2]
        assert np.allclose(H, H.T, rtol=0, atol=1e-12 * np.abs(H).max())
        assert np.linalg.eigvalsh(H).min() >= -1e-9 * np.abs(H).max()
        n = np.concatenate([[0.0], hy[i, 1:]])
        nulls.append(np.linalg.norm(H @ n) / np.linalg.norm(H))
        fd = _jtj_fd(o, xs[i], ys[i], hy[i], 0, 1)
        if fd is None:
            continue
        # depth and normal entries weighed alike: normal columns scaled by 1/depth
        sc = np.diag([1.0, 1 / hy[i, 0], 1 / hy[i, 0], 1 / hy[i, 0]])
        e = np.abs(sc @ (H - fd) @ sc).max() / np.abs(sc @ fd @ sc).max()
        rels.append(e)
    rels = np.array(rels)
    assert len(rels) > 100
    print(f"J^T J vs finite differences: {len(rels)} items, median {np.median(rels):.3g}, "
          f"max {rels.max():.3g}; |H n|/|H| max {max(nulls):.3g}")
    assert np.median(rels) < 1e-6 and rels.max() < 1e-4
    assert max(nulls) < 1e-6

