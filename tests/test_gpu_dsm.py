"""The surface model on the GPU (mvs_build_dsm) against the oracle's build_dsm
(recon/dsm.rs:128-290): the grid (x0, y0, cell, nx, ny), zmax and every cell
of the height grid bit-identical, NaN where the oracle has no data.

Clouds: the closed-form cases of the oracle's own tests (flat plane, the
second-highest rule, a spike, an enclosed pond), a terrain with buildings,
trees, dropouts and stray points, a wide sparse cloud whose cell edge comes
from the cell budget, the back-projected ground truth of the tiny survey, and
two builds in a row on one context, the second smaller (reused buffers).
"""
import sys

import numpy as np
import pytest

from aerial_mvs import capi

import oracle as orc

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _grid_points(x0, x1, y0, y1, step, z):
    y, x = np.mgrid[y0:y1:step, x0:x1:step]
    return np.c_[x.ravel(), y.ravel(), np.broadcast_to(z, x.shape).ravel()]


def _check(ctx, name, xyz, voxel, profile):
    _say(f"dsm case {name}: {len(xyz)} points, voxel {voxel}, profile {profile}")
    want = orc.build_dsm(xyz, voxel, orc.DSM_BUDGET[profile])
    info, h = ctx.build_dsm(xyz, voxel, profile)
    winfo, wh = want
    assert info == winfo, (name, info, winfo)
    assert h.shape == wh.shape
    nan = np.isnan(wh)
    assert np.array_equal(np.isnan(h), nan), (name, int((np.isnan(h) != nan).sum()))
    bad = (h != wh) & ~nan
    assert not bad.any(), (name, int(bad.sum()), np.argwhere(bad)[:5])
    return info, h


def _terrain(rng, n):
    """Rolling ground, flat-roofed buildings, trees (several returns per
    cell), dropout patches and a few stray points far above and outside."""
    x = rng.uniform(0, 300, n)
    y = rng.uniform(0, 200, n)
    z = 40 + 3 * np.sin(x / 37.0) + 2 * np.cos(y / 23.0)
    for bx, by, bw, bh, bz in ((50, 40, 30, 20, 12), (150, 120, 40, 35, 25), (220, 30, 15, 60, 8)):
        inb = (x > bx) & (x < bx + bw) & (y > by) & (y < by + bh)
        z[inb] += bz
    tree = rng.random(n) < 0.08
    z[tree] += rng.uniform(0.5, 9.0, tree.sum())
    keep = ~((np.hypot(x - 100, y - 150) < 9) | ((x > 260) & (y > 170)))
    p = np.c_[x, y, z][keep]
    stray = np.c_[rng.uniform(-200, 500, 300), rng.uniform(-200, 400, 300),
                  rng.uniform(0, 400, 300)]
    return np.concatenate([p, stray]).astype(np.float32)


def test_closed_form_clouds(gpu_ctx):
    rng = np.random.default_rng(1)
    p = np.c_[rng.uniform(0, 60, 40000), rng.uniform(0, 40, 40000), np.full(40000, 12.25)]
    _check(gpu_ctx, "plane", p, 0.3, 0)
    base = _grid_points(0.25, 30, 0.25, 20, 0.5, 0.0)
    for zs in ([1.0, 5.0, 3.0], [5.0, 5.0, 3.0], [2.0, 7.0], [4.5]):
        q = np.concatenate([base[:, :2] for _ in zs])
        z = np.repeat(np.asarray(zs), len(base))
        _check(gpu_ctx, f"second-highest {zs}", np.c_[q, z], 1.0 / 3.0, 0)
    p = _grid_points(0.125, 30, 0.125, 20, 0.25, 10.0)
    cellpts = (p[:, 0] > 15.0) & (p[:, 0] < 15.5) & (p[:, 1] > 10.0) & (p[:, 1] < 10.5)
    spike = np.array([[15.125, 10.125, 30.0], [15.375, 10.375, 31.0]])
    _check(gpu_ctx, "spike", np.concatenate([p[~cellpts], spike]), 1.0 / 3.0, 0)
    p = _grid_points(0.125, 40, 0.125, 40, 0.25, 3.0)
    hole = np.hypot(p[:, 0] - 20, p[:, 1] - 20) < 6.0
    _, h = _check(gpu_ctx, "pond", p[~hole], 1.0 / 3.0, 0)
    assert np.all(h == np.float32(3.0))


def test_terrain_with_buildings_and_trees(gpu_ctx):
    rng = np.random.default_rng(7)
    p = _terrain(rng, 400_000)
    for profile, voxel in ((0, 0.3), (1, 0.4)):
        info, h = _check(gpu_ctx, f"terrain profile {profile}", p, voxel, profile)
        assert 0.9 < np.isfinite(h).mean() < 1.0
    # 0.6 points per 0.2 m cell: every patch is an island below 25 m^2, the
    # grid ends empty and zmax stays at f32::MIN
    info, h = _check(gpu_ctx, "terrain, cells too fine", p, 0.137, 2)
    assert np.all(np.isnan(h)) and info["zmax"] == float(np.finfo(np.float32).min)


def test_cell_edge_from_budget(gpu_ctx):
    """2 km x 1.5 km with the Lite budget: sqrt(area / 3 M) > 1.5 voxel."""
    rng = np.random.default_rng(11)
    n = 1_200_000
    p = np.c_[rng.uniform(0, 2000, n), rng.uniform(0, 1500, n),
              rng.normal(100, 5, n)].astype(np.float32)
    info, _ = _check(gpu_ctx, "wide", p, 0.3, 0)
    assert info["cell"] > 0.45 and info["nx"] * info["ny"] <= 3_100_000


def test_extent_edge_cases(gpu_ctx):
    """The percentile extent (dsm.rs:140-148) over keys of both signs, heavy
    duplicates, a degenerate (one-column) extent and a NaN upper percentile."""
    rng = np.random.default_rng(17)
    n = 30_000
    p = np.c_[rng.uniform(-40, 40, n), rng.uniform(-1e-3, 1e-3, n) - 512.0, rng.normal(0, 1, n)]
    p[: n // 3, 0] = -0.0  # a third of the keys equal, around zero
    p[n // 3: n // 3 + 50, 0] = 0.0
    _check(gpu_ctx, "signs and duplicates", p, 0.3, 0)
    q = p.copy()
    q[:, 0] = 7.25  # every x equal: one column
    _check(gpu_ctx, "one column", q, 0.3, 0)
    r = p.copy()
    r[: n // 50, 0] = np.nan  # 2% NaN x: the 99.5th percentile is NaN, nx = 1
    info, _ = _check(gpu_ctx, "NaN percentile", r, 0.3, 0)
    assert info["nx"] == 1


def test_survey_ground_truth_cloud(gpu_ctx, tiny_survey):
    pts = []
    for v in tiny_survey.views:
        yy, xx = np.mgrid[0:v.h:2, 0:v.w:2]
        d = v.depth_gt[yy, xx]
        ok = d > 0
        q = np.stack([(xx + 0.5 - v.cx) / v.f * d, (yy + 0.5 - v.cy) / v.f * d, d], -1)[ok]
        pts.append(q @ np.asarray(v.R) + v.C)
    p = np.concatenate(pts).astype(np.float32)
    _check(gpu_ctx, "tiny survey ground truth", p, 0.25, 1)


def test_two_builds_in_a_row_and_refusal(gpu_ctx):
    rng = np.random.default_rng(3)
    big = _terrain(rng, 600_000)
    big[:5, 0] = np.nan  # NaN coordinates bin into cell index 0 (Rust's `as usize`)
    big[5:10, 1] = np.nan
    big[10:15, 2] = np.nan
    small = _terrain(rng, 50_000)[:, [1, 0, 2]] + np.float32(1000.0)
    _check(gpu_ctx, "big", big, 0.25, 1)
    _check(gpu_ctx, "small after big", small, 1.0, 1)
    with pytest.raises(capi.MvsError):
        gpu_ctx.build_dsm(small[:99], 0.3, 0)
