"""The oracle's DSM (recon/dsm.rs) pinned by the reference's own unit tests
(dsm.rs:395-445, restated) and by closed forms: a flat plane, the extent
percentiles, the second-highest rule, despiking, enclosed holes.
"""
import numpy as np

import oracle as orc

BUDGET_LITE = orc.DSM_BUDGET[0]


def test_reference_bilinear_height_and_gaps():
    """dsm.rs:395-418 (bilinear_height_and_gaps)."""
    info = dict(x0=0.0, y0=0.0, cell=1.0, nx=3, ny=3, zmax=12.0)
    h = np.full((3, 3), 10.0, np.float32)
    h[1, 1] = 12.0  # centre cell (1, 1) at world (1.5, 1.5)
    assert abs(orc.dsm_height(info, h, 1.5, 1.5) - 12.0) < 1e-5
    assert abs(orc.dsm_height(info, h, 1.0, 1.5) - 11.0) < 1e-5
    h[0, 0] = np.nan
    assert orc.dsm_height(info, h, 0.5, 0.5) is not None  # falls back to valid neighbours
    assert orc.dsm_height(info, h, -3.0, 0.0) is None


def test_reference_islands_removed_and_enclosed_holes_filled():
    """dsm.rs:420-445 (islands_removed_and_enclosed_holes_filled)."""
    nx = ny = 10
    h = np.full((ny, nx), np.nan, np.float32)
    h[2:8, 2:8] = 5.0  # a 6x6 block with a 2x2 hole, and a lone cell far away
    h[4:6, 4:6] = np.nan
    h[0, 0] = 9.0
    h = orc.remove_islands(h, 3)
    assert np.isnan(h[0, 0])
    h = orc.fill_enclosed(h)
    assert h[4, 4] == 5.0
    assert np.isnan(h[9, 9])  # outside stays empty


def _grid_points(x0, x1, y0, y1, step, z):
    # dyadic coordinates: cell indices floor((x - x0) / cell) are exact
    y, x = np.mgrid[y0:y1:step, x0:x1:step]
    return np.c_[x.ravel(), y.ravel(), np.broadcast_to(z, x.shape).ravel()]


def test_flat_plane_and_extent():
    rng = np.random.default_rng(1)
    p = np.c_[rng.uniform(0, 60, 40000), rng.uniform(0, 40, 40000), np.full(40000, 12.25)]
    info, h = orc.build_dsm(p, 0.3, BUDGET_LITE)
    xs = np.sort(p[:, 0].astype(np.float32))
    ys = np.sort(p[:, 1].astype(np.float32))
    n = len(p)
    xa, xb = float(xs[int((n - 1) * 0.005)]), float(xs[int((n - 1) * 0.995)])
    ya, yb = float(ys[int((n - 1) * 0.005)]), float(ys[int((n - 1) * 0.995)])
    cell = max(0.3 * 1.5, np.sqrt((xb - xa) * (yb - ya) / BUDGET_LITE))
    assert (info["x0"], info["y0"], info["cell"]) == (xa, ya, cell)
    assert info["nx"] == int(np.ceil((xb - xa) / cell)) + 1
    assert info["ny"] == int(np.ceil((yb - ya) / cell)) + 1
    assert np.all(h[np.isfinite(h)] == np.float32(12.25)) and info["zmax"] == np.float32(12.25)
    assert np.isfinite(h).mean() > 0.99


def test_second_highest_rule():
    """Each cell keeps its second-highest height (with multiplicity), or the
    only one: uniform fields survive the filters unchanged."""
    base = _grid_points(0.25, 30, 0.25, 20, 0.5, 0.0)  # one point per 0.5 m cell
    for zs, want in (([1.0, 5.0, 3.0], 3.0), ([5.0, 5.0, 3.0], 5.0), ([2.0, 7.0], 2.0),
                     ([4.5], 4.5)):
        p = np.concatenate([base[:, :2] for _ in zs])
        z = np.repeat(np.asarray(zs), len(base))
        info, h = orc.build_dsm(np.c_[p, z], 1.0 / 3.0, BUDGET_LITE)  # cell 0.5
        assert info["cell"] == 0.5
        v = h[np.isfinite(h)]
        assert len(v) > 0.9 * h.size and np.all(v == np.float32(want)), zs


def test_spike_is_removed():
    p = _grid_points(0.125, 30, 0.125, 20, 0.25, 10.0)  # four points per 0.5 m cell
    cellpts = (p[:, 0] > 15.0) & (p[:, 0] < 15.5) & (p[:, 1] > 10.0) & (p[:, 1] < 10.5)
    assert cellpts.sum() == 4
    spike = np.array([[15.125, 10.125, 30.0], [15.375, 10.375, 31.0]])  # a thin spike
    info, h = orc.build_dsm(np.concatenate([p[~cellpts], spike]), 1.0 / 3.0, BUDGET_LITE)
    assert np.all(h[np.isfinite(h)] == np.float32(10.0))


def test_enclosed_hole_is_filled_and_small_cloud_refused():
    p = _grid_points(0.125, 40, 0.125, 40, 0.25, 3.0)
    hole = np.hypot(p[:, 0] - 20, p[:, 1] - 20) < 6.0  # a pond 12 m across
    info, h = orc.build_dsm(p[~hole], 1.0 / 3.0, BUDGET_LITE)
    assert np.all(np.isfinite(h)) and np.all(h == np.float32(3.0))
    assert orc.build_dsm(p[:99], 0.3, BUDGET_LITE) is None


def test_nan_coordinates_land_in_the_first_row_and_column():
    """Rust's `as usize` turns a NaN cell index into 0 (dsm.rs:155-160): a
    point with NaN x bins like one on the grid's left edge (x = x0)."""
    rng = np.random.default_rng(5)
    p = np.c_[rng.uniform(0, 30, 20000), rng.uniform(0, 20, 20000), np.full(20000, 2.0)]
    p[:400, 0] = 0.0  # blocks of equal x at both ends hold the x percentiles
    p[400:800, 0] = 30.0  # in place whichever end the three extra points sort to
    info, _ = orc.build_dsm(p, 0.3, BUDGET_LITE)
    assert info["x0"] == 0.0
    # a strip of higher points, three per cell, down column 0: it survives
    # despiking (2.9 m step) and the 3x3 median (six values at the edge)
    ys = np.linspace(5.0, 15.0, 66)
    extra = np.c_[np.full(66, np.nan), ys, np.full(66, 4.9)]
    edge = extra.copy()
    edge[:, 0] = info["x0"]
    a = orc.build_dsm(np.concatenate([p, extra]), 0.3, BUDGET_LITE)
    b = orc.build_dsm(np.concatenate([p, edge]), 0.3, BUDGET_LITE)
    assert a[0] == b[0] and np.array_equal(a[1], b[1], equal_nan=True)
    assert np.sum(a[1][:, 0] == np.float32(4.9)) >= 15  # the strip is there
