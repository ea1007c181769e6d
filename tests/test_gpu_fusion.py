"""GPU fusion (dense.rs:542-651) against the oracle's per-pixel emission.

Bars: the voxel set equals the oracle's (integer keys, bit-exact work); per
voxel the mean position within 2e-5 m of the float64 mean of the oracle's
points, and the colour equal to the reference's truncated mean; the cloud is
identical whatever order the views are fused in.
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


def _maps(gpu_ctx, s):
    spec = s.spec
    prm = capi.params(taps=spec.taps, stride=spec.stride, iterations=spec.iterations)
    centres = s.centres()
    jobs = []
    for i in range(len(s.views)):
        nb = capi.select_neighbours(centres, i, s.neighbours[i], spec.depth_range[0],
                                    spec.neighbours)
        jobs.append((i, nb, spec.depth_range))
    sv = capi.Survey(gpu_ctx, s.views)
    sv.depth_maps(jobs, prm)
    maps = {i: sv.download(i)[0] for i in range(len(s.views))}
    return sv, maps


def _oracle_cloud(s, maps, refs, nbrs, voxel, rule, colours):
    keys, xs, cols = [], [], []
    for r, nb in zip(refs, nbrs):
        k, xw, pix = orc.fuse_points(s.views[r], maps[r], [(s.views[j], maps[j]) for j in nb],
                                     rule=rule, voxel=voxel)
        keys.append(k)
        xs.append(xw)
        cols.append(colours[r].reshape(-1, 3)[pix])
    keys = np.concatenate(keys)
    xs = np.concatenate(xs)
    cols = np.concatenate(cols).astype(np.int64)
    order = np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))
    keys, xs, cols = keys[order], xs[order], cols[order]
    uniq, start, count = np.unique(keys, axis=0, return_index=True, return_counts=True)
    sums = np.add.reduceat(xs, start, axis=0)
    csum = np.add.reduceat(cols, start, axis=0)
    mean = sums / count[:, None]
    col = (csum.astype(np.float32) / count[:, None].astype(np.float32)).astype(np.uint8)
    return uniq, mean, col, len(xs)


@pytest.mark.parametrize("rule", [capi.MVS_CONS_REFERENCE, capi.MVS_CONS_REPROJ])
def test_fusion_matches_oracle(gpu_ctx, tiny_survey, rule):
    s = tiny_survey
    sv, maps = _maps(gpu_ctx, s)
    try:
        refs = list(range(len(s.views)))
        nbrs = [list(s.neighbours[i][:s.spec.neighbours + 1]) for i in refs]
        voxel = 0.05
        rgb = {}
        rng = np.random.default_rng(3)
        for i in refs:
            # a colour plane with structure in each channel
            g = s.views[i].gray.astype(np.int64)
            rgb[i] = np.stack([g, (g * 7 + 13) % 256, rng.integers(0, 256, g.shape)],
                              axis=-1).astype(np.uint8)
            sv.set_colour(i, rgb[i])
        _say(f"fusion rule {rule}: {len(refs)} views, voxel {voxel}")
        xyz, col = sv.fuse(refs, nbrs, voxel, rule=rule)
        keys, mean, ocol, npts = _oracle_cloud(s, maps, refs, nbrs, voxel, rule, rgb)
        _say(f"  voxels gpu {len(xyz)} oracle {len(keys)} from {npts} points")
        assert len(xyz) == len(keys) and len(keys) > 1000
        err = np.abs(xyz.astype(np.float64) - mean).max()
        _say(f"  max |position - oracle mean| {err:.3g} m")
        assert err <= 2e-5
        assert np.array_equal(col, ocol)
        # every point lies in its voxel
        assert np.array_equal(np.floor(mean / voxel).astype(np.int64), keys)
        # order independence: the views fused in reverse give the same cloud
        xyz2, col2 = sv.fuse(refs[::-1], nbrs[::-1], voxel, rule=rule)
        assert np.array_equal(xyz2, xyz) and np.array_equal(col2, col)
    finally:
        sv.close()


def test_fusion_table_regrows_to_oracle_cloud(gpu_ctx, tiny_survey):
    """A voxel far below the ground sampling distance gives about one voxel
    per confirmed pixel: more than the first table (2^16 slots for the tiny
    survey, 70% limit) holds. The pass that overflows stops inserting, counts
    its points, and the repeat with a table sized from them must give the
    oracle's cloud."""
    s = tiny_survey
    sv, maps = _maps(gpu_ctx, s)
    try:
        refs = list(range(len(s.views)))
        nbrs = [list(s.neighbours[i][:s.spec.neighbours + 1]) for i in refs]
        voxel = 0.01
        rgb = {i: np.repeat(s.views[i].gray[:, :, None], 3, axis=2) for i in refs}
        gpu_ctx.reset_stats()
        _say(f"fusion regrow: {len(refs)} views, voxel {voxel}")
        xyz, col = sv.fuse(refs, nbrs, voxel)
        passes = gpu_ctx.stats()["fuse_passes"]
        keys, mean, ocol, npts = _oracle_cloud(s, maps, refs, nbrs, voxel,
                                               capi.MVS_CONS_REFERENCE, rgb)
        _say(f"  voxels gpu {len(xyz)} oracle {len(keys)} from {npts} points, passes {passes}")
        assert len(keys) > 0.7 * 65536
        assert passes == 2
        assert len(xyz) == len(keys)
        assert np.abs(xyz.astype(np.float64) - mean).max() <= 2e-5
        assert np.array_equal(col, ocol)
        # a larger voxel afterwards starts from the small table again: one pass
        gpu_ctx.reset_stats()
        xyz_c, _ = sv.fuse(refs, nbrs, 0.2)
        assert gpu_ctx.stats()["fuse_passes"] == 1 and 0 < len(xyz_c) < len(xyz)
    finally:
        sv.close()


def test_fused_cloud_to_ply(gpu_ctx, tiny_survey, tmp_path):
    s = tiny_survey
    sv, maps = _maps(gpu_ctx, s)
    try:
        refs = list(range(len(s.views)))
        nbrs = [list(s.neighbours[i][:s.spec.neighbours + 1]) for i in refs]
        xyz, col = sv.fuse(refs, nbrs, 0.08)
        # views without a colour plane take their 8-bit grey
        assert np.all(col[:, 0] == col[:, 1]) and np.all(col[:, 1] == col[:, 2])
        p = str(tmp_path / "dense.ply")
        capi.write_ply(p, xyz, col, "synthetic")
        x2, c2 = capi.read_ply(p)
        assert np.array_equal(x2, xyz) and np.array_equal(c2, col)
        # a smaller voxel gives more points
        xyz_f, _ = sv.fuse(refs, nbrs, 0.04)
        assert len(xyz_f) > len(xyz)
    finally:
        sv.close()
