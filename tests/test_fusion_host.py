"""Host side of the fusion (no GPU): the voxel-size rule and the PLY files.

The voxel rule (dense.rs:545-562) is checked against the oracle's restatement
and an independent numpy statement, on inputs where each branch decides
(ground sample distance vs the point budget, and the area clamp at 1 m^2).
The PLY writer is checked byte for byte against write_ply's layout
(ply.rs:11-40) and read back.
"""
import os

import numpy as np
import pytest

from aerial_mvs import capi

import oracle as orc


def _voxel_np(xs, ys, ranges, f, size, width, height, max_points):
    sx, sy = np.sort(xs), np.sort(ys)
    pct = lambda v, q: v[int((len(v) - 1) * q)]
    area = max((pct(sx, 0.99) - pct(sx, 0.01)) * (pct(sy, 0.99) - pct(sy, 0.01)), 1.0)
    dm = np.sort([(a + b) / 2 for a, b in ranges])
    f0 = f * size / max(width, height)
    gsd = dm[len(dm) // 2] / f0
    return max(gsd * 1.5, np.sqrt(area * 2.5 / max_points))


@pytest.mark.parametrize("case", range(6))
def test_fusion_voxel(case):
    rng = np.random.default_rng(40 + case)
    n = int(rng.integers(5, 400))
    spread = [200.0, 2000.0, 0.5, 50.0, 5000.0, 1.0][case]
    xs = rng.normal(0, spread, n)
    ys = rng.normal(0, spread * 0.7, n)
    ranges = [(rng.uniform(20, 60), rng.uniform(60, 200)) for _ in range(int(rng.integers(1, 30)))]
    f, size, w, h = 3000.0 + case, [640, 960, 1400][case % 3], 4000, 3000
    maxp = [2_000_000, 5_000_000, 12_000_000][case % 3]
    want = _voxel_np(xs, ys, ranges, f, size, w, h, maxp)
    assert capi.fusion_voxel(xs, ys, ranges, f, size, w, h, maxp) == pytest.approx(want, rel=1e-12)
    assert orc.fusion_voxel(xs, ys, ranges, f, size, w, h, maxp) == pytest.approx(want, rel=1e-12)


def test_ply_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(7)
    xyz = rng.normal(0, 50, (1000, 3)).astype(np.float32)
    rgb = rng.integers(0, 256, (1000, 3)).astype(np.uint8)
    p = str(tmp_path / "dense.ply")
    capi.write_ply(p, xyz, rgb, "origin 123.5 -45.25 7")
    assert not os.path.exists(str(tmp_path / "dense.ply.tmp"))
    raw = open(p, "rb").read()
    header = ("ply\nformat binary_little_endian 1.0\ncomment origin 123.5 -45.25 7\n"
              "element vertex 1000\nproperty float x\nproperty float y\nproperty float z\n"
              "property uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n")
    assert raw[:len(header)] == header.encode()
    body = np.frombuffer(raw[len(header):], np.uint8).reshape(1000, 15)
    assert np.array_equal(body[:, :12].copy().view("<f4"), xyz)
    assert np.array_equal(body[:, 12:], rgb)
    x2, c2 = capi.read_ply(p)
    assert np.array_equal(x2, xyz) and np.array_equal(c2, rgb)
    # an empty cloud is a valid file
    capi.write_ply(p, np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint8))
    x3, c3 = capi.read_ply(p)
    assert x3.shape == (0, 3)
