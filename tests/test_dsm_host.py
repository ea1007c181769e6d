"""Host side of the surface model over the C ABI (no GPU compute): the
``dsm.bin`` file (Dsm::save / Dsm::load, recon/dsm.rs:27-75) and the
bilinear height lookup (Dsm::height, dsm.rs:82-119), checked against the
reference's own unit test (dsm.rs:400-424, restated) and against the oracle's
lookup, which that test pins.
"""
import struct

import numpy as np

from aerial_mvs import capi

import oracle as orc


def _ref_grid():
    info = dict(x0=0.0, y0=0.0, cell=1.0, nx=3, ny=3, zmax=12.0)
    h = np.full((3, 3), 10.0, np.float32)
    h[1, 1] = 12.0
    return info, h


def test_reference_bilinear_height_and_gaps(tmp_path):
    """dsm.rs:400-424, through a save/load round trip as the reference does."""
    info, h = _ref_grid()
    p = str(tmp_path / "dsm.bin")
    capi.save_dsm(p, info, h)
    info2, h2 = capi.load_dsm(p)
    assert info2 == info and np.array_equal(h2, h)
    assert abs(capi.dsm_height(info2, h2, 1.5, 1.5) - 12.0) < 1e-5
    assert abs(capi.dsm_height(info2, h2, 1.0, 1.5) - 11.0) < 1e-5
    h2[0, 0] = np.nan
    assert capi.dsm_height(info2, h2, 0.5, 0.5) is not None
    assert capi.dsm_height(info2, h2, -3.0, 0.0) is None


def test_file_layout_and_nan(tmp_path):
    """"DSM1", x0 y0 cell as f64, nx ny as u64, zmax f32, then ny*nx f32, all
    little-endian; NaN (no data) survives the round trip bit for bit."""
    rng = np.random.default_rng(4)
    h = rng.normal(50, 5, (7, 11)).astype(np.float32)
    h[rng.random(h.shape) < 0.3] = np.nan
    info = dict(x0=512000.25, y0=-3.5, cell=0.4375, nx=11, ny=7, zmax=float(np.nanmax(h)))
    p = tmp_path / "dsm.bin"
    capi.save_dsm(str(p), info, h)
    raw = p.read_bytes()
    assert raw[:4] == b"DSM1" and len(raw) == 48 + 4 * h.size
    assert struct.unpack("<dddQQf", raw[4:48]) == (512000.25, -3.5, 0.4375, 11, 7,
                                                   np.float32(info["zmax"]))
    assert raw[48:] == h.astype("<f4").tobytes()
    info2, h2 = capi.load_dsm(str(p))
    assert info2 == dict(info, zmax=float(np.float32(info["zmax"])))
    assert h2.tobytes() == h.tobytes()


def test_damaged_and_truncated_files_are_refused(tmp_path):
    info, h = _ref_grid()
    p = tmp_path / "dsm.bin"
    capi.save_dsm(str(p), info, h)
    raw = p.read_bytes()
    for bad in (raw[:47], b"DSM2" + raw[4:], raw[:-1], b""):
        p.write_bytes(bad)
        assert capi.load_dsm(str(p)) is None
    assert capi.load_dsm(str(tmp_path / "missing.bin")) is None


def test_height_matches_oracle_everywhere():
    """Random lookups over a grid with gaps, inside, on the edges and outside."""
    rng = np.random.default_rng(9)
    ny, nx = 23, 31
    h = rng.normal(20, 3, (ny, nx)).astype(np.float32)
    h[rng.random(h.shape) < 0.25] = np.nan
    h[5:9, 5:9] = np.nan  # a block whose centre has no valid neighbour
    info = dict(x0=100.0, y0=200.0, cell=0.75, nx=nx, ny=ny, zmax=float(np.nanmax(h)))
    xs = rng.uniform(99.0, 100.0 + 0.75 * nx + 1.0, 3000)
    ys = rng.uniform(199.0, 200.0 + 0.75 * ny + 1.0, 3000)
    xs[:40] = 100.0 + 0.75 * np.arange(40) % (0.75 * nx)  # cell edges
    n_none = 0
    for x, y in zip(xs, ys):
        got = capi.dsm_height(info, h, x, y)
        want = orc.dsm_height(info, h, x, y)
        if want is None:
            n_none += 1
            assert got is None, (x, y)
        else:
            assert got is not None and np.float32(got) == np.float32(want), (x, y, got, want)
    assert 0 < n_none < len(xs)
