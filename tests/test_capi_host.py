"""Host-side checks of the C-ABI library (no GPU compute): exports, the host logic
that mirrors the reference (profile table, neighbour choice, `.depth` files) and
the loud failure when no gfx950 device is present."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from aerial_mvs import capi

import oracle as orc

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "aerial_mvs.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(mvs_[a-z0-9_]+)\s*\(", txt))
    return {n for n in names if not n.endswith("_fn")}


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (mvs_[a-z0-9_]+)", out))
    declared = _declared()
    assert declared, "no declarations parsed"
    assert declared <= exported, sorted(declared - exported)
    assert set(capi.EXPORTS) == declared
    L = capi.lib()
    for name in declared:
        assert getattr(L, name) is not None


def _header_arity():
    """{function: argument count} of the header's declarations."""
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for name, args in re.findall(r"\b(mvs_[a-z0-9_]+)\s*\(([^;]*?)\)\s*;", txt, flags=re.S):
        if name.endswith("_fn"):
            continue
        a = args.strip()
        out[name] = 0 if a in ("", "void") else a.count(",") + 1
    return out


def test_integration_binding_covers_the_header():
    """The Rust extern block a maintainer would add declares every function of
    the C ABI with the header's argument count."""
    md = open(os.path.join(REPO, "INTEGRATION.md")).read()
    block = md[md.index('extern "C" {'):]
    block = block[:block.index("\n}\n")]
    rust = {}
    for name, args in re.findall(r"pub fn (mvs_[a-z0-9_]+)\s*\(([^;]*?)\)\s*(?:->[^;]*)?;",
                                 block, flags=re.S):
        a = args.strip().rstrip(",")
        rust[name] = 0 if not a else a.count(",") + 1
    want = _header_arity()
    assert set(want) == _declared()
    assert rust == want


def test_default_params_are_the_reference_constants():
    p = capi.params()
    q = capi.MvsParams()
    capi.lib().mvs_default_params(C.byref(q))
    assert (q.taps, q.stride, q.iterations) == (5, 2, 4)           # dense.rs:170, :41-46
    assert q.texture_min == np.float32(0.008) and q.cost_max == np.float32(0.45)
    assert p.taps == 5


def test_profile_table():
    # dense.rs:32-55
    assert capi.dense_params_for_profile(0) == dict(size=640, iterations=3, neighbours=3,
                                                    max_points=2_000_000)
    assert capi.dense_params_for_profile(1) == dict(size=960, iterations=4, neighbours=4,
                                                    max_points=5_000_000)
    assert capi.dense_params_for_profile(2) == dict(size=1400, iterations=5, neighbours=5,
                                                    max_points=12_000_000)
    with pytest.raises(capi.MvsError):
        capi.dense_params_for_profile(3)


def test_select_neighbours_matches_oracle():
    rng = np.random.default_rng(0)
    for _ in range(200):
        n = int(rng.integers(2, 20))
        centres = rng.normal(0, 1.0, (n, 3)) * rng.choice([0.3, 3.0, 30.0])
        ci = int(rng.integers(0, n))
        nb = [int(j) for j in rng.permutation(n) if j != ci][: int(rng.integers(0, 9))]
        r0 = float(rng.uniform(0.1, 80))
        take = int(rng.integers(1, 9))
        assert capi.select_neighbours(centres, ci, nb, r0, take) == \
            orc.select_neighbours(centres, ci, nb, r0, take)


def test_depth_file_round_trip(tmp_path):
    d = np.random.default_rng(1).uniform(0, 80, (7, 11)).astype(np.float32)
    d[2, 3] = 0.0
    p = str(tmp_path / "abc_960.depth")
    capi.save_depth(p, d)
    assert not os.path.exists(str(tmp_path / "abc_960.tmp"))   # tmp + rename
    raw = open(p, "rb").read()
    assert len(raw) == 8 + 4 * d.size
    assert np.frombuffer(raw[:8], "<u4").tolist() == [11, 7]
    assert np.array_equal(np.frombuffer(raw[8:], "<f4").reshape(7, 11), d)
    assert np.array_equal(capi.load_depth(p), d)
    # load_depth (dense.rs:494-508): wrong length -> None
    open(p, "wb").write(raw[:-4])
    assert capi.load_depth(p) is None
    assert capi.load_depth(str(tmp_path / "missing.depth")) is None


def test_view_file_round_trip(tmp_path):
    """`.view` cache (dense.rs:92-154): u32 w, u32 h, w*h f32 grey, w*h*3 RGB;
    any other length is a damaged file."""
    rng = np.random.default_rng(2)
    rgb = rng.integers(0, 256, (5, 9, 3)).astype(np.uint8)
    f = rgb.astype(np.float32)
    gray = ((np.float32(0.299) * f[..., 0] + np.float32(0.587) * f[..., 1])
            + np.float32(0.114) * f[..., 2]) / np.float32(255.0)
    p = str(tmp_path / "IMG_0001_960.view")
    capi.save_view(p, gray, rgb)
    assert not os.path.exists(str(tmp_path / "IMG_0001_960.tmp"))
    raw = open(p, "rb").read()
    assert len(raw) == 8 + 7 * gray.size
    assert np.frombuffer(raw[:8], "<u4").tolist() == [9, 5]
    assert np.array_equal(np.frombuffer(raw[8:8 + 4 * gray.size], "<f4").reshape(5, 9), gray)
    assert raw[8 + 4 * gray.size:] == rgb.tobytes()
    g2, c2 = capi.load_view(p)
    assert np.array_equal(g2, gray) and np.array_equal(c2, rgb)
    open(p, "wb").write(raw + b"\0")
    assert capi.load_view(p) is None
    assert capi.load_view(str(tmp_path / "missing.view")) is None
    # working-size intrinsics: K scaled by w / full width (dense.rs:139-145)
    assert capi.view_intrinsics(3000.0, 2000.0, 1500.0, 4000, 960) == (720.0, 480.0, 360.0)


def test_no_device_fails_loudly():
    """Without a gfx950 device there is no fallback: creation fails with
    MVS_ERR_DEVICE and a message."""
    try:
        ctx = capi.Context(0)
    except capi.MvsError as e:
        assert e.code == -2 and "device" in str(e).lower()
        return
    ctx.close()
    pytest.skip("a gfx950 device is present")
