"""View preparation on the GPU (mvs_prepare_view) against the oracle: the
working-size RGB bytes and the grey values must be identical (integer box
sums, f64 bilinear weights and f32 grey in the reference's order), for every
branch of shrink_to (unchanged, box only, resize only, box + resize), odd
sizes, portrait photos and a padded row stride.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from aerial_mvs import capi

import oracle as orc

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _photo(h, w, seed):
    """A textured photo: smooth colour ramps plus noise (bilinear weights and
    rounding are exercised away from flat regions)."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    img = np.stack([(x * 255.0 / max(w - 1, 1)), (y * 255.0 / max(h - 1, 1)),
                    128 + 100 * np.sin(x / 7.0) * np.cos(y / 5.0)], -1)
    img = img + rng.normal(0, 12, img.shape)
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


@pytest.mark.parametrize("h,w,m", [
    (300, 400, 500),    # unchanged
    (600, 800, 400),    # box by 2, exact size
    (601, 803, 280),    # box by 2 (odd sizes truncate), then resize
    (500, 700, 400),    # 700 // 400 = 1: resize only
    (1200, 900, 250),   # portrait, box by 4, then resize
    (3000, 4000, 1400),  # a 12 Mpix photo to the High profile's working size
])
def test_prepare_view_identical_to_oracle(gpu_ctx, h, w, m):
    img = _photo(h, w, h * 7 + w)
    _say(f"prepare {w}x{h} -> max side {m}")
    rgb, grey = gpu_ctx.prepare_view(img, m)
    want_rgb, want_grey = orc.prepare_view(img, m)
    assert rgb.shape == want_rgb.shape
    assert np.array_equal(rgb, want_rgb)
    assert np.array_equal(grey.view(np.uint32), want_grey.view(np.uint32))


def test_prepare_view_with_padded_stride(gpu_ctx):
    h, w, pad = 210, 330, 17
    img = _photo(h, w, 5)
    buf = np.zeros((h, w * 3 + pad), np.uint8)
    buf[:, :w * 3] = img.reshape(h, w * 3)
    buf[:, w * 3:] = 255  # padding must not leak into the result
    ow, oh, _ = capi.prepared_size(w, h, 100)
    grey = np.zeros((oh, ow), np.float32)
    out = np.zeros((oh, ow, 3), np.uint8)
    gpu_ctx.check(capi.lib().mvs_prepare_view(gpu_ctx._h, buf.ctypes.data, w, h, w * 3 + pad, 100,
                                              grey.ctypes.data, out.ctypes.data))
    want_rgb, want_grey = orc.prepare_view(img, 100)
    assert np.array_equal(out, want_rgb) and np.array_equal(grey, want_grey)


def test_prepare_view_argument_errors(gpu_ctx):
    img = np.zeros((4, 4, 3), np.uint8)
    g = np.zeros(16, np.float32)
    o = np.zeros(48, np.uint8)
    L = capi.lib()
    assert L.mvs_prepare_view(gpu_ctx._h, img.ctypes.data, 4, 4, 11, 10, g.ctypes.data,
                              o.ctypes.data) == capi.MVS_ERR_ARG  # stride < 3 w
    assert L.mvs_prepare_view(gpu_ctx._h, None, 4, 4, 12, 10, g.ctypes.data,
                              o.ctypes.data) == capi.MVS_ERR_ARG
    assert L.mvs_prepare_view(gpu_ctx._h, img.ctypes.data, 4, 4, 12, 0, g.ctypes.data,
                              o.ctypes.data) == capi.MVS_ERR_ARG
