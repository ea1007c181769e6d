"""The oracle's view preparation (dense.rs:92-121, imgops.rs:89-171) pinned by
closed forms and independent computations, and the C ABI's host-side size
rule against it.

- shrink_to's sizes: hand-derived cases (unchanged, box only, box + resize,
  resize only, odd sizes, a long side that is a multiple of the target);
- box_down: numpy block sums over reshaped blocks with the same half-up
  integer rounding (a vectorised computation, not the oracle's loops);
- resize_bilinear: bilinear interpolation reproduces affine images, so a ramp
  in x (red), a ramp in y (green) and a constant (blue) must come out as the
  rounded ramps at the clamped source coordinates; identity at equal size; a
  2x2 -> 4x4 case worked by hand;
- grey: the 256 grey levels equal the grey table, pure primaries give
  0.299 / 0.587 / 0.114 to f32 rounding, and numpy f32 arithmetic in the
  reference's order agrees on random pixels;
- prepare_view = grey(resize(box(photo))) composed from the pieces.
"""
import numpy as np
import pytest

import oracle as orc
from aerial_mvs import capi


def test_shrink_sizes_by_hand():
    z = orc.shrink_size(1000, 800, 2000)
    assert (z["f"], z["w"], z["h"], z["scale"]) == (0, 1000, 800, 1.0)
    z = orc.shrink_size(4000, 3000, 2000)  # box by 2, exact size
    assert (z["f"], z["bw"], z["bh"], z["w"], z["h"], z["scale"]) == (2, 2000, 1500, 2000, 1500,
                                                                      0.5)
    z = orc.shrink_size(4000, 3000, 1400)  # box by 2, then 2000 -> 1400
    assert (z["f"], z["bw"], z["w"], z["h"]) == (2, 2000, 1400, 1050)
    z = orc.shrink_size(3000, 2000, 2000)  # 3000 // 2000 = 1: resize only
    assert (z["f"], z["w"], z["h"]) == (0, 2000, 1333)  # round(2000 * 2000 / 3000)
    z = orc.shrink_size(4001, 3001, 2000)  # box truncates: 2000 x 1500
    assert (z["f"], z["bw"], z["bh"], z["w"], z["h"]) == (2, 2000, 1500, 2000, 1500)
    z = orc.shrink_size(6000, 4000, 1400)  # 6000 // 1400 = 4: 1500 x 1000 -> 1400 x 933
    assert (z["f"], z["bw"], z["bh"], z["w"], z["h"]) == (4, 1500, 1000, 1400, 933)
    z = orc.shrink_size(3000, 4000, 1400)  # portrait
    assert (z["f"], z["bw"], z["bh"], z["w"], z["h"]) == (2, 1500, 2000, 1050, 1400)


def test_abi_prepared_size_matches_oracle():
    rng = np.random.default_rng(4)
    for _ in range(400):
        w, h = (int(v) for v in rng.integers(1, 9000, 2))
        m = int(rng.integers(1, 4000))
        z = orc.shrink_size(w, h, m)
        ow, oh, sc = capi.prepared_size(w, h, m)
        assert (ow, oh) == (z["w"], z["h"]) and sc == z["scale"], (w, h, m)


@pytest.mark.parametrize("f,h,w", [(2, 10, 14), (3, 11, 17), (4, 9, 9), (5, 23, 31)])
def test_box_down_is_the_rounded_block_mean(f, h, w):
    img = np.random.default_rng(f).integers(0, 256, (h, w, 3), dtype=np.uint8)
    oh, ow = h // f, w // f
    blocks = img[:oh * f, :ow * f].astype(np.uint32).reshape(oh, f, ow, f, 3).sum(axis=(1, 3))
    area = f * f
    want = ((blocks + area // 2) // area).astype(np.uint8)
    assert np.array_equal(orc.box_down(img, f), want)


def test_box_down_of_a_constant_is_the_constant():
    img = np.full((12, 18, 3), 201, np.uint8)
    assert np.all(orc.box_down(img, 3) == 201)


@pytest.mark.parametrize("sh,sw,h,w", [(40, 60, 28, 41), (33, 47, 33, 20), (50, 50, 17, 23),
                                       (21, 30, 40, 57)])
def test_resize_reproduces_ramps(sh, sw, h, w):
    a, b = 2.0, 7.0
    img = np.zeros((sh, sw, 3), np.uint8)
    img[..., 0] = (a * np.arange(sw) + b)[None, :].astype(np.uint8)   # ramp in x
    img[..., 1] = (3.0 * np.arange(sh) + 1.0)[:, None].astype(np.uint8)  # ramp in y
    img[..., 2] = 99
    out = orc.resize_bilinear(img, w, h)
    fx = np.clip((np.arange(w) + 0.5) * (sw / w) - 0.5, 0.0, sw - 1)
    fy = np.clip((np.arange(h) + 0.5) * (sh / h) - 0.5, 0.0, sh - 1)
    rx = np.floor(a * fx + b + 0.5).astype(np.uint8)
    ry = np.floor(3.0 * fy + 1.0 + 0.5).astype(np.uint8)
    # rounding ties can go either way after the f64 weights; allow exactly
    # those, and nothing else
    dx = out[..., 0].astype(int) - rx[None, :].astype(int)
    dy = out[..., 1].astype(int) - ry[:, None].astype(int)
    assert np.all(np.abs(dx) <= 1) and np.all(np.abs(dy) <= 1)
    tie_x = np.abs((a * fx + b) % 1.0 - 0.5) < 1e-9
    tie_y = np.abs((3.0 * fy + 1.0) % 1.0 - 0.5) < 1e-9
    assert np.all(dx[:, ~tie_x] == 0) and np.all(dy[~tie_y, :] == 0)
    assert np.all(out[..., 2] == 99)


def test_resize_to_the_same_size_is_the_identity():
    img = np.random.default_rng(2).integers(0, 256, (13, 19, 3), dtype=np.uint8)
    assert np.array_equal(orc.resize_bilinear(img, 19, 13), img)


def test_resize_2x2_to_4x4_by_hand():
    img = np.zeros((2, 2, 3), np.uint8)
    img[..., 0] = [[0, 100], [200, 40]]
    out = orc.resize_bilinear(img, 4, 4)[..., 0]
    # source coordinates of the output centres: 0 (clamped), 0.25, 0.75, 1 (clamped)
    t = [0.0, 0.25, 0.75, 1.0]
    want = np.zeros((4, 4))
    for y in range(4):
        for x in range(4):
            ax, ay = t[x], t[y]
            top = 0 * (1 - ax) + 100 * ax
            bot = 200 * (1 - ax) + 40 * ax
            want[y, x] = top * (1 - ay) + bot * ay
    assert np.array_equal(out, np.floor(want + 0.5).astype(np.uint8))


def test_grey_levels_and_primaries():
    lut = orc.gray_lut()
    for g in range(256):
        img = np.full((1, 1, 3), g, np.uint8)
        _, grey = orc.prepare_view(img, 10)
        assert grey[0, 0] == lut[g]
    for c, k in enumerate((0.299, 0.587, 0.114)):
        img = np.zeros((1, 1, 3), np.uint8)
        img[0, 0, c] = 255
        _, grey = orc.prepare_view(img, 10)
        assert abs(float(grey[0, 0]) - k) <= 2e-7


def test_grey_matches_numpy_f32_in_reference_order():
    img = np.random.default_rng(8).integers(0, 256, (37, 29, 3), dtype=np.uint8)
    _, grey = orc.prepare_view(img, 100)
    p = img.astype(np.float32)
    v = np.float32(0.299) * p[..., 0] + np.float32(0.587) * p[..., 1]
    v = v + np.float32(0.114) * p[..., 2]
    assert np.array_equal(grey, v / np.float32(255.0))


@pytest.mark.parametrize("h,w,m", [(61, 83, 40), (120, 90, 50), (64, 96, 48), (50, 75, 60),
                                   (30, 20, 100)])
def test_prepare_view_composes_the_pieces(h, w, m):
    img = np.random.default_rng(h + w).integers(0, 256, (h, w, 3), dtype=np.uint8)
    rgb, grey = orc.prepare_view(img, m)
    z = orc.shrink_size(w, h, m)
    b = orc.box_down(img, z["f"]) if z["f"] >= 2 else img
    want = b if (z["w"], z["h"]) == (b.shape[1], b.shape[0]) else orc.resize_bilinear(
        b, z["w"], z["h"])
    assert np.array_equal(rgb, want)
    _, grey2 = orc.prepare_view(want, max(z["w"], z["h"]))  # unchanged size: grey only
    assert np.array_equal(grey, grey2)
