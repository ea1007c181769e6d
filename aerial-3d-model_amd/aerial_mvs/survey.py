"""Seeded synthetic aerial surveys (workload for tests and the benchmark).

Scene: fractal heightfield (smooth value-noise fBm) textured with band-limited
value noise, rendered by the C generator ``libsynth.so`` with pinhole
projection, per-view Gaussian sensor noise, and the FP64 camera-z depth of every
pixel (ground truth). Cameras follow the reference's convention and its
``Pose::from_angles`` (crates/stitch-core/src/recon/geometry.rs:76-93); the
lawnmower layout with alternating yaw follows its synthetic BA survey
(recon/ba.rs:411-421). Neighbours are nearest by camera position (GPS), best
first, at most 8 as in recon/sparse.rs:536.
"""
import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import build as _build

_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")

_lib = None


def _synth():
    global _lib
    if _lib is None:
        path = os.environ.get("AERIAL_MVS_SYNTH_LIB") or _build.LIB_SYNTH
        if not os.path.exists(path):
            raise RuntimeError("libsynth.so not built")
        L = C.CDLL(path)
        L.synth_scene_create.argtypes = [C.c_double] * 6 + [C.c_int32, C.c_double, C.c_double,
                                                            C.c_uint64, C.c_double, C.c_int32,
                                                            C.c_int32]
        L.synth_scene_create.restype = C.c_void_p
        L.synth_scene_add_boxes.argtypes = [C.c_void_p, _f64p, C.c_int32]
        L.synth_scene_destroy.argtypes = [C.c_void_p]
        L.synth_scene_info.argtypes = [C.c_void_p, _f64p]
        L.synth_height.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.synth_height.restype = C.c_double
        L.synth_render.argtypes = [C.c_void_p, C.c_int32, C.c_int32] + [C.c_double] * 5 + [
            _f64p, _f64p, C.c_uint64, C.c_double, _u8p, C.c_void_p, C.c_int32]
        _lib = L
    return _lib


def rot_axis(axis, angle):
    """geometry.rs:104-108 (Rodrigues about a unit axis)."""
    a = np.asarray(axis, np.float64)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]], np.float64)
    return np.eye(3) + K * np.sin(angle) + K @ K * (1.0 - np.cos(angle))


def from_angles(c, yaw_deg, pitch_deg, roll_deg):
    """Pose::from_angles (geometry.rs:76-93): returns (R world->camera, C)."""
    sy, cy = np.sin(np.radians(yaw_deg)), np.cos(np.radians(yaw_deg))
    sp, cp = np.sin(np.radians(pitch_deg)), np.cos(np.radians(pitch_deg))
    heading = np.array([sy, cy, 0.0])
    up = np.array([0.0, 0.0, 1.0])
    x = np.array([cy, -sy, 0.0])
    z = heading * cp + up * sp
    y = np.cross(z, x)
    r_wc = np.stack([x, y, z], axis=1)
    if abs(roll_deg) > 1e-9:
        r_wc = r_wc @ rot_axis([0.0, 0.0, 1.0], np.radians(roll_deg))
    return np.ascontiguousarray(r_wc.T), np.asarray(c, np.float64)


@dataclass
class View:
    gray: np.ndarray                 # (h, w) uint8
    f: float
    cx: float
    cy: float
    k1: float
    k2: float
    R: np.ndarray                    # world -> camera
    C: np.ndarray
    depth_gt: Optional[np.ndarray] = None   # (h, w) float64 camera z, 0 = no hit

    @property
    def w(self):
        return self.gray.shape[1]

    @property
    def h(self):
        return self.gray.shape[0]


@dataclass
class SurveySpec:
    name: str
    width: int
    height: int
    focal: float
    altitude: float
    lines: int
    per_line: int
    forward_overlap: float = 0.75
    side_overlap: float = 0.65
    patch: tuple = (200.0, 150.0)
    relief: float = 15.0
    grid: float = 0.1
    base_lambda: float = 50.0
    octaves: int = 6
    persistence: float = 0.5
    tex_lambda0: float = 0.2
    tex_octaves: int = 5
    pos_sigma: float = 0.05
    ang_sigma_deg: float = 0.05
    noise_dn: float = 1.0
    depth_range: tuple = (40.0, 80.0)
    neighbours: int = 4
    taps: int = 7
    stride: int = 1
    iterations: int = 4
    k1: float = 0.0
    k2: float = 0.0
    seed: int = 0xA3D00000


# The reference-window shape: 5 taps at stride 2 (dense.rs:170).
SPECS = {
    # c1: 16 nadir views 1024x768, 4 neighbours, 7x7, 4 iterations
    "c1": SurveySpec("c1", 1024, 768, 900.0, 60.0, 4, 4, neighbours=4, iterations=4,
                     seed=0xA3D00001),
    # c2: 77 nadir views 2000x1500, 8 neighbours, 7x7, 6 iterations;
    # focal scaled with the width to keep c1's field of view; 60 m
    # altitude, depth range [40, 80], the 7 x 11 block over the scene centre.
    "c2": SurveySpec("c2", 2000, 1500, 900.0 * 2000 / 1024, 60.0, 7, 11, patch=(500.0, 350.0),
                     neighbours=8, iterations=6, seed=0xA3D00002),
    # c4: c2's survey shape (its own scene seed) with an 11x11 window
    # (N = 121), refined by Gauss-Newton steps
    "c4": SurveySpec("c4", 2000, 1500, 900.0 * 2000 / 1024, 60.0, 7, 11, patch=(500.0, 350.0),
                     neighbours=8, taps=11, iterations=6, seed=0xA3D00004),
    # small surveys for tests: the c1 camera geometry with a cropped sensor
    "tiny": SurveySpec("tiny", 160, 120, 900.0, 60.0, 3, 3, patch=(60.0, 50.0), neighbours=4,
                       taps=5, stride=2, iterations=3, seed=0xA3D0000A),
    "small": SurveySpec("small", 320, 240, 900.0, 60.0, 3, 4, patch=(80.0, 60.0), neighbours=4,
                        taps=7, stride=1, iterations=4, seed=0xA3D0000B),
    # the same with radial distortion (geometry.rs:26-32) in every camera
    "small_dist": SurveySpec("small_dist", 320, 240, 900.0, 60.0, 3, 4, patch=(80.0, 60.0),
                             neighbours=4, taps=7, stride=1, iterations=4, k1=-0.02, k2=0.005,
                             seed=0xA3D0000E),
}


@dataclass
class Survey:
    spec: SurveySpec
    views: List[View]
    neighbours: List[List[int]]          # up to 8 per view, nearest first
    ranges: List[tuple]
    ground: float
    extra: dict = field(default_factory=dict)

    def centres(self):
        return np.array([v.C for v in self.views], np.float64)

    def reference_index(self):
        """The view nearest the survey centre (lowest index on ties)."""
        c = self.centres()[:, :2]
        m = c.mean(axis=0)
        d = np.hypot(c[:, 0] - m[0], c[:, 1] - m[1])
        return int(np.argmin(d))


def _layout(spec, rng):
    fw = spec.width / spec.focal * spec.altitude      # across track (image x)
    fh = spec.height / spec.focal * spec.altitude     # along track (image y)
    along = (1.0 - spec.forward_overlap) * fh
    across = (1.0 - spec.side_overlap) * fw
    cx0, cy0 = spec.patch[0] / 2, spec.patch[1] / 2
    poses = []
    for r in range(spec.lines):
        yaw = 90.0 if r % 2 == 0 else 270.0
        cols = range(spec.per_line) if r % 2 == 0 else range(spec.per_line - 1, -1, -1)
        for cidx in cols:
            x = cx0 + (cidx - (spec.per_line - 1) / 2) * along
            y = cy0 + (r - (spec.lines - 1) / 2) * across
            poses.append((x, y, yaw))
    out = []
    for (x, y, yaw) in poses:
        dp = rng.normal(0.0, spec.pos_sigma, 3)
        da = rng.normal(0.0, spec.ang_sigma_deg, 3)
        out.append(((x + dp[0], y + dp[1], dp[2]), yaw + da[0], -90.0 + da[1], da[2]))
    return out


def nearest_neighbours(centres, k=8):
    out = []
    for i in range(len(centres)):
        d = np.hypot(centres[:, 0] - centres[i, 0], centres[:, 1] - centres[i, 1])
        order = sorted((float(d[j]), j) for j in range(len(centres)) if j != i)
        out.append([j for _, j in order[:k]])
    return out


def make_survey(spec, nthreads=None, with_depth=True, views=None):
    """Render a survey. `views` (optional) restricts which views get pixels
    (others keep an empty image), to save time when only some are needed."""
    if isinstance(spec, str):
        spec = SPECS[spec]
    nthreads = nthreads or min(16, os.cpu_count() or 1)
    L = _synth()
    margin = spec.altitude
    scene = L.synth_scene_create(-margin, -margin, spec.patch[0] + 2 * margin,
                                 spec.patch[1] + 2 * margin, spec.grid, spec.base_lambda,
                                 spec.octaves, spec.persistence, spec.relief, spec.seed,
                                 spec.tex_lambda0, spec.tex_octaves, nthreads)
    try:
        info = np.zeros(6)
        L.synth_scene_info(scene, info)
        ground = 0.5 * (info[0] + info[1])
        rng = np.random.default_rng(spec.seed + 1)
        layout = _layout(spec, rng)
        cx, cy = spec.width / 2.0, spec.height / 2.0
        vs = []
        for vi, (c, yaw, pitch, roll) in enumerate(layout):
            R, Cc = from_angles((c[0], c[1], ground + spec.altitude + c[2]), yaw, pitch, roll)
            img = np.zeros((spec.height, spec.width), np.uint8)
            dep = np.zeros((spec.height, spec.width), np.float64) if with_depth else None
            if views is None or vi in views:
                L.synth_render(scene, spec.width, spec.height, spec.focal, cx, cy, spec.k1, spec.k2,
                               np.ascontiguousarray(R.ravel()), Cc, spec.seed + 1000 + vi,
                               spec.noise_dn, img.ravel(),
                               None if dep is None else dep.ctypes.data, nthreads)
            vs.append(View(img, spec.focal, cx, cy, spec.k1, spec.k2, R, Cc, dep))
    finally:
        L.synth_scene_destroy(scene)
    cent = np.array([v.C for v in vs])
    nb = nearest_neighbours(cent, 8)
    ranges = [tuple(spec.depth_range)] * len(vs)
    return Survey(spec, vs, nb, ranges, ground)


def params_for(spec, api_params):
    return api_params(taps=spec.taps, stride=spec.stride, iterations=spec.iterations)


@dataclass
class UnevenSpec:
    """A survey of nadir and oblique views of unequal sizes, with per-view
    neighbour counts and window sizes: the batched-uneven workload (300 views:
    220 nadir 1600x1200 over [40, 80] m, 80 oblique 45-degree edge-ring views
    1920x1080 over [50, 200] m, K uniform in 3..12, windows in {5, 7, 11}),
    scaled by `scale` for tests."""
    name: str
    nadir_lines: int
    nadir_per_line: int
    nadir_size: tuple
    oblique: int
    oblique_size: tuple
    scale: float = 1.0
    altitude: float = 60.0
    patch: tuple = (300.0, 220.0)
    k_range: tuple = (3, 12)
    windows: tuple = (5, 7, 11)
    nadir_range: tuple = (40.0, 80.0)
    oblique_range: tuple = (50.0, 200.0)
    iterations: int = 4
    seed: int = 0xA3D00003


UNEVEN = {
    "c3": UnevenSpec("c3", 11, 20, (1600, 1200), 80, (1920, 1080)),
    # the same geometry with sensors and patch scaled down, for tests
    "uneven_small": UnevenSpec("uneven_small", 3, 5, (320, 240), 6, (384, 216), scale=0.2,
                               patch=(60.0, 44.0), iterations=3, seed=0xA3D0000C),
}


def make_uneven_survey(spec, nthreads=None, views=None):
    """Render an uneven survey; returns (Survey, per-view dict(k, taps,
    stride, range)). Nadir views fly a lawnmower block (C1 field of view,
    75% / 65% overlap); oblique views sit on a ring around the block at the
    same altitude, pitched 45 degrees down and facing the patch centre.
    `views` (optional) restricts which views get pixels and depth (others
    keep an empty image); cameras and per-view draws are the same."""
    if isinstance(spec, str):
        spec = UNEVEN[spec]
    nthreads = nthreads or min(16, os.cpu_count() or 1)
    L = _synth()
    rng = np.random.default_rng(spec.seed + 1)
    alt = spec.altitude * spec.scale
    margin = 3.5 * alt
    px, py = spec.patch
    scene = L.synth_scene_create(-margin, -margin, px + 2 * margin, py + 2 * margin,
                                 0.1 * max(spec.scale, 0.2), 50.0 * spec.scale, 6, 0.5,
                                 15.0 * spec.scale, spec.seed, 0.2, 5, nthreads)
    rendered, cams = [], []
    try:
        info = np.zeros(6)
        L.synth_scene_info(scene, info)
        ground = 0.5 * (info[0] + info[1])
        nw, nh = spec.nadir_size
        fn = 900.0 * nw / 1024
        fw, fh = nw / fn * alt, nh / fn * alt
        along, across = 0.25 * fh, 0.35 * fw
        for r in range(spec.nadir_lines):
            yaw = 90.0 if r % 2 == 0 else 270.0
            for cidx in range(spec.nadir_per_line):
                x = px / 2 + (cidx - (spec.nadir_per_line - 1) / 2) * along
                y = py / 2 + (r - (spec.nadir_lines - 1) / 2) * across
                dp = rng.normal(0.0, 0.05, 3)
                da = rng.normal(0.0, 0.05, 3)
                cams.append(("nadir", (x + dp[0], y + dp[1], alt + dp[2]), yaw + da[0],
                             -90.0 + da[1], da[2]))
        ow, oh = spec.oblique_size
        ring = 0.5 * np.hypot(px, py) + 0.6 * alt
        for o in range(spec.oblique):
            a = 2 * np.pi * o / spec.oblique
            x, y = px / 2 + ring * np.cos(a), py / 2 + ring * np.sin(a)
            # heading (sin yaw, cos yaw) points from the camera to the centre
            yaw = np.degrees(np.arctan2(px / 2 - x, py / 2 - y))
            cams.append(("oblique", (x, y, alt), yaw, -45.0, 0.0))
        for vi, (kind, c, yaw, pitch, roll) in enumerate(cams):
            w, h = (nw, nh) if kind == "nadir" else (ow, oh)
            f = (900.0 * w / 1024) if kind == "nadir" else (900.0 * w / 1024)
            R, Cc = from_angles((c[0], c[1], ground + c[2]), yaw, pitch, roll)
            img = np.zeros((h, w), np.uint8)
            dep = np.zeros((h, w), np.float64)
            if views is None or vi in views:
                L.synth_render(scene, w, h, f, w / 2.0, h / 2.0, 0.0, 0.0,
                               np.ascontiguousarray(R.ravel()), Cc, spec.seed + 1000 + vi, 1.0,
                               img.ravel(), dep.ctypes.data, nthreads)
            rendered.append(View(img, f, w / 2.0, h / 2.0, 0.0, 0.0, R, Cc, dep))
    finally:
        L.synth_scene_destroy(scene)
    cent = np.array([v.C for v in rendered])
    nb = nearest_neighbours(cent, spec.k_range[1])
    per = []
    for vi, (kind, *_rest) in enumerate(cams):
        k = int(rng.integers(spec.k_range[0], spec.k_range[1] + 1))
        taps = int(rng.choice(spec.windows))
        rng_v = spec.nadir_range if kind == "nadir" else spec.oblique_range
        per.append(dict(kind=kind, k=k, taps=taps, stride=1 if taps > 5 else 2,
                        range=tuple(r * spec.scale for r in rng_v)))
    surv = Survey(SPECS["small"], rendered, nb, [p["range"] for p in per], ground)
    return surv, per


# ---- large surveys rendered on the GPU (libsynth_gpu.so) ----

_glib = None


def _synth_gpu():
    global _glib
    if _glib is None:
        if not os.path.exists(_build.LIB_SYNTH_GPU):
            raise RuntimeError("libsynth_gpu.so not built")
        L = C.CDLL(_build.LIB_SYNTH_GPU)
        L.sg_scene_create.argtypes = [C.c_double] * 6 + [C.c_int32, C.c_double, C.c_double,
                                                         C.c_uint64, C.c_double, C.c_int32,
                                                         C.POINTER(C.c_void_p)]
        L.sg_scene_add_boxes.argtypes = [C.c_void_p, _f64p, C.c_int32]
        L.sg_scene_info.argtypes = [C.c_void_p, _f64p]
        L.sg_scene_info.restype = None
        L.sg_render.argtypes = [C.c_void_p, C.c_int32, C.c_int32] + [C.c_double] * 3 + [
            _f64p, _f64p, C.c_uint64, C.c_double, _u8p, C.c_void_p]
        L.sg_scene_destroy.argtypes = [C.c_void_p]
        L.sg_scene_destroy.restype = None
        _glib = L
    return _glib


class GpuScene:
    """The generator's scene on the GPU (same arguments as synth_scene_create)."""

    def __init__(self, x0, y0, width, height, step, base_lambda, octaves, persistence, relief,
                 seed, tex_lambda0, tex_octaves):
        L = _synth_gpu()
        h = C.c_void_p()
        rc = L.sg_scene_create(x0, y0, width, height, step, base_lambda, octaves, persistence,
                               relief, seed, tex_lambda0, tex_octaves, C.byref(h))
        if rc:
            raise RuntimeError(f"sg_scene_create failed ({rc})")
        self._h = h

    def add_boxes(self, boxes):
        b = np.ascontiguousarray(boxes, np.float64).reshape(-1, 5)
        if _synth_gpu().sg_scene_add_boxes(self._h, b.ravel(), len(b)):
            raise RuntimeError("sg_scene_add_boxes failed")

    def info(self):
        out = np.zeros(6)
        _synth_gpu().sg_scene_info(self._h, out)
        return out

    def render(self, w, h, f, cx, cy, R, Cc, noise_seed, noise_dn, with_depth):
        img = np.zeros((h, w), np.uint8)
        dep = np.zeros((h, w), np.float32) if with_depth else None
        rc = _synth_gpu().sg_render(self._h, w, h, f, cx, cy,
                                    np.ascontiguousarray(np.asarray(R, np.float64).ravel()),
                                    np.ascontiguousarray(Cc, np.float64), noise_seed, noise_dn,
                                    img.ravel(), None if dep is None else dep.ctypes.data)
        if rc:
            raise RuntimeError(f"sg_render failed ({rc})")
        return img, dep

    def close(self):
        if self._h:
            _synth_gpu().sg_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class LargeSpec:
    """A large survey: nadir views on a lawnmower block and oblique stations
    of four 45-degree views (north, east, south, west) over a fractal
    heightfield with building-like boxes; rendered on the GPU.

    c5: 2000 views 2400x1800 (1400 nadir in 28 lines of 50, 150 oblique
    stations in 10 x 15), 2 km x 2 km block. The altitude, 183 m, is the one
    at which 1400 nadir views of the c1 field of view cover the block with
    ~75% forward and ~65% side overlap; the oblique stations fly at the same
    height. Depth ranges: nadir [100, 230] m, oblique [100, 520] m. Boxes
    stand on the highest node under them, so overlapping ones stack."""
    name: str
    size: tuple = (2400, 1800)
    nadir_lines: int = 28
    nadir_per_line: int = 50
    oblique_grid: tuple = (10, 15)
    block: tuple = (2000.0, 2000.0)
    altitude: float = 183.0
    margin: float = 640.0  # the far corner rays of an oblique view land ~575 m out
    grid: float = 0.1
    base_lambda: float = 400.0
    octaves: int = 8
    relief: float = 30.0
    boxes: int = 1200
    box_half: tuple = (4.0, 15.0)
    box_height: tuple = (3.0, 30.0)
    tex_lambda0: float = 0.2
    tex_octaves: int = 5
    noise_dn: float = 1.0
    neighbours: int = 10
    taps: int = 7
    stride: int = 1
    iterations: int = 6
    nadir_range: tuple = (100.0, 230.0)
    oblique_range: tuple = (100.0, 520.0)
    depth_every: int = 25  # ground-truth depth for every n-th view
    seed: int = 0xA3D00005


LARGE = {
    "c5": LargeSpec("c5"),
    # the same layout scaled down, for tests
    "large_small": LargeSpec("large_small", size=(320, 240), nadir_lines=4, nadir_per_line=6,
                             oblique_grid=(2, 2), block=(64.0, 72.0), altitude=60.0,
                             margin=200.0, base_lambda=60.0, octaves=6, relief=10.0, boxes=60,
                             box_half=(2.0, 6.0), box_height=(3.0, 12.0), neighbours=6,
                             iterations=4, nadir_range=(30.0, 80.0),
                             oblique_range=(30.0, 220.0), depth_every=1, seed=0xA3D0000D),
}


def make_large_survey(spec, progress=None, views=None):
    """Render a large survey on the GPU. Returns (Survey, per-view dict(kind,
    range)); depth_gt (f32) only on every spec.depth_every-th view. `views`
    (optional) restricts which views get pixels (others keep an empty image
    and no depth); cameras, boxes and neighbours are the same."""
    if isinstance(spec, str):
        spec = LARGE[spec]
    rng = np.random.default_rng(spec.seed + 1)
    bx, by = spec.block
    m = spec.margin
    scene = GpuScene(-m, -m, bx + 2 * m, by + 2 * m, spec.grid, spec.base_lambda, spec.octaves,
                     0.5, spec.relief, spec.seed, spec.tex_lambda0, spec.tex_octaves)
    try:
        ground = 0.5 * spec.relief  # before the boxes: the terrain spans [0, relief]
        half = rng.uniform(spec.box_half[0], spec.box_half[1], (spec.boxes, 2))
        boxes = np.c_[rng.uniform(-0.25 * m, bx + 0.25 * m, spec.boxes),
                      rng.uniform(-0.25 * m, by + 0.25 * m, spec.boxes), half,
                      rng.uniform(spec.box_height[0], spec.box_height[1], spec.boxes)]
        scene.add_boxes(boxes)
        w, h = spec.size
        f = 900.0 * w / 1024  # the c1 field of view
        alt = ground + spec.altitude
        cams = []
        along = bx / (spec.nadir_per_line - 1)
        across = by / (spec.nadir_lines - 1)
        for r in range(spec.nadir_lines):
            yaw = 90.0 if r % 2 == 0 else 270.0
            cols = range(spec.nadir_per_line) if r % 2 == 0 else range(spec.nadir_per_line - 1,
                                                                      -1, -1)
            for cidx in cols:
                dp = rng.normal(0.0, 0.05, 3)
                da = rng.normal(0.0, 0.05, 3)
                cams.append(("nadir", (cidx * along + dp[0], r * across + dp[1], alt + dp[2]),
                             yaw + da[0], -90.0 + da[1], da[2]))
        gx, gy = spec.oblique_grid
        for r in range(gy):
            for cidx in range(gx):
                x = (cidx + 0.5) * bx / gx
                y = (r + 0.5) * by / gy
                for yaw in (0.0, 90.0, 180.0, 270.0):
                    cams.append(("oblique", (x, y, alt), yaw, -45.0, 0.0))
        rendered, per = [], []
        for vi, (kind, c, yaw, pitch, roll) in enumerate(cams):
            R, Cc = from_angles(c, yaw, pitch, roll)
            want_depth = vi % spec.depth_every == 0
            if views is None or vi in views:
                img, dep = scene.render(w, h, f, w / 2.0, h / 2.0, R, Cc, spec.seed + 1000 + vi,
                                        spec.noise_dn, want_depth)
            else:
                img, dep = np.zeros((h, w), np.uint8), None
            rendered.append(View(img, f, w / 2.0, h / 2.0, 0.0, 0.0, R, Cc, dep))
            per.append(dict(kind=kind, yaw=yaw,
                            range=spec.nadir_range if kind == "nadir" else spec.oblique_range))
            if progress and (vi + 1) % 200 == 0:
                progress(vi + 1, len(cams))
    finally:
        scene.close()
    # neighbours: nearest by position among views of the same kind and heading
    cent = np.array([v.C for v in rendered])
    nb = []
    for i in range(len(rendered)):
        same = np.array([j for j in range(len(rendered)) if j != i and per[j]["kind"] ==
                         per[i]["kind"] and (per[i]["kind"] == "nadir" or
                                             per[j]["yaw"] == per[i]["yaw"])], np.int64)
        d = np.hypot(cent[same, 0] - cent[i, 0], cent[same, 1] - cent[i, 1])
        nb.append([int(j) for j in same[np.argsort(d, kind="stable")[:spec.neighbours + 2]]])
    surv = Survey(SPECS["small"], rendered, nb, [p["range"] for p in per], ground)
    surv.extra["large_spec"] = spec
    return surv, per
