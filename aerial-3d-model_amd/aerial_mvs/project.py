"""A reference project's dense stage on the GPU: its files in, its files out.

Mirrors, over the C ABI:
- ``result/3d/cameras.json`` — serde_json of ``Sparse`` (crates/stitch-core/
  src/recon/sparse.rs:26-54, written by recon/mod.rs:26-31, read by
  ``load_sparse`` recon/mod.rs:55-60);
- the ``.view`` cache ``cache/dense/{key}_{size}.view`` (recon/dense.rs:88-113,
  123-154) and the ``.depth`` cache ``{key}_{size}.depth`` (dense.rs:311-313,
  481-508);
- ``reconstruct_dense`` (dense.rs:516-651): prepare missing views, compute the
  missing depth maps (here all in one survey batch), fuse, drop the views;
- ``run_dense`` (recon/mod.rs:66-87): ``dense.ply``, ``view.ply`` and
  ``dense.voxel``;
- the surface-model step of ``run_ortho`` (recon/mod.rs:91-98): ``dense.ply``
  and ``dense.voxel`` in, ``dsm.bin`` out (the orthophoto is not built here).

JSON layout. ``Pose.r`` is a nalgebra ``Matrix3<f64>``; nalgebra 0.33's serde
writes a matrix's storage, a flat array in column-major order, although the
reference's comment (recon/geometry.rs:49) says row-major. No ``cameras.json``
from a real run exists to settle it, so ``load_sparse`` reads column-major by
default and takes ``order="row"`` for the other reading. (The file's own sparse
points do not settle it either: for near-nadir cameras the two readings differ
mostly by the sign of the yaw, and both put many points inside the image.)
"""
import json
import math
import os
from dataclasses import dataclass, field
from decimal import Decimal

import numpy as np

from . import capi

VIEW_MAX_POINTS = 1_500_000  # recon/mod.rs:52


@dataclass
class Intrinsics:
    f: float
    cx: float
    cy: float
    k1: float
    k2: float
    width: int
    height: int


@dataclass
class Camera:
    id: int
    path: str
    key: str
    group: int
    R: np.ndarray  # world -> camera, row-major 3x3
    C: np.ndarray  # centre in world coordinates
    oblique: bool
    neighbours: list
    depth_range: tuple


@dataclass
class Sparse:
    zone: int
    north: bool
    origin_e: float
    origin_n: float
    height_ref: str
    intrinsics: list
    cameras: list
    points: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    colors: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.uint8))
    rms_px: float = 0.0
    skipped: list = field(default_factory=list)
    pose_order: str = "column"

    def epsg(self):
        """geo.rs:27-33."""
        return (32600 if self.north else 32700) + self.zone


def rust_f64(x):
    """Rust's ``Display`` of an f64: the shortest digits that round-trip, never
    in exponent form, no trailing ``.0`` on integers."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    r = repr(x)
    if "e" in r or "E" in r:
        r = format(Decimal(r), "f")
    if r.endswith(".0"):
        r = r[:-2]
    return r


def origin_comment(sp):
    """recon/mod.rs:37-45."""
    return (f"origin EPSG:{sp.epsg()} E {rust_f64(sp.origin_e)} N {rust_f64(sp.origin_n)} "
            f"height {sp.height_ref}")


def load_sparse(path, order="column"):
    """``cameras.json`` -> Sparse. order: how ``Pose.r``'s nine numbers are
    laid out, "column" (nalgebra's storage order) or "row"."""
    with open(path, "rb") as f:
        d = json.loads(f.read())
    intr = [Intrinsics(float(k["f"]), float(k["cx"]), float(k["cy"]), float(k["k1"]),
                       float(k["k2"]), int(k["width"]), int(k["height"]))
            for k in d["intrinsics"]]
    if order not in ("column", "row"):
        raise ValueError("order must be 'column' or 'row'")
    first = order
    cams = []
    for c in d["cameras"]:
        r = np.asarray(c["pose"]["r"], np.float64).reshape(3, 3)
        cams.append(Camera(int(c["id"]), str(c["path"]), str(c["key"]), int(c["group"]),
                           r if first == "row" else r.T, np.asarray(c["pose"]["c"], np.float64),
                           bool(c["oblique"]), [int(j) for j in c["neighbours"]],
                           (float(c["depth_range"][0]), float(c["depth_range"][1]))))
    sp = Sparse(int(d["zone"]["zone"]), bool(d["zone"]["north"]), float(d["origin_e"]),
                float(d["origin_n"]), str(d["height_ref"]), intr, cams,
                np.asarray(d["points"], np.float64).reshape(-1, 3),
                np.asarray(d["colors"], np.uint8).reshape(-1, 3), float(d["rms_px"]),
                [int(s) for s in d["skipped"]], first)
    return sp


def save_sparse(path, sp, order="column"):
    """Sparse -> ``cameras.json`` in the serde layout ``load_sparse`` reads
    (``Pose.r`` in ``order``)."""
    def r_of(c):
        return (c.R if order == "row" else c.R.T).ravel().tolist()
    d = dict(
        zone=dict(zone=sp.zone, north=sp.north), origin_e=sp.origin_e, origin_n=sp.origin_n,
        height_ref=sp.height_ref,
        intrinsics=[dict(f=k.f, cx=k.cx, cy=k.cy, k1=k.k1, k2=k.k2, width=k.width,
                         height=k.height) for k in sp.intrinsics],
        cameras=[dict(id=c.id, path=c.path, key=c.key, group=c.group,
                      pose=dict(r=r_of(c), c=np.asarray(c.C, np.float64).tolist()),
                      oblique=c.oblique, neighbours=list(c.neighbours),
                      depth_range=list(c.depth_range)) for c in sp.cameras],
        points=np.asarray(sp.points, np.float64).tolist(),
        colors=np.asarray(sp.colors, np.uint8).tolist(), rms_px=sp.rms_px,
        skipped=list(sp.skipped))
    with open(path, "w") as f:
        json.dump(d, f)


def cache_dir(project_dir):
    """project.rs:156-159 for kind "dense"."""
    return os.path.join(project_dir, "cache", "dense")


def out_dir(project_dir):
    """recon/mod.rs:21-23."""
    return os.path.join(project_dir, "result", "3d")


def view_path(dirname, key, size):
    """dense.rs:88-90."""
    return os.path.join(dirname, f"{key}_{size}.view")


def depth_path(dirname, key, size):
    """dense.rs:311-313."""
    return os.path.join(dirname, f"{key}_{size}.depth")


class _View:
    """A working-size view for the C ABI (f32 grey of the `.view` cache)."""

    def __init__(self, grey, rgb, k, cam):
        self.gray_f32 = grey
        self.rgb = rgb
        self.h, self.w = grey.shape
        # load_view (dense.rs:139-145): f, cx, cy scaled by w / full width
        self.f, self.cx, self.cy = capi.view_intrinsics(k.f, k.cx, k.cy, k.width, self.w)
        self.k1, self.k2 = k.k1, k.k2
        self.R, self.C = cam.R, cam.C


@dataclass
class DenseCloud:
    points: np.ndarray
    colors: np.ndarray
    voxel_m: float


def reconstruct_dense(project_dir, sp, ctx, kind=1, decode=None, prm=None, keep_views=False,
                      progress=None):
    """dense.rs:516-651 on the GPU. kind: profile (0 Lite, 1 Standard, 2 High,
    ``DenseParams::for_profile``); decode(path) -> (h, w, 3) uint8 photo,
    needed only for views missing from the cache; prm: MvsParams of the depth
    maps (default: the reference's window with the profile's iterations)."""
    dp = capi.dense_params_for_profile(kind)
    size = dp["size"]
    cdir = cache_dir(project_dir)
    os.makedirs(cdir, exist_ok=True)
    n = len(sp.cameras)
    # prepare_views (dense.rs:92-121): the missing .view files, on the GPU
    for i, c in enumerate(sp.cameras):
        p = view_path(cdir, c.key, size)
        if not os.path.exists(p):
            if decode is None:
                raise FileNotFoundError(f"{p} missing and no decoder given")
            rgb, grey = ctx.prepare_view(decode(c.path), size)
            capi.save_view(p, grey, rgb)
        if progress:
            progress("dense-prep", i + 1, n)
    views = []
    for c in sp.cameras:
        got = capi.load_view(view_path(cdir, c.key, size))
        if got is None:
            raise ValueError(f"damaged cache file for {c.path}")  # dense.rs:130-132
        views.append(_View(got[0], got[1], sp.intrinsics[c.group], c))
    if prm is None:
        prm = capi.params(iterations=dp["iterations"])
    sv = capi.Survey(ctx, views)
    try:
        for i, v in enumerate(views):
            sv.set_colour(i, v.rgb)
        # depth maps, cached per photo (dense.rs:528-540); the missing ones in
        # one batch, the neighbour filter of dense.rs:324-334
        centres = np.array([c.C for c in sp.cameras], np.float64)
        jobs, has = [], [False] * n
        for i, c in enumerate(sp.cameras):
            p = depth_path(cdir, c.key, size)
            if os.path.exists(p):  # cached: reused as is; a damaged file fuses nothing
                d = capi.load_depth(p)
                if d is not None and d.shape == (views[i].h, views[i].w):
                    sv.set_depth(i, d)
                    has[i] = True
                continue
            nb = capi.select_neighbours(centres, i, c.neighbours, c.depth_range[0],
                                        dp["neighbours"])
            if nb:
                jobs.append((i, nb, c.depth_range))
        if jobs:
            sv.depth_maps(jobs, prm)
            for (i, _, _) in jobs:
                d, _, _ = sv.download(i)
                capi.save_depth(depth_path(cdir, sp.cameras[i].key, size), d)
                has[i] = True
        if progress:
            progress("depth", n, n)
        # fusion (dense.rs:542-651): the voxel edge, then every map with the
        # neighbours that have one
        k0 = sp.intrinsics[0]
        voxel = capi.fusion_voxel(sp.points[:, 0], sp.points[:, 1],
                                  [c.depth_range for c in sp.cameras], k0.f, size, k0.width,
                                  k0.height, dp["max_points"])
        refs = [i for i in range(n) if has[i]]
        nbrs = [[j for j in sp.cameras[i].neighbours[:dp["neighbours"] + 1] if has[j]]
                for i in refs]
        xyz, rgb = sv.fuse(refs, nbrs, voxel, rule=capi.MVS_CONS_REFERENCE)
    finally:
        sv.close()
    if not keep_views:  # dense.rs:643-647
        for c in sp.cameras:
            p = view_path(cdir, c.key, size)
            if os.path.exists(p):
                os.remove(p)
    return DenseCloud(xyz, rgb, voxel)


def run_dense(project_dir, ctx, kind=1, decode=None, prm=None, order="column", progress=None):
    """recon/mod.rs:66-87: cameras.json in; dense.ply, view.ply (at most
    1.5 M points, every step-th) and dense.voxel out. Returns (points, voxel)."""
    odir = out_dir(project_dir)
    sp = load_sparse(os.path.join(odir, "cameras.json"), order=order)
    cloud = reconstruct_dense(project_dir, sp, ctx, kind=kind, decode=decode, prm=prm,
                              progress=progress)
    comment = origin_comment(sp)
    capi.write_ply(os.path.join(odir, "dense.ply"), cloud.points, cloud.colors, comment)
    step = max(1, -(-len(cloud.points) // VIEW_MAX_POINTS))
    capi.write_ply(os.path.join(odir, "view.ply"), cloud.points[::step], cloud.colors[::step],
                   comment)
    with open(os.path.join(odir, "dense.voxel"), "w") as f:
        f.write(rust_f64(cloud.voxel_m))
    return len(cloud.points), cloud.voxel_m


def dense_voxel(project_dir):
    """recon/mod.rs:102-109: the voxel edge ``run_dense`` recorded, or None
    (missing file, or text Rust's f64 parser refuses)."""
    try:
        with open(os.path.join(out_dir(project_dir), "dense.voxel")) as f:
            t = f.read().strip()
    except OSError:
        return None
    if "_" in t:  # Python's float() takes digit separators, Rust's parse does not
        return None
    try:
        return float(t)
    except ValueError:
        return None


def run_dsm(project_dir, ctx, kind=1):
    """The surface model of ``run_ortho`` (recon/mod.rs:91-98) on the GPU:
    ``dense.ply`` and ``dense.voxel`` (0.25 m when absent) in, ``dsm.bin`` out.
    kind: the profile (0 Lite, 1 Standard, 2 High: 3, 8, 20 M cells,
    dsm.rs:121-128). Returns (info, heights)."""
    odir = out_dir(project_dir)
    voxel = dense_voxel(project_dir)
    if voxel is None:
        voxel = 0.25
    xyz, _ = capi.read_ply(os.path.join(odir, "dense.ply"))
    info, h = ctx.build_dsm(xyz, voxel, kind)
    capi.save_dsm(os.path.join(odir, "dsm.bin"), info, h)
    return info, h
