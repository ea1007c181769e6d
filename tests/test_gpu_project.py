"""A reference project's dense stage end to end (aerial_mvs/project.py over the
C ABI): cameras.json + photos in; the `.view` cache (prepared on the GPU), the
`.depth` cache, dense.ply, view.ply and dense.voxel out (recon/mod.rs:66-87,
recon/dense.rs:516-651).

The project is the tiny synthetic survey written in the reference's layout,
its photos grey RGB (R = G = B). Checks: every camera with a usable neighbour
gets a `.depth` file identical to the depth map the survey API computes
directly on the same views (the `.view` grey of R = G = B equals the grey
table, so the maps must be bit-identical); dense.ply holds exactly the cloud
of a direct fusion with the same voxel and the origin comment; dense.voxel
is the voxel edge in Rust's float format; the `.view` files are gone; a
second run reuses the `.depth` cache and writes the same cloud; the surface
model built from dense.ply and dense.voxel (dsm.bin) equals the oracle's.
"""
import os
import sys

import numpy as np
import pytest

from aerial_mvs import capi
from aerial_mvs import project as P

import oracle as orc

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _write_project(root, s):
    views = s.views
    k = P.Intrinsics(views[0].f, views[0].cx, views[0].cy, views[0].k1, views[0].k2,
                     views[0].w, views[0].h)
    cams = [P.Camera(i, f"photos/IMG_{i:03d}.JPG", f"IMG_{i:03d}", 0, np.asarray(v.R),
                     np.asarray(v.C), False, list(s.neighbours[i]), tuple(s.ranges[i]))
            for i, v in enumerate(views)]
    pts = []
    for v in views[::2]:
        yy, xx = np.mgrid[0:v.h:9, 0:v.w:9]
        d = v.depth_gt[yy, xx]
        ok = d > 0
        q = np.stack([(xx + 0.5 - v.cx) / v.f * d, (yy + 0.5 - v.cy) / v.f * d, d], -1)[ok]
        pts.append(q @ np.asarray(v.R) + v.C)
    pts = np.concatenate(pts)
    sp = P.Sparse(31, True, 512345.25, 4412345.5, "takeoff", [k], cams, pts,
                  np.full((len(pts), 3), 128, np.uint8), 0.3, [])
    os.makedirs(P.out_dir(root), exist_ok=True)
    P.save_sparse(os.path.join(P.out_dir(root), "cameras.json"), sp)
    return sp


def test_run_dense_on_a_project(gpu_ctx, tiny_survey, tmp_path):
    s = tiny_survey
    root = str(tmp_path / "proj")
    sp = _write_project(root, s)
    key_to_view = {c.path: i for i, c in enumerate(sp.cameras)}

    def decode(path):
        g = s.views[key_to_view[path]].gray
        return np.repeat(g[:, :, None], 3, axis=2)

    _say("run_dense on the tiny project (Lite profile)")
    n, voxel = P.run_dense(root, gpu_ctx, kind=0, decode=decode)
    dp = capi.dense_params_for_profile(0)
    size = dp["size"]
    cdir = P.cache_dir(root)
    odir = P.out_dir(root)
    assert n > 1000
    assert not any(f.endswith(".view") for f in os.listdir(cdir))
    with open(os.path.join(odir, "dense.voxel")) as f:
        assert f.read() == P.rust_f64(voxel)

    # the same depth maps computed directly on the 8-bit views
    prm = capi.params(iterations=dp["iterations"])
    sv = capi.Survey(gpu_ctx, s.views)
    try:
        centres = s.centres()
        jobs = []
        for i, c in enumerate(sp.cameras):
            nb = capi.select_neighbours(centres, i, c.neighbours, c.depth_range[0],
                                        dp["neighbours"])
            if nb:
                jobs.append((i, nb, c.depth_range))
        sv.depth_maps(jobs, prm)
        have = set()
        for (i, _, _) in jobs:
            want, _, _ = sv.download(i)
            got = capi.load_depth(P.depth_path(cdir, sp.cameras[i].key, size))
            assert got is not None and np.array_equal(got, want), i
            have.add(i)
        for i, v in enumerate(s.views):
            sv.set_colour(i, np.repeat(v.gray[:, :, None], 3, axis=2))
        refs = sorted(have)
        nbrs = [[j for j in sp.cameras[i].neighbours[:dp["neighbours"] + 1] if j in have]
                for i in refs]
        xyz, rgb = sv.fuse(refs, nbrs, voxel, rule=capi.MVS_CONS_REFERENCE)
    finally:
        sv.close()
    pxyz, prgb, comment = _read_ply_with_comment(os.path.join(odir, "dense.ply"))
    assert comment == "origin EPSG:32631 E 512345.25 N 4412345.5 height takeoff"
    assert np.array_equal(pxyz, xyz) and np.array_equal(prgb, rgb)
    vxyz, _, _ = _read_ply_with_comment(os.path.join(odir, "view.ply"))
    assert np.array_equal(vxyz, xyz)  # under the viewer's budget: every point
    _say(f"  {len(refs)} depth maps, {n} voxels of {voxel:.4f} m")

    # second run: the .depth cache is reused (the views are prepared again)
    stamp = {f: os.path.getmtime(os.path.join(cdir, f)) for f in os.listdir(cdir)}
    n2, voxel2 = P.run_dense(root, gpu_ctx, kind=0, decode=decode)
    assert (n2, voxel2) == (n, voxel)
    for f, t in stamp.items():
        assert os.path.getmtime(os.path.join(cdir, f)) == t
    pxyz2, _, _ = _read_ply_with_comment(os.path.join(odir, "dense.ply"))
    assert np.array_equal(pxyz2, pxyz)

    # the surface model from dense.ply at the recorded voxel (recon/mod.rs:91-98)
    _say("run_dsm on the project")
    info, h = P.run_dsm(root, gpu_ctx, kind=0)
    winfo, wh = orc.build_dsm(pxyz, float(P.rust_f64(voxel)), orc.DSM_BUDGET[0])
    assert info == winfo and np.array_equal(h, wh, equal_nan=True)
    linfo, lh = capi.load_dsm(os.path.join(odir, "dsm.bin"))
    assert linfo == info and lh.tobytes() == h.tobytes()
    assert np.isfinite(h).mean() > 0.5


def _read_ply_with_comment(path):
    xyz, rgb = capi.read_ply(path)
    with open(path, "rb") as f:
        head = f.read(512).split(b"end_header")[0].decode()
    comment = [ln[len("comment "):] for ln in head.splitlines() if ln.startswith("comment ")]
    return xyz, rgb, comment[0] if comment else None
