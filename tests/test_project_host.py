"""Host side of the project interop (aerial_mvs/project.py): the cameras.json
layout (serde of recon/sparse.rs:26-54 with nalgebra's column-major matrix
storage, or row-major on request), Rust's float formatting for the origin
comment and dense.voxel (recon/mod.rs:37-45, 85), and the cache paths.
No GPU.
"""
import json
import os

import numpy as np

from aerial_mvs import project as P


def _rot(yaw, pitch, roll):
    cy, sy, cp, sp_, cr, sr = (np.cos(yaw), np.sin(yaw), np.cos(pitch), np.sin(pitch),
                               np.cos(roll), np.sin(roll))
    Rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    Ry = np.array([[cp, 0, sp_], [0, 1, 0], [-sp_, 0, cp]])
    Rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    return Rz @ Ry @ Rx


def _sparse(n_cam=4, n_pts=300, seed=3):
    """Nadir-ish cameras (R = world->camera, looking down -z) over a patch of
    ground points."""
    rng = np.random.default_rng(seed)
    down = np.array([[1.0, 0, 0], [0, -1, 0], [0, 0, -1]])
    cams = []
    for i in range(n_cam):
        R = down @ _rot(0.4 + 0.7 * i, 0.25, 0.1)  # yawed and tilted
        C = np.array([10.0 * i, 5.0 * (i % 2), 60.0])
        cams.append(P.Camera(i, f"/photos/IMG_{i:04d}.JPG", f"IMG_{i:04d}", 0, R, C, False,
                             [j for j in range(n_cam) if j != i], (40.0, 80.0)))
    pts = np.c_[rng.uniform(-20, 50, n_pts), rng.uniform(-20, 30, n_pts), rng.uniform(-3, 3, n_pts)]
    return P.Sparse(32, True, 500000.0, 4649776.5, "takeoff",
                    [P.Intrinsics(900.0, 512.0, 384.0, -0.01, 0.002, 1024, 768)], cams, pts,
                    rng.integers(0, 256, (n_pts, 3)).astype(np.uint8), 0.42, [7])


def test_cameras_json_layout(tmp_path):
    """A hand-written file in the serde layout: Pose.r is nalgebra's storage,
    column-major, so r[j*3 + i] is row i, column j."""
    d = dict(zone=dict(zone=33, north=False), origin_e=1.5, origin_n=2.0, height_ref="gps",
             intrinsics=[dict(f=1000.0, cx=500.5, cy=400.25, k1=0.01, k2=-0.001, width=1000,
                              height=800)],
             cameras=[dict(id=5, path="a/b.jpg", key="b", group=0,
                           pose=dict(r=[1, 2, 3, 4, 5, 6, 7, 8, 9], c=[1.0, 2.0, 3.0]),
                           oblique=True, neighbours=[2, 0], depth_range=[12.5, 80.0])],
             points=[[1.0, 2.0, 3.0]], colors=[[9, 8, 7]], rms_px=0.5, skipped=[1, 4])
    p = tmp_path / "cameras.json"
    p.write_text(json.dumps(d))
    sp = P.load_sparse(str(p))
    c = sp.cameras[0]
    assert np.array_equal(c.R, [[1, 4, 7], [2, 5, 8], [3, 6, 9]])
    assert np.array_equal(P.load_sparse(str(p), order="row").cameras[0].R,
                          [[1, 2, 3], [4, 5, 6], [7, 8, 9]])
    assert (c.id, c.path, c.key, c.group, c.oblique, c.neighbours, c.depth_range) == (
        5, "a/b.jpg", "b", 0, True, [2, 0], (12.5, 80.0))
    assert sp.epsg() == 32733 and sp.skipped == [1, 4] and sp.rms_px == 0.5
    k = sp.intrinsics[0]
    assert (k.f, k.cx, k.cy, k.k1, k.k2, k.width, k.height) == (1000.0, 500.5, 400.25, 0.01,
                                                                 -0.001, 1000, 800)
    assert np.array_equal(sp.colors, [[9, 8, 7]])


def test_save_load_round_trip(tmp_path):
    sp = _sparse()
    for order in ("column", "row"):
        p = str(tmp_path / f"cams_{order}.json")
        P.save_sparse(p, sp, order=order)
        back = P.load_sparse(p, order=order)
        for a, b in zip(sp.cameras, back.cameras):
            assert np.array_equal(a.R, b.R) and np.array_equal(a.C, b.C)
            assert (a.key, a.neighbours, a.depth_range) == (b.key, b.neighbours, b.depth_range)
        assert np.array_equal(back.points, sp.points)
        assert np.array_equal(back.colors, sp.colors)


def test_rust_float_display():
    cases = {500000.0: "500000", 4649776.5: "4649776.5", 0.1: "0.1", 1e-7: "0.0000001",
             1e21: "1000000000000000000000", -0.0: "-0", 0.30000000000000004:
             "0.30000000000000004", 2.5e-5: "0.000025", 123456789.0: "123456789"}
    for x, want in cases.items():
        assert P.rust_f64(x) == want, x


def test_origin_comment_and_paths():
    sp = _sparse()
    assert P.origin_comment(sp) == "origin EPSG:32632 E 500000 N 4649776.5 height takeoff"
    assert P.view_path("/p/cache/dense", "IMG_1", 960) == "/p/cache/dense/IMG_1_960.view"
    assert P.depth_path("/p/cache/dense", "IMG_1", 960) == "/p/cache/dense/IMG_1_960.depth"
    assert P.cache_dir("/p") == os.path.join("/p", "cache", "dense")
    assert P.out_dir("/p") == os.path.join("/p", "result", "3d")


def test_dense_voxel_file(tmp_path):
    """recon/mod.rs:102-109: trimmed text parsed as f64, None otherwise."""
    root = str(tmp_path)
    assert P.dense_voxel(root) is None
    os.makedirs(P.out_dir(root))
    path = os.path.join(P.out_dir(root), "dense.voxel")
    for text, want in (("0.1875", 0.1875), (" 0.3\n", 0.3), ("1e-1", 0.1), ("abc", None),
                       ("1_0", None), ("", None)):
        with open(path, "w") as f:
            f.write(text)
        assert P.dense_voxel(root) == want, text
