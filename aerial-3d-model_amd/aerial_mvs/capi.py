"""ctypes binding of ``libaerial_mvs.so`` (the C ABI declared in include/aerial_mvs.h).

This is glue for the tests and the benchmark; the engine itself is the HIP
library. Loading fails loudly when the library is missing — there is no other
implementation behind these calls.
"""
import ctypes as C
import os

import numpy as np

from . import build as _build

# AERIAL_MVS_LIB selects another build of the same library (the sanitizer
# build of the host code, for instance)
LIB_PATH = os.environ.get("AERIAL_MVS_LIB") or _build.LIB_MVS

MVS_OK = 0
MVS_NO_NEIGHBOURS = 1
MVS_ERR_ARG = -1
MVS_ERR_DEVICE = -2
MVS_ERR_CANCELLED = -3
MVS_ERR_NOMEM = -4
MVS_ERR_IO = -5
MVS_ERR_FORMAT = -6
MVS_GRAY_U8 = 0
MVS_GRAY_F32 = 1
MVS_CONS_REFERENCE = 0
MVS_CONS_REPROJ = 1
MVS_MODE_EXACT = 0
MVS_MODE_FAST = 1
MVS_SURVEY_LEAN_VIEWS = 1
MVS_SURVEY_DEPTH_ONLY = 2

EXPORTS = [
    "mvs_default_params", "mvs_dense_params_for_profile", "mvs_select_neighbours",
    "mvs_create", "mvs_destroy", "mvs_last_error", "mvs_synchronize", "mvs_depth_map",
    "mvs_ncc_cost", "mvs_patchmatch_steps", "mvs_survey_create", "mvs_survey_destroy",
    "mvs_survey_set_view", "mvs_survey_depth_map", "mvs_survey_depth_maps",
    "mvs_survey_download", "mvs_survey_consistency", "mvs_set_timing", "mvs_get_stats",
    "mvs_reset_stats", "mvs_save_depth", "mvs_load_depth", "mvs_selftest",
    "mvs_set_schedule", "mvs_tuned_schedules", "mvs_ncc_pair", "mvs_set_graphs",
    "mvs_survey_depth_maps_ex", "mvs_save_view", "mvs_load_view", "mvs_view_intrinsics",
    "mvs_survey_set_colour", "mvs_survey_fuse", "mvs_survey_fused_points", "mvs_fusion_voxel",
    "mvs_write_ply", "mvs_read_ply", "mvs_prepared_size", "mvs_prepare_view",
    "mvs_survey_set_depth", "mvs_build_dsm", "mvs_dsm_download", "mvs_save_dsm", "mvs_survey_set_options",
    "mvs_load_dsm", "mvs_dsm_height", "mvs_refine_terms", "mvs_survey_consistency_noalloc",
]

# Limits of the allocation-free consistency filter (aerial_mvs.h)
MVS_NOALLOC_MAX_PIXELS = 1 << 23
MVS_NOALLOC_MAX_NBRS = 16


class MvsView(C.Structure):
    _fields_ = [("w", C.c_int32), ("h", C.c_int32), ("stride", C.c_int32),
                ("format", C.c_int32), ("pixels", C.c_void_p),
                ("f", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
                ("k1", C.c_double), ("k2", C.c_double),
                ("R", C.c_double * 9), ("C", C.c_double * 3)]


class MvsParams(C.Structure):
    _fields_ = [("taps", C.c_int32), ("stride", C.c_int32), ("iterations", C.c_int32),
                ("texture_min", C.c_float), ("cost_max", C.c_float),
                ("refine_steps", C.c_int32), ("mode", C.c_int32), ("reserved", C.c_int32 * 1)]


class MvsStats(C.Structure):
    _fields_ = [("sweep_launches", C.c_int64), ("sweep_ms", C.c_double),
                ("init_launches", C.c_int64), ("init_ms", C.c_double),
                ("evals", C.c_int64), ("pixels", C.c_int64),
                ("init_evals", C.c_int64), ("sweep_evals", C.c_int64),
                ("split", C.c_int32), ("gpc", C.c_int32),
                ("refine_launches", C.c_int64), ("refine_ms", C.c_double),
                ("refine_pixels", C.c_int64), ("refine_accepted", C.c_int64),
                ("graph_captures", C.c_int64), ("graph_replays", C.c_int64),
                ("fuse_passes", C.c_int64), ("sweep_kernels", C.c_int64),
                ("sweep_skipped", C.c_int64)]


PROGRESS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_char_p, C.c_int32, C.c_int32)

_lib = None


class MvsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"aerial_mvs error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP engine not built ({LIB_PATH} missing): run "
                               "`python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double
        P = C.POINTER
        L.mvs_default_params.argtypes = [P(MvsParams)]
        L.mvs_dense_params_for_profile.argtypes = [i32, P(C.c_int64)]
        L.mvs_select_neighbours.argtypes = [P(C.c_double), i32, i32, P(C.c_int32), i32, f64, i32,
                                            P(C.c_int32)]
        L.mvs_select_neighbours.restype = i32
        L.mvs_create.argtypes = [i32, P(vp), C.c_char_p, C.c_size_t]
        L.mvs_destroy.argtypes = [vp]
        L.mvs_destroy.restype = None
        L.mvs_last_error.argtypes = [vp]
        L.mvs_last_error.restype = C.c_char_p
        L.mvs_synchronize.argtypes = [vp]
        L.mvs_depth_map.argtypes = [vp, P(MvsView), P(MvsView), i32, f64, f64, P(MvsParams),
                                    vp, vp, vp, vp, PROGRESS_FN, vp]
        L.mvs_ncc_cost.argtypes = [vp, P(MvsView), P(MvsView), i32, f64, f64, P(MvsParams), i64,
                                   vp, vp, vp, vp]
        L.mvs_ncc_pair.argtypes = [vp, P(MvsView), P(MvsView), i32, f64, f64, P(MvsParams), i64,
                                   vp, vp, vp, vp, vp, vp, vp]
        L.mvs_refine_terms.argtypes = [vp, P(MvsView), P(MvsView), i32, f64, f64,
                                       P(MvsParams), i64, vp, vp, vp, vp, vp, vp, vp]
        L.mvs_patchmatch_steps.argtypes = [vp, P(MvsView), P(MvsView), i32, f64, f64,
                                           P(MvsParams), i32, i32, vp, vp, P(C.c_int64)]
        L.mvs_survey_create.argtypes = [vp, i32, P(vp)]
        L.mvs_survey_destroy.argtypes = [vp]
        L.mvs_survey_destroy.restype = None
        L.mvs_survey_set_view.argtypes = [vp, i32, P(MvsView)]
        L.mvs_survey_depth_map.argtypes = [vp, i32, vp, i32, f64, f64, P(MvsParams)]
        L.mvs_survey_depth_maps.argtypes = [vp, i32, vp, vp, vp, vp, P(MvsParams), vp,
                                            PROGRESS_FN, vp]
        L.mvs_survey_depth_maps_ex.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp, PROGRESS_FN, vp]
        L.mvs_survey_download.argtypes = [vp, i32, vp, vp, vp]
        L.mvs_survey_consistency.argtypes = [vp, i32, vp, i32, i32, vp, P(C.c_int64)]
        L.mvs_survey_consistency_noalloc.argtypes = [vp, i32, vp, i32, i32, vp, P(C.c_int64)]
        L.mvs_set_timing.argtypes = [vp, i32]
        L.mvs_get_stats.argtypes = [vp, P(MvsStats)]
        L.mvs_reset_stats.argtypes = [vp]
        L.mvs_set_schedule.argtypes = [vp, i32, i32, i32, i32]
        L.mvs_set_graphs.argtypes = [vp, i32]
        L.mvs_tuned_schedules.argtypes = [vp, vp, i32]
        L.mvs_tuned_schedules.restype = i32
        L.mvs_selftest.argtypes = [vp, i32, i64, C.c_uint64, P(C.c_int64)]
        L.mvs_save_depth.argtypes = [C.c_char_p, i32, i32, vp]
        L.mvs_load_depth.argtypes = [C.c_char_p, P(C.c_int32), P(C.c_int32), vp, i64]
        L.mvs_save_view.argtypes = [C.c_char_p, i32, i32, vp, vp]
        L.mvs_load_view.argtypes = [C.c_char_p, P(C.c_int32), P(C.c_int32), vp, vp, i64]
        L.mvs_view_intrinsics.argtypes = [f64, f64, f64, i32, i32, P(C.c_double)]
        L.mvs_prepared_size.argtypes = [i32, i32, i32, P(C.c_int32), P(C.c_int32),
                                        P(C.c_double)]
        L.mvs_prepare_view.argtypes = [vp, vp, i32, i32, i32, i32, vp, vp]
        L.mvs_survey_set_depth.argtypes = [vp, i32, vp]
        L.mvs_build_dsm.argtypes = [vp, vp, i64, f64, i32, P(MvsDsmInfo), vp, i64]
        L.mvs_dsm_download.argtypes = [vp, vp, i64]
        L.mvs_survey_set_options.argtypes = [vp, i32]
        L.mvs_save_dsm.argtypes = [C.c_char_p, P(MvsDsmInfo), vp]
        L.mvs_load_dsm.argtypes = [C.c_char_p, P(MvsDsmInfo), vp, i64]
        L.mvs_dsm_height.argtypes = [P(MvsDsmInfo), vp, f64, f64, P(C.c_float)]
        L.mvs_survey_set_colour.argtypes = [vp, i32, vp]
        L.mvs_survey_fuse.argtypes = [vp, i32, vp, vp, vp, i32, f64, P(C.c_int64)]
        L.mvs_survey_fused_points.argtypes = [vp, vp, vp, i64, P(C.c_int64)]
        L.mvs_fusion_voxel.argtypes = [vp, vp, i64, vp, i32, f64, i32, i32, i32, i64,
                                       P(C.c_double)]
        L.mvs_write_ply.argtypes = [C.c_char_p, vp, vp, i64, C.c_char_p]
        L.mvs_read_ply.argtypes = [C.c_char_p, vp, vp, i64, P(C.c_int64)]
        _lib = L
    return _lib


def params(taps=5, stride=2, iterations=4, texture_min=0.008, cost_max=0.45, refine_steps=0,
           fast=False):
    p = MvsParams()
    lib().mvs_default_params(C.byref(p))
    p.taps, p.stride, p.iterations = taps, stride, iterations
    p.refine_steps = refine_steps
    p.mode = MVS_MODE_FAST if fast else MVS_MODE_EXACT
    p.texture_min, p.cost_max = texture_min, cost_max
    return p


def _ptr(a):
    return None if a is None else a.ctypes.data


def make_view(v, keep):
    """mvs_view of a survey View (u8 grey, or f32 grey when `gray_f32` is set)."""
    mv = MvsView()
    g32 = getattr(v, "gray_f32", None)
    if g32 is not None:
        arr = np.ascontiguousarray(g32, np.float32)
        mv.format = MVS_GRAY_F32
    else:
        arr = np.ascontiguousarray(v.gray, np.uint8)
        mv.format = MVS_GRAY_U8
    keep.append(arr)
    mv.h, mv.w = arr.shape
    mv.stride = arr.shape[1]
    mv.pixels = arr.ctypes.data
    mv.f, mv.cx, mv.cy, mv.k1, mv.k2 = v.f, v.cx, v.cy, v.k1, v.k2
    mv.R[:] = [float(a) for a in np.asarray(v.R, np.float64).ravel()]
    mv.C[:] = [float(a) for a in np.asarray(v.C, np.float64).ravel()]
    return mv


def dense_params_for_profile(kind):
    out = (C.c_int64 * 4)()
    rc = lib().mvs_dense_params_for_profile(kind, out)
    if rc:
        raise MvsError(rc, "unknown profile")
    return dict(size=out[0], iterations=out[1], neighbours=out[2], max_points=out[3])


def select_neighbours(centres, ci, neighbours, range0, take):
    centres = np.ascontiguousarray(centres, np.float64)
    nb = np.ascontiguousarray(neighbours, np.int32)
    out = np.zeros(max(len(nb), 1), np.int32)
    k = lib().mvs_select_neighbours(
        centres.ctypes.data_as(C.POINTER(C.c_double)), len(centres), ci,
        nb.ctypes.data_as(C.POINTER(C.c_int32)), len(nb), range0, take,
        out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out[:k].tolist()


def save_depth(path, depth):
    d = np.ascontiguousarray(depth, np.float32)
    rc = lib().mvs_save_depth(os.fsencode(path), d.shape[1], d.shape[0], d.ctypes.data)
    if rc:
        raise MvsError(rc, "save_depth failed")


def load_depth(path):
    w, h = C.c_int32(), C.c_int32()
    rc = lib().mvs_load_depth(os.fsencode(path), C.byref(w), C.byref(h), None, 0)
    if rc:
        return None
    d = np.zeros((h.value, w.value), np.float32)
    rc = lib().mvs_load_depth(os.fsencode(path), C.byref(w), C.byref(h), d.ctypes.data, d.size)
    return None if rc else d


def save_view(path, gray, rgb):
    """A `.view` cache file: (h, w) f32 grey and (h, w, 3) u8 RGB."""
    g = np.ascontiguousarray(gray, np.float32)
    c = np.ascontiguousarray(rgb, np.uint8)
    assert c.shape == g.shape + (3,)
    rc = lib().mvs_save_view(os.fsencode(path), g.shape[1], g.shape[0], g.ctypes.data,
                             c.ctypes.data)
    if rc:
        raise MvsError(rc, "save_view failed")


def load_view(path):
    """(grey (h, w) f32, rgb (h, w, 3) u8), or None when missing or damaged."""
    w, h = C.c_int32(), C.c_int32()
    if lib().mvs_load_view(os.fsencode(path), C.byref(w), C.byref(h), None, None, 0):
        return None
    g = np.zeros((h.value, w.value), np.float32)
    c = np.zeros((h.value, w.value, 3), np.uint8)
    if lib().mvs_load_view(os.fsencode(path), C.byref(w), C.byref(h), g.ctypes.data,
                           c.ctypes.data, g.size):
        return None
    return g, c


class MvsDsmInfo(C.Structure):
    _fields_ = [("x0", C.c_double), ("y0", C.c_double), ("cell", C.c_double),
                ("nx", C.c_int64), ("ny", C.c_int64), ("zmax", C.c_float),
                ("reserved", C.c_int32)]

    def as_dict(self):
        return dict(x0=self.x0, y0=self.y0, cell=self.cell, nx=self.nx, ny=self.ny,
                    zmax=self.zmax)


def _dsm_info(d):
    return MvsDsmInfo(d["x0"], d["y0"], d["cell"], d["nx"], d["ny"], d.get("zmax", 0.0), 0)


def save_dsm(path, info, h):
    """``dsm.bin`` (Dsm::save, dsm.rs:27-46)."""
    g = np.ascontiguousarray(h, np.float32)
    rc = lib().mvs_save_dsm(os.fsencode(path), C.byref(_dsm_info(info)), g.ctypes.data)
    if rc:
        raise MvsError(rc, "save_dsm failed")


def load_dsm(path):
    """(info, h (ny, nx) f32), or None for a missing, damaged or truncated file."""
    d = MvsDsmInfo()
    if lib().mvs_load_dsm(os.fsencode(path), C.byref(d), None, 0):
        return None
    h = np.zeros((d.ny, d.nx), np.float32)
    if lib().mvs_load_dsm(os.fsencode(path), C.byref(d), h.ctypes.data, h.size):
        return None
    return d.as_dict(), h


def dsm_height(info, h, x, y):
    """Dsm::height (dsm.rs:82-119): the height, or None outside."""
    g = np.ascontiguousarray(h, np.float32)
    out = C.c_float()
    rc = lib().mvs_dsm_height(C.byref(_dsm_info(info)), g.ctypes.data, x, y, C.byref(out))
    if rc < 0:
        raise MvsError(rc, "bad surface model")
    return out.value if rc == 1 else None


def prepared_size(w, h, max_side):
    """shrink_to's output size and scale (imgops.rs:145-171)."""
    ow, oh, sc = C.c_int32(), C.c_int32(), C.c_double()
    rc = lib().mvs_prepared_size(w, h, max_side, C.byref(ow), C.byref(oh), C.byref(sc))
    if rc:
        raise MvsError(rc, "bad sizes")
    return ow.value, oh.value, sc.value


def view_intrinsics(f, cx, cy, full_width, w):
    out = (C.c_double * 3)()
    rc = lib().mvs_view_intrinsics(f, cx, cy, full_width, w, out)
    if rc:
        raise MvsError(rc, "bad intrinsics")
    return tuple(out)


def fusion_voxel(xs, ys, ranges, f, size, width, height, max_points):
    xs = np.ascontiguousarray(xs, np.float64)
    ys = np.ascontiguousarray(ys, np.float64)
    rg = np.ascontiguousarray(ranges, np.float64).ravel()
    out = C.c_double()
    rc = lib().mvs_fusion_voxel(xs.ctypes.data, ys.ctypes.data, len(xs), rg.ctypes.data,
                                len(rg) // 2, f, size, width, height, max_points, C.byref(out))
    if rc:
        raise MvsError(rc, "bad fusion inputs")
    return out.value


def write_ply(path, xyz, rgb, comment=""):
    p = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    c = np.ascontiguousarray(rgb, np.uint8).reshape(-1, 3)
    rc = lib().mvs_write_ply(os.fsencode(path), p.ctypes.data, c.ctypes.data, len(p),
                             comment.encode())
    if rc:
        raise MvsError(rc, "write_ply failed")


def read_ply(path):
    n = C.c_int64()
    rc = lib().mvs_read_ply(os.fsencode(path), None, None, 0, C.byref(n))
    if rc:
        raise MvsError(rc, "read_ply failed")
    p = np.zeros((n.value, 3), np.float32)
    c = np.zeros((n.value, 3), np.uint8)
    rc = lib().mvs_read_ply(os.fsencode(path), p.ctypes.data, c.ctypes.data, n.value, C.byref(n))
    if rc:
        raise MvsError(rc, "read_ply failed")
    return p, c


class Context:
    """One mvs_ctx: a HIP stream and its device buffers on one GPU."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = lib().mvs_create(device, C.byref(self._h), err, len(err))
        if rc:
            raise MvsError(rc, err.value.decode())

    def close(self):
        if self._h:
            lib().mvs_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def check(self, rc, allow=(MVS_OK,)):
        if rc not in allow:
            raise MvsError(rc, lib().mvs_last_error(self._h).decode())
        return rc

    def synchronize(self):
        self.check(lib().mvs_synchronize(self._h))

    def _views(self, ref, srcs):
        keep = []
        r = make_view(ref, keep)
        arr = (MvsView * max(len(srcs), 1))()
        for i, s in enumerate(srcs):
            arr[i] = make_view(s, keep)
        return r, arr, keep

    def depth_map(self, ref, srcs, range0, range1, prm, cancel=None, progress=None):
        """Returns (depth, normal, cost) or None when there are no neighbours.
        progress(stage, done, total) -> truthy to cancel."""
        r, arr, keep = self._views(ref, srcs)
        h, w = ref.h, ref.w
        d = np.zeros((h, w), np.float32)
        n = np.zeros((h, w, 3), np.float32)
        c = np.zeros((h, w), np.float32)
        cflag = None if cancel is None else C.cast(C.pointer(cancel), C.c_void_p)
        cb = PROGRESS_FN()
        if progress is not None:
            cb = PROGRESS_FN(lambda user, stage, done, total:
                             1 if progress(stage.decode(), done, total) else 0)
        rc = lib().mvs_depth_map(self._h, C.byref(r), arr, len(srcs), range0, range1,
                                 C.byref(prm), d.ctypes.data, n.ctypes.data, c.ctypes.data, cflag,
                                 cb, None)
        if self.check(rc, (MVS_OK, MVS_NO_NEIGHBOURS)) == MVS_NO_NEIGHBOURS:
            return None
        return d, n, c

    def ncc_cost(self, ref, srcs, range0, range1, prm, xs, ys, hyps):
        r, arr, keep = self._views(ref, srcs)
        xs = np.ascontiguousarray(xs, np.int32)
        ys = np.ascontiguousarray(ys, np.int32)
        hy = np.ascontiguousarray(hyps, np.float32).reshape(-1, 4)
        out = np.zeros(len(xs), np.float32)
        self.check(lib().mvs_ncc_cost(self._h, C.byref(r), arr, len(srcs), range0, range1,
                                      C.byref(prm), len(xs), xs.ctypes.data, ys.ctypes.data,
                                      hy.ctypes.data, out.ctypes.data))
        return out

    def ncc_pair(self, ref, srcs, range0, range1, prm, xs, ys, hyps, pairs=None, grad_out=None,
                 backward=True):
        """Fixed-pair cost 1 - (NCC_s0 + NCC_s1)/2 of (pixel, plane) items and,
        with backward, its gradient w.r.t. (d, n0, n1, n2) times grad_out.
        pairs: (n, 2) source indices; None or rows of -1 select the best pair.
        Returns (cost, pairs, grad or None)."""
        r, arr, keep = self._views(ref, srcs)
        xs = np.ascontiguousarray(xs, np.int32)
        ys = np.ascontiguousarray(ys, np.int32)
        hy = np.ascontiguousarray(hyps, np.float32).reshape(-1, 4)
        n = len(xs)
        pr = (np.full((n, 2), -1, np.int32) if pairs is None
              else np.array(pairs, np.int32).reshape(n, 2).copy())
        go = None if grad_out is None else np.ascontiguousarray(grad_out, np.float32)
        out = np.zeros(n, np.float32)
        grad = np.zeros((n, 4), np.float32) if backward else None
        self.check(lib().mvs_ncc_pair(self._h, C.byref(r), arr, len(srcs), range0, range1,
                                      C.byref(prm), n, xs.ctypes.data, ys.ctypes.data,
                                      hy.ctypes.data, pr.ctypes.data, _ptr(go), out.ctypes.data,
                                      _ptr(grad)))
        return out, pr, grad

    def refine_terms(self, ref, srcs, range0, range1, prm, xs, ys, hyps, pairs=None):
        """The refinement kernels' own Gauss-Newton terms of (pixel, plane,
        pair) items: (cost, pairs, grad (n, 4) w.r.t. (d, n0, n1, n2), J^T J
        (n, 4, 4)). pairs: None or rows of -1 select the best pair."""
        r, arr, keep = self._views(ref, srcs)
        xs = np.ascontiguousarray(xs, np.int32)
        ys = np.ascontiguousarray(ys, np.int32)
        hy = np.ascontiguousarray(hyps, np.float32).reshape(-1, 4)
        n = len(xs)
        pr = (np.full((n, 2), -1, np.int32) if pairs is None
              else np.array(pairs, np.int32).reshape(n, 2).copy())
        out = np.zeros(n, np.float32)
        grad = np.zeros((n, 4), np.float32)
        j10 = np.zeros((n, 10), np.float32)
        self.check(lib().mvs_refine_terms(self._h, C.byref(r), arr, len(srcs), range0, range1,
                                          C.byref(prm), n, xs.ctypes.data, ys.ctypes.data,
                                          hy.ctypes.data, pr.ctypes.data, out.ctypes.data,
                                          grad.ctypes.data, j10.ctypes.data))
        iu = np.triu_indices(4)
        jtj = np.zeros((n, 4, 4), np.float32)
        jtj[:, iu[0], iu[1]] = j10
        jtj[:, iu[1], iu[0]] = j10
        return out, pr, grad, jtj

    def patchmatch_steps(self, ref, srcs, range0, range1, prm, first_step, n_steps, state=None):
        """Returns (state (h,w,5), textured, evals). first_step = -1 initialises."""
        r, arr, keep = self._views(ref, srcs)
        h, w = ref.h, ref.w
        st = (np.zeros((h, w, 5), np.float32) if state is None
              else np.array(state, np.float32, order="C", copy=True))
        tex = np.zeros((h, w), np.uint8)
        ev = C.c_int64()
        self.check(lib().mvs_patchmatch_steps(self._h, C.byref(r), arr, len(srcs), range0, range1,
                                              C.byref(prm), first_step, n_steps, st.ctypes.data,
                                              tex.ctypes.data, C.byref(ev)))
        return st, tex, ev.value

    def set_timing(self, on=True):
        self.check(lib().mvs_set_timing(self._h, 1 if on else 0))

    def stats(self):
        s = MvsStats()
        self.check(lib().mvs_get_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in MvsStats._fields_}

    def reset_stats(self):
        self.check(lib().mvs_reset_stats(self._h))

    def build_dsm(self, xyz, voxel_m, profile=1):
        """build_dsm (dsm.rs:128-290) on the GPU: (info dict, h (ny, nx) f32)."""
        p = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        d = MvsDsmInfo()
        self.check(lib().mvs_build_dsm(self._h, p.ctypes.data, len(p), voxel_m, profile,
                                       C.byref(d), None, 0))
        h = np.empty((d.ny, d.nx), np.float32)
        self.check(lib().mvs_dsm_download(self._h, h.ctypes.data, h.size))
        return d.as_dict(), h

    def prepare_view(self, rgb, max_side):
        """prepare_views' per-photo work on the GPU (dense.rs:100-112): an
        (h, w, 3) uint8 photo -> (working-size (h', w', 3) uint8 RGB, (h', w')
        f32 grey), the contents of a `.view` file."""
        img = np.ascontiguousarray(rgb, np.uint8)
        h, w = img.shape[:2]
        ow, oh, _ = prepared_size(w, h, max_side)
        grey = np.zeros((oh, ow), np.float32)
        out = np.zeros((oh, ow, 3), np.uint8)
        self.check(lib().mvs_prepare_view(self._h, img.ctypes.data, w, h, w * 3, max_side,
                                          grey.ctypes.data, out.ctypes.data))
        return out, grey

    def set_schedule(self, split=0, gpc=0, per_cu=16, autotune=True):
        self.check(lib().mvs_set_schedule(self._h, split, gpc, per_cu, 1 if autotune else 0))

    def set_graphs(self, enable=True):
        self.check(lib().mvs_set_graphs(self._h, 1 if enable else 0))

    def tuned_schedules(self):
        n = lib().mvs_tuned_schedules(self._h, None, 0)
        out = np.zeros((max(n, 1), 9), np.int32)
        lib().mvs_tuned_schedules(self._h, out.ctypes.data, n)
        keys = ("taps", "srcl", "pinhole", "log2px", "iteration", "mode", "split", "gpc",
                "per_cu")
        return [dict(zip(keys, map(int, r))) for r in out[:n]]

    def selftest_division(self, n, seed=1):
        bad = C.c_int64()
        self.check(lib().mvs_selftest(self._h, 0, n, seed, C.byref(bad)))
        return bad.value


class Survey:
    """Views resident in HBM; depth maps stay on the device."""

    def __init__(self, ctx, views, options=0):
        """options: MVS_SURVEY_LEAN_VIEWS | MVS_SURVEY_DEPTH_ONLY (residency
        for large surveys)."""
        self.ctx = ctx
        self._h = C.c_void_p()
        ctx.check(lib().mvs_survey_create(ctx.handle, len(views), C.byref(self._h)))
        self.options = options
        if options:
            ctx.check(lib().mvs_survey_set_options(self._h, options))
        self.shapes = [(v.h, v.w) for v in views]
        for i, v in enumerate(views):
            keep = []
            mv = make_view(v, keep)
            ctx.check(lib().mvs_survey_set_view(self._h, i, C.byref(mv)))

    def close(self):
        if self._h:
            lib().mvs_survey_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def depth_map(self, ref, srcs, range0, range1, prm):
        s = np.ascontiguousarray(srcs, np.int32)
        return self.ctx.check(lib().mvs_survey_depth_map(self._h, ref, s.ctypes.data, len(s),
                                                         range0, range1, C.byref(prm)),
                              (MVS_OK, MVS_NO_NEIGHBOURS))

    def depth_maps(self, jobs, prm, progress=None):
        """jobs: list of (ref, [srcs], (range0, range1)). prm: one MvsParams
        for all jobs, or a list with one per job."""
        refs = np.array([j[0] for j in jobs], np.int32)
        off = np.zeros(len(jobs) + 1, np.int32)
        for i, j in enumerate(jobs):
            off[i + 1] = off[i] + len(j[1])
        srcs = np.array([s for j in jobs for s in j[1]] or [0], np.int32)
        rng = np.array([x for j in jobs for x in j[2]], np.float64)
        cb = PROGRESS_FN(progress) if progress else PROGRESS_FN()
        if isinstance(prm, (list, tuple)):
            arr = (MvsParams * max(1, len(prm)))(*prm)
            self.ctx.check(lib().mvs_survey_depth_maps_ex(self._h, len(jobs), refs.ctypes.data,
                                                          off.ctypes.data, srcs.ctypes.data,
                                                          rng.ctypes.data, C.addressof(arr),
                                                          None, cb, None))
        else:
            self.ctx.check(lib().mvs_survey_depth_maps(self._h, len(jobs), refs.ctypes.data,
                                                       off.ctypes.data, srcs.ctypes.data,
                                                       rng.ctypes.data, C.byref(prm), None, cb,
                                                       None))

    def download(self, view):
        """(depth, normal, cost); normal and cost are None on a depth-only
        survey."""
        h, w = self.shapes[view]
        d = np.zeros((h, w), np.float32)
        if self.options & MVS_SURVEY_DEPTH_ONLY:
            self.ctx.check(lib().mvs_survey_download(self._h, view, d.ctypes.data, None, None))
            return d, None, None
        n = np.zeros((h, w, 3), np.float32)
        c = np.zeros((h, w), np.float32)
        self.ctx.check(lib().mvs_survey_download(self._h, view, d.ctypes.data, n.ctypes.data,
                                                 c.ctypes.data))
        return d, n, c

    def set_depth(self, view, depth):
        """A depth map computed earlier (the `.depth` cache) as the view's map."""
        h, w = self.shapes[view]
        d = np.ascontiguousarray(depth, np.float32)
        assert d.shape == (h, w)
        self.ctx.check(lib().mvs_survey_set_depth(self._h, view, d.ctypes.data))

    def set_colour(self, view, rgb):
        c = np.ascontiguousarray(rgb, np.uint8)
        self.ctx.check(lib().mvs_survey_set_colour(self._h, view, c.ctypes.data))

    def fuse(self, refs, nbrs, voxel, rule=MVS_CONS_REFERENCE):
        """Fuse the maps of refs; nbrs[r] lists ref r's neighbours. Returns the
        cloud (xyz (n, 3) f32, rgb (n, 3) u8) in voxel-index order."""
        refs_a = np.ascontiguousarray(refs, np.int32)
        off = np.zeros(len(refs) + 1, np.int32)
        for i, nb in enumerate(nbrs):
            off[i + 1] = off[i] + len(nb)
        flat = np.array([j for nb in nbrs for j in nb] or [0], np.int32)
        nvox = C.c_int64()
        self.ctx.check(lib().mvs_survey_fuse(self._h, len(refs_a), refs_a.ctypes.data,
                                             off.ctypes.data, flat.ctypes.data, rule, voxel,
                                             C.byref(nvox)))
        n = C.c_int64()
        self.ctx.check(lib().mvs_survey_fused_points(self._h, None, None, 0, C.byref(n)))
        xyz = np.zeros((n.value, 3), np.float32)
        rgb = np.zeros((n.value, 3), np.uint8)
        self.ctx.check(lib().mvs_survey_fused_points(self._h, xyz.ctypes.data, rgb.ctypes.data,
                                                     n.value, C.byref(n)))
        return xyz, rgb

    def consistency(self, ref, nbrs, rule=MVS_CONS_REFERENCE, download=True, noalloc=False):
        """(agree counts or None, number of consistent pixels). noalloc: the
        allocation-free entry point (mvs_survey_consistency_noalloc)."""
        h, w = self.shapes[ref]
        nb = np.ascontiguousarray(nbrs, np.int32)
        agree = np.zeros((h, w), np.int32) if download else None
        n = C.c_int64()
        fn = lib().mvs_survey_consistency_noalloc if noalloc else lib().mvs_survey_consistency
        self.ctx.check(fn(self._h, ref, nb.ctypes.data, len(nb), rule, _ptr(agree), C.byref(n)))
        return agree, n.value
