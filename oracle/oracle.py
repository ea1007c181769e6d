"""ctypes wrapper of the CPU oracle (``liboracle.so``) — TEST INFRASTRUCTURE ONLY.

Imported only by ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg
of ``bench.py``, always as the checker, never as the thing measured or shipped.
The C restatement it wraps follows crates/stitch-core/src/recon/dense.rs; see
``dense_oracle.c`` for the line-by-line citations.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("AERIAL_MVS_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


class OrcView(C.Structure):
    _fields_ = [("w", C.c_int32), ("h", C.c_int32), ("gray", C.c_void_p),
                ("f", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
                ("k1", C.c_double), ("k2", C.c_double),
                ("R", C.c_double * 9), ("C", C.c_double * 3)]


class OrcParams(C.Structure):
    _fields_ = [("taps", C.c_int32), ("stride", C.c_int32), ("iterations", C.c_int32),
                ("texture_min", C.c_float), ("cost_max", C.c_float)]


class OrcNbrDepth(C.Structure):
    _fields_ = [("view", C.POINTER(OrcView)), ("depth", C.c_void_p)]


_lib = None


class OrcDsm(C.Structure):
    _fields_ = [("x0", C.c_double), ("y0", C.c_double), ("cell", C.c_double),
                ("nx", C.c_int64), ("ny", C.c_int64), ("zmax", C.c_float)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `python oracle/build.py`")
        L = C.CDLL(LIB_PATH)
        L.orc_gray_lut.argtypes = [_f32p]
        L.orc_gray_rgb.argtypes = [C.c_uint8] * 3
        L.orc_gray_rgb.restype = C.c_float
        L.orc_rnd.argtypes = [_u64p]
        L.orc_rnd.restype = C.c_float
        L.orc_seed_init.argtypes = [C.c_uint64]
        L.orc_seed_init.restype = C.c_uint64
        L.orc_seed_refine.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_seed_refine.restype = C.c_uint64
        L.orc_random_normal.argtypes = [_u64p, _f32p, _f32p, C.c_float, _f32p]
        L.orc_distort.argtypes = [C.c_double] * 7 + [_f64p]
        L.orc_undistort.argtypes = [C.c_double] * 7 + [_f64p]
        L.orc_relative.argtypes = [_f64p, _f64p, _f64p, _f64p, _f32p, _f32p]
        L.orc_gray_at.argtypes = [_f32p, C.c_int32, C.c_int32, C.c_float, C.c_float,
                                  C.POINTER(C.c_float)]
        L.orc_gray_at.restype = C.c_int
        L.orc_mean_sd.argtypes = [_f32p, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_select_neighbours.argtypes = [_f64p, C.c_int32, C.c_int32, _i32p, C.c_int32,
                                            C.c_double, C.c_int32, _i32p]
        L.orc_select_neighbours.restype = C.c_int
        L.orc_ctx_create.argtypes = [C.POINTER(OrcView), C.POINTER(OrcView), C.c_int32,
                                     C.POINTER(OrcParams), C.c_double, C.c_double]
        L.orc_ctx_create.restype = C.c_void_p
        L.orc_ctx_destroy.argtypes = [C.c_void_p]
        L.orc_ctx_rays.argtypes = [C.c_void_p]
        L.orc_ctx_rays.restype = C.POINTER(C.c_float)
        L.orc_ctx_relpose.argtypes = [C.c_void_p, C.c_int32, _f32p, _f32p]
        L.orc_ctx_limits.argtypes = [C.c_void_p, _f32p]
        L.orc_cost.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _f32p]
        L.orc_cost.restype = C.c_float
        L.orc_cost_batch.argtypes = [C.c_void_p, C.c_int64, _i32p, _i32p, _f32p, _f32p, C.c_int32]
        L.orc_init.argtypes = [C.c_void_p, _f32p, _u8p, C.c_int32]
        L.orc_half_sweep.argtypes = [C.c_void_p, _f32p, _u8p, C.c_int32, C.c_int32, C.c_int32]
        L.orc_half_sweep.restype = C.c_int64
        L.orc_threshold.argtypes = [C.c_void_p, _f32p, _f32p]
        L.orc_depth_map.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _f32p, _u8p, _f32p]
        L.orc_depth_map.restype = C.c_int64
        L.orc_consistency.argtypes = [C.POINTER(OrcView), _f32p, C.POINTER(OrcNbrDepth),
                                      C.c_int32, _i32p, C.c_int32]
        L.orc_consistency.restype = C.c_int64
        L.orc_consistency_reproj.argtypes = [C.POINTER(OrcView), _f32p, C.POINTER(OrcNbrDepth),
                                             C.c_int32, C.c_int32, _i32p, C.c_int32]
        L.orc_consistency_reproj.restype = C.c_int64
        L.orc_pair_cost64.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _f64p, C.c_int32,
                                      C.c_int32, C.POINTER(C.c_double), _f64p,
                                      C.POINTER(C.c_uint64)]
        L.orc_pair_cost64.restype = C.c_int
        L.orc_pair_terms64.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _f64p, C.c_int32,
                                       C.c_int32, C.POINTER(C.c_double), _f64p, _f64p,
                                       C.POINTER(C.c_uint64)]
        L.orc_pair_terms64.restype = C.c_int
        L.orc_pair_unit64.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _f64p, C.c_int32, _f64p,
                                      C.POINTER(C.c_uint64)]
        L.orc_pair_unit64.restype = C.c_int
        L.orc_fuse_points.argtypes = [C.POINTER(OrcView), _f32p, C.POINTER(OrcNbrDepth),
                                      C.c_int32, C.c_int32, C.c_int32, C.c_double, _i32p,
                                      _f64p, _i32p]
        L.orc_fuse_points.restype = C.c_int64
        L.orc_fusion_voxel.argtypes = [_f64p, _f64p, C.c_int64, _f64p, C.c_int32, C.c_double,
                                       C.c_int32, C.c_int32, C.c_int32, C.c_int64]
        L.orc_fusion_voxel.restype = C.c_double
        _i32o = C.POINTER(C.c_int32)
        L.orc_shrink_size.argtypes = [C.c_int32, C.c_int32, C.c_int32, _i32o, _i32o, _i32o,
                                      _i32o, _i32o, C.POINTER(C.c_double)]
        L.orc_box_down.argtypes = [_u8p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _u8p]
        L.orc_resize_bilinear.argtypes = [_u8p, C.c_int32, C.c_int32, C.c_int32, _u8p,
                                          C.c_int32, C.c_int32]
        L.orc_prepare_view.argtypes = [_u8p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _u8p,
                                       _f32p]
        L.orc_build_dsm.argtypes = [_f32p, C.c_int64, C.c_double, C.c_int64,
                                    C.POINTER(OrcDsm), C.POINTER(C.POINTER(C.c_float))]
        L.orc_build_dsm.restype = C.c_int
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_remove_islands.argtypes = [_f32p, C.c_int64, C.c_int64, C.c_int64]
        L.orc_fill_enclosed.argtypes = [_f32p, C.c_int64, C.c_int64]
        L.orc_dsm_height.argtypes = [C.POINTER(OrcDsm), _f32p, C.c_double, C.c_double,
                                     C.POINTER(C.c_float)]
        L.orc_dsm_height.restype = C.c_int
        _lib = L
    return _lib


def gray_lut():
    out = np.zeros(256, np.float32)
    lib().orc_gray_lut(out)
    return out


def gray_rgb(r, g, b):
    return lib().orc_gray_rgb(r, g, b)


def rnd_stream(seed, n):
    """n successive draws of the reference's xorshift from `seed`; returns
    (draws as float32, final state)."""
    s = np.array([seed], np.uint64)
    out = np.array([lib().orc_rnd(s) for _ in range(n)], np.float32)
    return out, int(s[0])


def seed_init(i):
    return lib().orc_seed_init(i)


def seed_refine(i, it, parity):
    return lib().orc_seed_refine(i, it, parity)


def random_normal(seed, ray, around, spread):
    s = np.array([seed], np.uint64)
    out = np.zeros(3, np.float32)
    lib().orc_random_normal(s, np.asarray(ray, np.float32), np.asarray(around, np.float32),
                            spread, out)
    return out, int(s[0])


def distort(f, cx, cy, k1, k2, x, y):
    out = np.zeros(2)
    lib().orc_distort(f, cx, cy, k1, k2, x, y, out)
    return out


def undistort(f, cx, cy, k1, k2, u, v):
    out = np.zeros(2)
    lib().orc_undistort(f, cx, cy, k1, k2, u, v, out)
    return out


def relative(Rr, Cr, Rs, Cs):
    R = np.zeros(9, np.float32)
    t = np.zeros(3, np.float32)
    lib().orc_relative(np.ascontiguousarray(Rr, np.float64).ravel(), np.asarray(Cr, np.float64),
                       np.ascontiguousarray(Rs, np.float64).ravel(), np.asarray(Cs, np.float64),
                       R, t)
    return R.reshape(3, 3), t


def gray_at(gray, x, y):
    g = np.ascontiguousarray(gray, np.float32)
    out = C.c_float()
    ok = lib().orc_gray_at(g.ravel(), g.shape[1], g.shape[0], x, y, C.byref(out))
    return out.value if ok else None


def mean_sd(v):
    v = np.ascontiguousarray(v, np.float32)
    m, s = C.c_float(), C.c_float()
    lib().orc_mean_sd(v, len(v), C.byref(m), C.byref(s))
    return m.value, s.value


def select_neighbours(centres, ci, neighbours, range0, take):
    centres = np.ascontiguousarray(centres, np.float64)
    nb = np.ascontiguousarray(neighbours, np.int32)
    out = np.zeros(max(len(nb), 1), np.int32)
    k = lib().orc_select_neighbours(centres.ravel(), len(centres), ci, nb, len(nb), range0, take,
                                    out)
    return out[:k].tolist()


def _orc_view(v, gray_f32):
    ov = OrcView()
    ov.w = v.w
    ov.h = v.h
    ov.gray = gray_f32.ctypes.data if gray_f32 is not None else None
    ov.f, ov.cx, ov.cy, ov.k1, ov.k2 = v.f, v.cx, v.cy, v.k1, v.k2
    ov.R[:] = [float(a) for a in np.asarray(v.R, np.float64).ravel()]
    ov.C[:] = [float(a) for a in np.asarray(v.C, np.float64).ravel()]
    return ov


def view_gray_f32(v):
    """The reference's f32 grey plane of a view (dense.rs:108 for grey u8)."""
    g = getattr(v, "gray_f32", None)
    if g is not None:
        return np.ascontiguousarray(g, np.float32)
    return gray_lut()[np.asarray(v.gray)]


class OracleDepth:
    """One reference view with its sources: the state of depth_map
    (dense.rs:316-479) on the CPU."""

    def __init__(self, ref, srcs, range0, range1, taps=5, stride=2, iterations=4,
                 texture_min=0.008, cost_max=0.45):
        self._keep = []
        g = view_gray_f32(ref)
        self._keep.append(g)
        self.ref = _orc_view(ref, g)
        arr = (OrcView * max(len(srcs), 1))()
        for i, s in enumerate(srcs):
            gs = view_gray_f32(s)
            self._keep.append(gs)
            arr[i] = _orc_view(s, gs)
        self._srcs = arr
        p = OrcParams(taps, stride, iterations, texture_min, cost_max)
        self.params = p
        self.w, self.h = ref.w, ref.h
        self.iterations = iterations
        self.ctx = lib().orc_ctx_create(C.byref(self.ref), arr, len(srcs), C.byref(p), range0,
                                        range1)
        if not self.ctx:
            raise ValueError("oracle context rejected the parameters")

    def __del__(self):
        if getattr(self, "ctx", None):
            lib().orc_ctx_destroy(self.ctx)
            self.ctx = None

    def rays(self):
        p = lib().orc_ctx_rays(self.ctx)
        return np.ctypeslib.as_array(p, shape=(self.h, self.w, 2)).copy()

    def relpose(self, s):
        R = np.zeros(9, np.float32)
        t = np.zeros(3, np.float32)
        lib().orc_ctx_relpose(self.ctx, s, R, t)
        return R.reshape(3, 3), t

    def limits(self):
        out = np.zeros(4, np.float32)
        lib().orc_ctx_limits(self.ctx, out)
        return out

    def cost(self, xs, ys, hyps, nthreads=8):
        xs = np.ascontiguousarray(xs, np.int32)
        ys = np.ascontiguousarray(ys, np.int32)
        hy = np.ascontiguousarray(hyps, np.float32).reshape(-1, 4)
        out = np.zeros(len(xs), np.float32)
        lib().orc_cost_batch(self.ctx, len(xs), xs, ys, hy.ravel(), out, nthreads)
        return out

    def pair_cost64(self, x, y, hyp, s0, s1):
        """f64 fixed-pair cost 1 - (NCC_s0 + NCC_s1)/2, its analytic gradient
        w.r.t. (d, n0, n1, n2) and the bilinear-cell signature of every sample;
        None where the cost is undefined. (s, s) gives 1 - NCC_s."""
        h = np.ascontiguousarray(hyp, np.float64).reshape(4)
        c = C.c_double(0.0)
        g = np.zeros(4, np.float64)
        cell = C.c_uint64(0)
        ok = lib().orc_pair_cost64(self.ctx, int(x), int(y), h, int(s0), int(s1), C.byref(c), g,
                                   C.byref(cell))
        if not ok:
            return None
        return c.value, g, cell.value

    def pair_terms64(self, x, y, hyp, s0, s1):
        """f64 fixed-pair cost, gradient w.r.t. (d, n0, n1, n2), the 4 x 4
        Gauss-Newton matrix J^T J of the pair's stacked residuals
        a/|a| - b_s/|b_s| and the cell signature; None where undefined."""
        h = np.ascontiguousarray(hyp, np.float64).reshape(4)
        c = C.c_double(0.0)
        g = np.zeros(4, np.float64)
        jtj = np.zeros(16, np.float64)
        cell = C.c_uint64(0)
        ok = lib().orc_pair_terms64(self.ctx, int(x), int(y), h, int(s0), int(s1), C.byref(c),
                                    g, jtj, C.byref(cell))
        if not ok:
            return None
        return c.value, g, jtj.reshape(4, 4), cell.value

    def pair_unit64(self, x, y, hyp, s):
        """Unit centred sample window b_s/|b_s| of source s (f64, tap order)
        and the cell signature; None where undefined."""
        h = np.ascontiguousarray(hyp, np.float64).reshape(4)
        u = np.zeros(self.params.taps * self.params.taps, np.float64)
        cell = C.c_uint64(0)
        if not lib().orc_pair_unit64(self.ctx, int(x), int(y), h, int(s), u, C.byref(cell)):
            return None
        return u, cell.value

    def pair_costs64(self, xs, ys, hyps, pairs):
        """Vector form of pair_cost64: (cost (nan if undefined), grad (n, 4))."""
        n = len(xs)
        cost = np.full(n, np.nan)
        grad = np.full((n, 4), np.nan)
        for i in range(n):
            r = self.pair_cost64(xs[i], ys[i], hyps[i], pairs[i][0], pairs[i][1])
            if r is not None:
                cost[i], grad[i] = r[0], r[1]
        return cost, grad

    def init(self, nthreads=8):
        st = np.zeros((self.h, self.w, 5), np.float32)
        tex = np.zeros((self.h, self.w), np.uint8)
        lib().orc_init(self.ctx, st.ravel(), tex.ravel(), nthreads)
        return st, tex

    def half_sweep(self, state, textured, it, parity, nthreads=8):
        assert state.flags.c_contiguous and state.dtype == np.float32
        return lib().orc_half_sweep(self.ctx, state.ravel(), np.ascontiguousarray(textured).ravel(),
                                    it, parity, nthreads)

    def threshold(self, state):
        d = np.zeros((self.h, self.w), np.float32)
        lib().orc_threshold(self.ctx, np.ascontiguousarray(state).ravel(), d.ravel())
        return d

    def depth_map(self, iterations=None, nthreads=8):
        it = self.iterations if iterations is None else iterations
        st = np.zeros((self.h, self.w, 5), np.float32)
        tex = np.zeros((self.h, self.w), np.uint8)
        d = np.zeros((self.h, self.w), np.float32)
        evals = lib().orc_depth_map(self.ctx, it, nthreads, st.ravel(), tex.ravel(), d.ravel())
        return d, st, tex, evals


def consistency(ref, depth, nbrs, nthreads=8, rule=0, need=3):
    """Agree counts of the fusion filter (dense.rs:585-615; rule 0) or of the
    reprojection rule (rule 1: <= 1 px, consistent when agree >= need). nbrs:
    list of (view, depth map). Returns (agree int32 map, number consistent)."""
    rv = _orc_view(ref, None)
    keep = []
    arr = (OrcNbrDepth * max(len(nbrs), 1))()
    views = (OrcView * max(len(nbrs), 1))()
    for i, (v, d) in enumerate(nbrs):
        d = np.ascontiguousarray(d, np.float32)
        keep.append(d)
        views[i] = _orc_view(v, None)
        arr[i].view = C.pointer(views[i])
        arr[i].depth = d.ctypes.data
    depth = np.ascontiguousarray(depth, np.float32)
    agree = np.zeros(depth.shape, np.int32)
    if rule == 0:
        n = lib().orc_consistency(C.byref(rv), depth.ravel(), arr, len(nbrs), agree.ravel(),
                                  nthreads)
    else:
        n = lib().orc_consistency_reproj(C.byref(rv), depth.ravel(), arr, len(nbrs), need,
                                         agree.ravel(), nthreads)
    return agree, n


def _nbr_arrays(nbrs):
    keep = []
    arr = (OrcNbrDepth * max(len(nbrs), 1))()
    views = (OrcView * max(len(nbrs), 1))()
    for i, (v, d) in enumerate(nbrs):
        d = np.ascontiguousarray(d, np.float32)
        keep.append(d)
        views[i] = _orc_view(v, None)
        arr[i].view = C.pointer(views[i])
        arr[i].depth = d.ctypes.data
    return arr, views, keep


def fuse_points(ref, depth, nbrs, rule=0, need=None, voxel=0.1):
    """Fusion's per-pixel emission for one reference view (dense.rs:589-632):
    (keys (n, 3) int32, world points (n, 3) f64, pixel indices (n,))."""
    if need is None:
        need = (2 if len(nbrs) >= 3 else 1) if rule == 0 else 3
    rv = _orc_view(ref, None)
    arr, views, keep = _nbr_arrays(nbrs)
    depth = np.ascontiguousarray(depth, np.float32)
    n = depth.size
    keys = np.zeros(3 * n, np.int32)
    xw = np.zeros(3 * n, np.float64)
    pix = np.zeros(n, np.int32)
    m = lib().orc_fuse_points(C.byref(rv), depth.ravel(), arr, len(nbrs), rule, need,
                              1.0 / voxel, keys, xw, pix)
    return keys[:3 * m].reshape(-1, 3), xw[:3 * m].reshape(-1, 3), pix[:m]


def fusion_voxel(xs, ys, ranges, f, size, width, height, max_points):
    """dense.rs:545-562."""
    xs = np.ascontiguousarray(xs, np.float64)
    ys = np.ascontiguousarray(ys, np.float64)
    rg = np.ascontiguousarray(ranges, np.float64).ravel()
    return lib().orc_fusion_voxel(xs, ys, len(xs), rg, len(rg) // 2, f, size, width, height,
                                  max_points)


# ---- view preparation (dense.rs:92-121, imgops.rs:89-171) --------------------

def shrink_size(w, h, max_side):
    """imgops.rs:145-171: box factor (0 = none), boxed size, output size, scale."""
    v = [C.c_int32() for _ in range(5)]
    sc = C.c_double()
    lib().orc_shrink_size(w, h, max_side, *[C.byref(x) for x in v], C.byref(sc))
    f, bw, bh, ow, oh = (x.value for x in v)
    return dict(f=f, bw=bw, bh=bh, w=ow, h=oh, scale=sc.value)


def box_down(img, f):
    """imgops.rs:89-118 on an (h, w, 3) uint8 image."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape[:2]
    out = np.zeros((h // f, w // f, 3), np.uint8)
    lib().orc_box_down(img.ravel(), w, h, w * 3, f, out.ravel())
    return out


def resize_bilinear(img, w, h):
    """imgops.rs:120-143 on an (h, w, 3) uint8 image."""
    img = np.ascontiguousarray(img, np.uint8)
    sh, sw = img.shape[:2]
    out = np.zeros((h, w, 3), np.uint8)
    lib().orc_resize_bilinear(img.ravel(), sw, sh, sw * 3, out.ravel(), w, h)
    return out


def prepare_view(img, max_side):
    """dense.rs:100-112: (working-size RGB8 (h, w, 3), grey f32 (h, w))."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape[:2]
    z = shrink_size(w, h, max_side)
    rgb = np.zeros((z["h"], z["w"], 3), np.uint8)
    grey = np.zeros((z["h"], z["w"]), np.float32)
    lib().orc_prepare_view(img.ravel(), w, h, w * 3, max_side, rgb.ravel(), grey.ravel())
    return rgb, grey


# ---- DSM (recon/dsm.rs) ------------------------------------------------------

DSM_BUDGET = {0: 3_000_000, 1: 8_000_000, 2: 20_000_000}  # dsm.rs:122-128


def build_dsm(xyz, voxel_m, budget):
    """dsm.rs:128-290 -> (dict(x0, y0, cell, nx, ny, zmax), h (ny, nx) f32), or
    None for fewer than 100 points."""
    p = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    d = OrcDsm()
    hp = C.POINTER(C.c_float)()
    if lib().orc_build_dsm(p.ravel(), len(p), voxel_m, budget, C.byref(d), C.byref(hp)):
        return None
    n = d.nx * d.ny
    h = np.ctypeslib.as_array(hp, shape=(n,)).copy().reshape(d.ny, d.nx)
    lib().orc_free(hp)
    return dict(x0=d.x0, y0=d.y0, cell=d.cell, nx=d.nx, ny=d.ny, zmax=d.zmax), h


def remove_islands(h, min_cells):
    g = np.ascontiguousarray(h, np.float32).copy()
    lib().orc_remove_islands(g.ravel(), g.shape[1], g.shape[0], min_cells)
    return g


def fill_enclosed(h):
    g = np.ascontiguousarray(h, np.float32).copy()
    lib().orc_fill_enclosed(g.ravel(), g.shape[1], g.shape[0])
    return g


def dsm_height(info, h, x, y):
    d = OrcDsm(info["x0"], info["y0"], info["cell"], info["nx"], info["ny"], info["zmax"])
    out = C.c_float()
    g = np.ascontiguousarray(h, np.float32)
    ok = lib().orc_dsm_height(C.byref(d), g.ravel(), x, y, C.byref(out))
    return out.value if ok else None
