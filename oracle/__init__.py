"""CPU oracle for the DiffTetVR tetrahedral volume-rendering hot path (ctypes wrapper over oracle.c).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this package, and only as the checker.  The product package ``difftetvr_amd`` never imports it.

The C sources restate the reference's INTERSECTION renderer; every routine cites the reference file:line it
follows.  ``build()`` compiles ``oracle.c`` with gcc (``oracle/Makefile``).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


def build():
    """Compile the oracle with gcc (always rebuilds when sources are newer)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    c = ctypes
    lib.orc_camera.argtypes = [_f32p, c.c_float, c.c_float, c.c_float, c.c_int, c.c_int, _f32p, _f64p]
    lib.orc_camera.restype = c.c_int
    lib.orc_build_fragments.argtypes = [_i32p, c.c_int64, _f32p, _f32p, c.c_float, c.c_int, c.c_int, c.c_int, c.c_int,
                                        c.c_int]
    lib.orc_build_fragments.restype = c.c_void_p
    lib.orc_frags_total.argtypes = [c.c_void_p]
    lib.orc_frags_total.restype = c.c_int64
    lib.orc_frags_num_tris.argtypes = [c.c_void_p]
    lib.orc_frags_num_tris.restype = c.c_int64
    lib.orc_frags_export.argtypes = [c.c_void_p, _i64p, _i32p]
    lib.orc_frags_export_tris.argtypes = [c.c_void_p, _u32p, _i32p]
    lib.orc_frags_free.argtypes = [c.c_void_p]
    vp = c.c_void_p  # optional arrays (None = NULL)
    lib.orc_forward_f.argtypes = [vp, _i32p, _f32p, vp, vp, _f32p, c.c_float, _f32p, c.c_int, c.c_float, c.c_int,
                                  _f32p, _f32p, _i32p]
    lib.orc_forward_d.argtypes = [vp, _i32p, _f64p, vp, vp, _f64p, c.c_double, _f64p, c.c_int, c.c_double, c.c_int,
                                  c.c_int, vp, _f64p, _f64p, _i32p]
    lib.orc_adjoint_exact_f.argtypes = [vp, _i32p, _f32p, vp, vp, _f32p, c.c_float, _f32p, c.c_int, c.c_float,
                                        c.c_int, _f32p, c.c_int, _f32p, _f32p]
    lib.orc_adjoint_exact_d.argtypes = [vp, _i32p, _f64p, vp, vp, _f64p, c.c_double, _f64p, c.c_int, c.c_double,
                                        c.c_int, c.c_int, _f64p, c.c_int, _f64p, _f64p]
    lib.orc_adjoint_reference_f.argtypes = [c.c_void_p, _i32p, _f32p, _f32p, _f32p, c.c_float, _f32p, _f32p, c.c_int,
                                            _f32p, _f32p]
    lib.orc_proj_fragments.argtypes = [_i32p, c.c_int64, _f32p, vp, vp, _f32p, c.c_float, c.c_int, c.c_int]
    lib.orc_proj_fragments.restype = c.c_void_p
    lib.orc_proj_forward_f.argtypes = [vp, _i32p, _f32p, vp, vp, _f32p, c.c_float, _f32p, _f32p, _f32p]
    lib.orc_proj_forward_d.argtypes = [vp, _i32p, _f64p, vp, vp, _f64p, c.c_double, _f64p, _f64p, _f64p]
    lib.orc_proj_tet_d.argtypes = [_f64p, _f64p, c.c_int, _f64p, _f32p, c.c_float, c.c_double, c.POINTER(c.c_int),
                                   c.POINTER(c.c_int), _i32p, _f64p, _f64p, _f64p]
    lib.orc_accumulate_linear_const_d.argtypes = [c.c_double, _f64p, _f64p, c.c_double, _f64p]
    lib.orc_ray_tet_d.argtypes = [_f64p, _f64p, _f64p, c.POINTER(c.c_int), c.POINTER(c.c_int),
                                  c.POINTER(c.c_double), c.POINTER(c.c_double)]
    lib.orc_ray_tet_d.restype = c.c_int
    lib.orc_ray_dir_f.argtypes = [_f32p, c.c_int, c.c_int, c.c_int, c.c_int, _f32p]
    lib.orc_adam_f.argtypes = [_f32p, _f32p, _f32p, _f32p, c.c_int64, c.c_double, c.c_double, c.c_double, c.c_double,
                               c.c_int64, c.c_int]
    lib.orc_mse_grad_f.argtypes = [_f32p, _f32p, c.c_int64, _f32p]
    lib.orc_mse_grad_f.restype = c.c_double
    lib.orc_tet_regularizer_f.argtypes = [_i32p, c.c_int64, _f32p, c.c_float, c.c_float, _f32p]
    lib.orc_tet_regularizer_d.argtypes = [_i32p, c.c_int64, _f64p, c.c_double, c.c_double, _f64p]
    lib.orc_tet_regularizer_loss_d.argtypes = [_i32p, c.c_int64, _f64p, c.c_double, c.c_double]
    lib.orc_tet_regularizer_loss_d.restype = c.c_double
    _lib = lib
    return lib


def camera_uniforms(view, fovy, znear, zfar, width, height):
    """Returns (float32[68], float64[68]) uniforms: viewProj, view, invProj, invView (column-major), camPos."""
    lib = _load()
    uf = np.zeros(68, np.float32)
    ud = np.zeros(68, np.float64)
    rc = lib.orc_camera(np.ascontiguousarray(view, np.float32), fovy, znear, zfar, width, height, uf, ud)
    if rc != 0:
        raise RuntimeError("singular view matrix")
    return uf, ud


class Fragments:
    """Ordered per-pixel fragment lists (fragment id = tet*4 + face), front to back."""

    def __init__(self, cells, positions, cam_f, znear, width, height, y_begin=0, y_end=None, y_step=1):
        self._lib = _load()
        self.cells = np.ascontiguousarray(cells, np.int32).reshape(-1)
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
        self.width, self.height = width, height
        self.handle = self._lib.orc_build_fragments(self.cells, self.cells.size // 4, pos,
                                                    np.ascontiguousarray(cam_f, np.float32), znear, width, height,
                                                    y_begin, height if y_end is None else y_end, y_step)

    def __del__(self):
        if getattr(self, "handle", None):
            self._lib.orc_frags_free(self.handle)
            self.handle = None

    @property
    def total(self):
        return self._lib.orc_frags_total(self.handle)

    def lists(self):
        P = self.width * self.height
        off = np.zeros(P + 1, np.int64)
        ids = np.zeros(max(self.total, 1), np.int32)
        self._lib.orc_frags_export(self.handle, off, ids)
        return off, ids[: self.total]

    def counts(self):
        off, _ = self.lists()
        return np.diff(off).astype(np.int32).reshape(self.height, self.width)

    def triangles(self):
        n = self._lib.orc_frags_num_tris(self.handle)
        keys = np.zeros(max(n, 1), np.uint32)
        ids = np.zeros(max(n, 1), np.int32)
        self._lib.orc_frags_export_tris(self.handle, keys, ids)
        return keys[:n], ids[:n]


def _colour_args(colors, cell_colors, dtype):
    """(per-vertex colours, per-cell colours) as NULL-able pointers, plus the arrays that keep them alive."""
    if (colors is None) == (cell_colors is None):
        raise ValueError("give exactly one of colors (per vertex) and cell_colors (per cell)")
    arr = np.ascontiguousarray(colors if cell_colors is None else cell_colors, dtype).reshape(-1)
    ptr = arr.ctypes.data_as(ctypes.c_void_p)
    return (ptr, None, arr) if cell_colors is None else (None, ptr, arr)


def forward(frags, positions, colors, cam, attenuation, clear=(0, 0, 0, 0), use_ert=False, ert_alpha=1.0 - 1e-4,
            dtype=np.float32, linear_depth=False, f32_decisions=False, decision_positions=None, cell_colors=None):
    """Renders; returns (image[H,W,4], ray colour C[H,W,4] before the clear blend, fragment count[H,W]).
    linear_depth: segment lengths as differences of linear eye depths (the PPLL renderer's) instead of the
    Euclidean lengths of the INTERSECTION renderer.  f32_decisions (float64 only; cam = the float64 uniforms):
    the ray-tet decisions of the FP32 evaluation, the continuous maths in FP64; decision_positions (float32
    [V,3]) takes those decisions at other positions (decisions frozen, e.g. for differences along one branch).
    cell_colors [T,4] (with colors None): one colour per tet (ColorStorage.PER_CELL)."""
    lib = _load()
    H, W = frags.height, frags.width
    img = np.zeros((H, W, 4), dtype)
    C = np.zeros((H, W, 4), dtype)
    cnt = np.zeros((H, W), np.int32)
    vcol, ccol, keep = _colour_args(colors, cell_colors, dtype)
    dp = None if decision_positions is None else np.ascontiguousarray(decision_positions, np.float32).reshape(-1)
    extra = () if dtype == np.float32 else (int(bool(f32_decisions)),
                                            None if dp is None else dp.ctypes.data_as(ctypes.c_void_p))
    fn = lib.orc_forward_f if dtype == np.float32 else lib.orc_forward_d
    fn(frags.handle, frags.cells, np.ascontiguousarray(positions, dtype).reshape(-1), vcol, ccol,
       np.ascontiguousarray(cam, dtype), attenuation, np.ascontiguousarray(clear, dtype), int(use_ert), ert_alpha,
       int(bool(linear_depth)), *extra, img, C, cnt)
    del keep
    return img, C, cnt


def adjoint_exact(frags, positions, colors, cam, attenuation, dimg, clear=(0, 0, 0, 0), use_ert=False,
                  ert_alpha=1.0 - 1e-4, use_abs_grad=False, dtype=np.float32, linear_depth=False, f32_decisions=False,
                  cell_colors=None):
    """Exact gradient of sum(dimg * image) w.r.t. vertex positions [V,3] and colours: [V,4] per vertex, or [T,4]
    per cell when cell_colors is given (colors None)."""
    lib = _load()
    V = np.asarray(positions).reshape(-1, 3).shape[0]
    vcol, ccol, keep = _colour_args(colors, cell_colors, dtype)
    gpos = np.zeros((V, 3), dtype)
    gcol = np.zeros((V if cell_colors is None else keep.size // 4, 4), dtype)
    fn = lib.orc_adjoint_exact_f if dtype == np.float32 else lib.orc_adjoint_exact_d
    fn(frags.handle, frags.cells, np.ascontiguousarray(positions, dtype).reshape(-1), vcol, ccol,
       np.ascontiguousarray(cam, dtype), attenuation, np.ascontiguousarray(clear, dtype), int(use_ert), ert_alpha,
       int(bool(linear_depth)), *(() if dtype == np.float32 else (int(bool(f32_decisions)),)),
       np.ascontiguousarray(dimg, dtype).reshape(-1), int(use_abs_grad), gpos.reshape(-1), gcol.reshape(-1))
    del keep
    return gpos, gcol


def adjoint_reference(frags, positions, colors, cam, attenuation, img, dimg, use_abs_grad=False):
    """The reference's own adjoint (AdjointIntersectRasterization.glsl), quirks kept, FP32."""
    lib = _load()
    V = np.asarray(positions).reshape(-1, 3).shape[0]
    gpos = np.zeros((V, 3), np.float32)
    gcol = np.zeros((V, 4), np.float32)
    lib.orc_adjoint_reference_f(frags.handle, frags.cells, np.ascontiguousarray(positions, np.float32).reshape(-1),
                                np.ascontiguousarray(colors, np.float32).reshape(-1),
                                np.ascontiguousarray(cam, np.float32), attenuation,
                                np.ascontiguousarray(img, np.float32).reshape(-1),
                                np.ascontiguousarray(dimg, np.float32).reshape(-1), int(use_abs_grad),
                                gpos.reshape(-1), gcol.reshape(-1))
    return gpos, gcol


class ProjFragments(Fragments):
    """The Projection renderer's ordered per-pixel lists (fragment id = tet * 4 + triangle), built with the float
    decisions of the mesh at `positions`: projected-tetrahedra triangles, centroid keys, sort, raster.  Kept
    fixed, they also freeze every discrete choice for finite differences of proj_forward."""

    def __init__(self, cells, positions, colors, cam_f, znear, width, height, cell_colors=None):
        self._lib = _load()
        self.cells = np.ascontiguousarray(cells, np.int32).reshape(-1)
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
        vcol, ccol, keep = _colour_args(colors, cell_colors, np.float32)
        self.width, self.height = width, height
        self.handle = self._lib.orc_proj_fragments(self.cells, self.cells.size // 4, pos, vcol, ccol,
                                                   np.ascontiguousarray(cam_f, np.float32), znear, width, height)
        del keep


def proj_forward(frags, positions, colors, cam, attenuation, clear=(0, 0, 0, 0), dtype=np.float32,
                 cell_colors=None):
    """Projection renderer image [H,W,4] and composite before the clear blend, over the lists of `frags`
    (ProjFragments), values at `positions` / colours in precision dtype (cam: uniforms of that precision)."""
    lib = _load()
    H, W = frags.height, frags.width
    img = np.zeros((H, W, 4), dtype)
    C = np.zeros((H, W, 4), dtype)
    vcol, ccol, keep = _colour_args(colors, cell_colors, dtype)
    fn = lib.orc_proj_forward_f if dtype == np.float32 else lib.orc_proj_forward_d
    fn(frags.handle, frags.cells, np.ascontiguousarray(positions, dtype).reshape(-1), vcol, ccol,
       np.ascontiguousarray(cam, dtype), attenuation, np.ascontiguousarray(clear, dtype), img.reshape(-1),
       C.reshape(-1))
    del keep
    return img, C


def proj_tet(P, C, cam_d, cam_f, znear, attenuation, per_cell=False):
    """One tet's projected-tetrahedra vertices in double (float decisions): dict with ntri, cls2, corners [4,3]
    (-1 beyond ntri), ndc [5,3], col [5,4], dep [5]."""
    c = ctypes
    ntri, cls2 = c.c_int(), c.c_int()
    corners = np.zeros(12, np.int32)
    ndc, col, dep = np.zeros(15), np.zeros(20), np.zeros(5)
    _load().orc_proj_tet_d(np.ascontiguousarray(P, np.float64).reshape(-1),
                           np.ascontiguousarray(C, np.float64).reshape(-1), int(per_cell),
                           np.ascontiguousarray(cam_d, np.float64), np.ascontiguousarray(cam_f, np.float32), znear,
                           attenuation, c.byref(ntri), c.byref(cls2), corners, ndc, col, dep)
    return {"ntri": ntri.value, "cls2": bool(cls2.value), "corners": corners.reshape(4, 3),
            "ndc": ndc.reshape(5, 3), "col": col.reshape(5, 4), "dep": dep}


def accumulate_linear_const(t, c0, c1, a):
    out = np.zeros(4, np.float64)
    _load().orc_accumulate_linear_const_d(t, np.asarray(c0, np.float64), np.asarray(c1, np.float64), a, out)
    return out


def ray_tet(ro, rd, verts):
    c = ctypes
    f0, f1 = c.c_int(), c.c_int()
    t0, t1 = c.c_double(), c.c_double()
    hit = _load().orc_ray_tet_d(np.asarray(ro, np.float64), np.asarray(rd, np.float64),
                                np.ascontiguousarray(verts, np.float64).reshape(-1), c.byref(f0), c.byref(f1),
                                c.byref(t0), c.byref(t1))
    return bool(hit), f0.value, f1.value, t0.value, t1.value


def ray_dir(cam_f, width, height, x, y):
    out = np.zeros(3, np.float32)
    _load().orc_ray_dir_f(np.ascontiguousarray(cam_f, np.float32), width, height, x, y, out)
    return out


def adam(param, grad, m, v, lr, step, betas=(0.9, 0.999), eps=1e-8, clamp_nonneg=False):
    """torch.optim.Adam single-tensor step (in place on param, m, v)."""
    _load().orc_adam_f(param.reshape(-1), np.ascontiguousarray(grad, np.float32).reshape(-1), m.reshape(-1),
                       v.reshape(-1), param.size, lr, betas[0], betas[1], eps, step, int(clamp_nonneg))


def mse_grad(img, tgt):
    img = np.ascontiguousarray(img, np.float32)
    d = np.zeros_like(img)
    loss = _load().orc_mse_grad_f(img.reshape(-1), np.ascontiguousarray(tgt, np.float32).reshape(-1), img.size,
                                  d.reshape(-1))
    return loss, d


def tet_regularizer_grad(cells, positions, reg_lambda, softplus_beta, dtype=np.float32):
    """Gradient of the tet-quality regulariser w.r.t. the vertex positions (TetRegularizer.glsl:108-166)."""
    cells = np.ascontiguousarray(cells, np.int32).reshape(-1)
    pos = np.ascontiguousarray(positions, dtype).reshape(-1)
    g = np.zeros_like(pos)
    fn = _load().orc_tet_regularizer_f if dtype == np.float32 else _load().orc_tet_regularizer_d
    fn(cells, cells.size // 4, pos, reg_lambda, softplus_beta, g)
    return g.reshape(-1, 3)


def tet_regularizer_loss(cells, positions, reg_lambda, softplus_beta):
    """sum over tets of (lambda / beta) log(1 + exp(-beta q)) in double precision."""
    cells = np.ascontiguousarray(cells, np.int32).reshape(-1)
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
    return _load().orc_tet_regularizer_loss_d(cells, cells.size // 4, pos, reg_lambda, softplus_beta)
