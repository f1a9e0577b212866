"""Coarse-to-fine mesh refinement on host arrays through the C ABI (``dtv_split_*``, host code, no device work).

``TetMesh.split_by_largest_gradient_magnitudes`` copies the mesh and its gradient to the host, calls
``split_by_largest_gradient_magnitudes`` here and uploads the refined mesh, as TetMesh.cpp:1786-1879 does.
"""
import ctypes

import numpy as np

from . import _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _mesh_arrays(cells, positions, colors):
    cells = np.ascontiguousarray(np.asarray(cells, dtype=np.int32).reshape(-1, 4))
    positions = np.ascontiguousarray(np.asarray(positions, dtype=np.float32).reshape(-1, 3))
    colors = np.ascontiguousarray(np.asarray(colors, dtype=np.float32).reshape(-1, 4))
    if colors.shape[0] != positions.shape[0]:
        raise RuntimeError("refine: colours must be per vertex ([V,4]).")
    return cells, positions, colors


def _take(lib, handle):
    try:
        t, v = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(lib.dtv_split_result_sizes(handle, ctypes.byref(t), ctypes.byref(v)), "dtv_split_result_sizes")
        cells = np.empty((t.value, 4), np.int32)
        pos = np.empty((v.value, 3), np.float32)
        col = np.empty((v.value, 4), np.float32)
        _lib.check(lib.dtv_split_result_copy(handle, _ptr(cells), _ptr(pos), _ptr(col)), "dtv_split_result_copy")
        return cells, pos, col
    finally:
        lib.dtv_split_result_destroy(handle)


def split_by_largest_gradient_magnitudes(cells, positions, colors, gradients, splits_ratio):
    """Refines at the ceil(splits_ratio * V) vertices whose gradient rows ([V,3] or [V,4]) are longest.
    Returns (cells int32 [T',4], positions float32 [V',3], colors float32 [V',4])."""
    lib = _lib.load()
    cells, positions, colors = _mesh_arrays(cells, positions, colors)
    g = np.ascontiguousarray(np.asarray(gradients, dtype=np.float32))
    if g.ndim != 2 or g.shape[0] != positions.shape[0] or g.shape[1] not in (3, 4):
        raise RuntimeError("split_by_largest_gradient_magnitudes: gradients must be [V,3] or [V,4].")
    h = ctypes.c_void_p()
    _lib.check(lib.dtv_split_by_largest_gradient_magnitudes(_ptr(cells), cells.shape[0], _ptr(positions),
                                                            _ptr(colors), positions.shape[0], _ptr(g), g.shape[1],
                                                            float(splits_ratio), ctypes.byref(h)),
               "dtv_split_by_largest_gradient_magnitudes")
    return _take(lib, h)


def subdivide_vertices(cells, positions, colors, gradient_magnitudes, num_splits):
    """TetMesh::subdivideVertices (TetMesh.cpp:1757-1784): the num_splits vertices of largest magnitude."""
    lib = _lib.load()
    cells, positions, colors = _mesh_arrays(cells, positions, colors)
    m = np.ascontiguousarray(np.asarray(gradient_magnitudes, dtype=np.float32).reshape(-1))
    if m.shape[0] != positions.shape[0]:
        raise RuntimeError("subdivide_vertices: one magnitude per vertex expected.")
    h = ctypes.c_void_p()
    _lib.check(lib.dtv_subdivide_vertices(_ptr(cells), cells.shape[0], _ptr(positions), _ptr(colors),
                                          positions.shape[0], _ptr(m), int(num_splits), ctypes.byref(h)),
               "dtv_subdivide_vertices")
    return _take(lib, h)
