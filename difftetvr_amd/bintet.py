"""``.bintet`` mesh files (src/Tet/Loaders/BinTetLoader.cpp:37-67, the matching writer BinTetWriter::saveToFile
in the same file, :92-138).

Layout, little endian: u32 version (1, or 2 with cell colours); then length-prefixed arrays, each a u32 element
count followed by the elements: u32 cell indices (4 per tet), vec3 f32 positions, vec4 f32 vertex colours
[, vec4 f32 cell colours].
"""
import struct

import numpy as np


def _read_array(buf, off, dtype, width):
    (n,) = struct.unpack_from("<I", buf, off)
    off += 4
    nbytes = n * width * np.dtype(dtype).itemsize
    if off + nbytes > len(buf):
        raise RuntimeError("bintet: truncated file")
    a = np.frombuffer(buf, dtype=dtype, count=n * width, offset=off).reshape(n, width) if width > 1 else \
        np.frombuffer(buf, dtype=dtype, count=n, offset=off)
    return a.copy(), off + nbytes


def read(path):
    """Returns (cells int32 [T,4], positions float32 [V,3], vertex colours float32 [V,4] or None, cell colours
    float32 [T,4] or None) (BinTetLoader::loadFromFile, BinTetLoader.cpp:37-67).  Version 1 holds vertex
    colours; version 2 adds a cell colour array, and its vertex colour array may be empty."""
    with open(path, "rb") as f:
        buf = f.read()
    if len(buf) < 4:
        raise RuntimeError(f"bintet: '{path}' is too short")
    (version,) = struct.unpack_from("<I", buf, 0)
    if version not in (1, 2):
        raise RuntimeError(f"Error in BinTetLoader::loadFromFile: Invalid version number in file \"{path}\".")
    off = 4
    cells, off = _read_array(buf, off, "<u4", 1)
    pos, off = _read_array(buf, off, "<f4", 3)
    col, off = _read_array(buf, off, "<f4", 4)
    cell_col = None
    if version == 2:
        cell_col, off = _read_array(buf, off, "<f4", 4)
    if cells.size % 4:
        raise RuntimeError("bintet: cell index count is not a multiple of 4")
    ntets = cells.size // 4
    if cell_col is not None and cell_col.shape[0] == ntets and col.shape[0] == 0:
        col = None
    elif col.shape[0] != pos.shape[0]:
        raise RuntimeError("bintet: the vertex colour count differs from the vertex count")
    else:
        cell_col = None if cell_col is None or cell_col.shape[0] == 0 else cell_col
    if cell_col is not None and cell_col.shape[0] != ntets:
        raise RuntimeError("bintet: the cell colour count differs from the cell count")
    return (cells.astype(np.int32).reshape(-1, 4), pos.astype(np.float32),
            None if col is None else col.astype(np.float32), None if cell_col is None else cell_col.astype(np.float32))


def write(path, cells, positions, colors, cell_colors=None):
    """Writes a version-1 file (vertex colours), or version 2 when cell_colors is given (BinTetWriter::saveToFile,
    BinTetLoader.cpp:92-138; colors may then be None, written as an empty array)."""
    cells = np.ascontiguousarray(cells, "<u4").reshape(-1)
    pos = np.ascontiguousarray(positions, "<f4").reshape(-1, 3)
    col = np.zeros((0, 4), "<f4") if colors is None else np.ascontiguousarray(colors, "<f4").reshape(-1, 4)
    arrays = [cells, pos, col]
    if cell_colors is not None:
        arrays.append(np.ascontiguousarray(cell_colors, "<f4").reshape(-1, 4))
    with open(path, "wb") as f:
        f.write(struct.pack("<I", 1 if cell_colors is None else 2))
        for a in arrays:
            f.write(struct.pack("<I", a.shape[0]))
            f.write(a.tobytes())
