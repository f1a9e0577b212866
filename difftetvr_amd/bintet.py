"""``.bintet`` mesh files (src/Tet/Loaders/BinTetLoader.cpp:37-66, the matching writer at
src/Tet/Writers/BinTetWriter.cpp).

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
    """Returns (cells int32 [T,4], positions float32 [V,3], vertex colours float32 [V,4])."""
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
    if cells.size % 4:
        raise RuntimeError("bintet: cell index count is not a multiple of 4")
    if col.shape[0] != pos.shape[0]:
        raise RuntimeError("bintet: only per-vertex colours are supported")
    return cells.astype(np.int32).reshape(-1, 4), pos.astype(np.float32), col.astype(np.float32)


def write(path, cells, positions, colors):
    """Writes a version-1 file (per-vertex colours)."""
    cells = np.ascontiguousarray(cells, "<u4").reshape(-1)
    pos = np.ascontiguousarray(positions, "<f4").reshape(-1, 3)
    col = np.ascontiguousarray(colors, "<f4").reshape(-1, 4)
    with open(path, "wb") as f:
        f.write(struct.pack("<I", 1))
        f.write(struct.pack("<I", cells.size))
        f.write(cells.tobytes())
        f.write(struct.pack("<I", pos.shape[0]))
        f.write(pos.tobytes())
        f.write(struct.pack("<I", col.shape[0]))
        f.write(col.tobytes())
