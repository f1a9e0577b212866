"""Synthetic meshes, colour fields and cameras for the benchmark and test workloads (host-side numpy).

* ``hex_grid`` restates TetMesh::setHexMeshConst (src/Tet/TetMesh.cpp:1882-1936) with HEX_TO_TET_TABLE
  (:104-117): vertices x-fastest, 6 tets per cell.
* ``single_tetrahedron`` / ``cube_central_gradient`` are the built-in test meshes of TetMesh::loadTestData
  (:969-1077).
* ``look_at_view_matrix`` restates pytests/datasets/sample_view.py:98-130 (camera looks at the origin,
  global up +y, column-major output).
* ``delaunay_ball`` tetrahedralises seeded points in the unit ball (scipy Qhull) and reorients every tet to the
  reference convention dot(cross(p1-p0, p2-p0), p3-p0) < 0 (TetMesh.cpp:532-541).
All randomness comes from ``numpy.random.default_rng(seed)``.
"""
import math

import numpy as np

HEX_TO_TET_TABLE = np.array([[0, 4, 7, 6], [0, 4, 6, 5], [0, 3, 6, 7], [0, 3, 2, 6], [0, 1, 5, 6], [0, 1, 6, 2]],
                            np.int64)


def hex_grid(aabb_min, aabb_max, xs, ys, zs):
    """Returns (positions float32 [xs*ys*zs, 3], cells int32 [6*(xs-1)*(ys-1)*(zs-1), 4])."""
    mn = np.asarray(aabb_min, np.float32)
    mx = np.asarray(aabb_max, np.float32)
    ix = np.arange(xs, dtype=np.float32) / np.float32(xs - 1)
    iy = np.arange(ys, dtype=np.float32) / np.float32(ys - 1)
    iz = np.arange(zs, dtype=np.float32) / np.float32(zs - 1)
    px = mn[0] + ix * (mx[0] - mn[0])
    py = mn[1] + iy * (mx[1] - mn[1])
    pz = mn[2] + iz * (mx[2] - mn[2])
    Z, Y, X = np.meshgrid(pz, py, px, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1).astype(np.float32)

    def idx(x, y, z):
        return z * xs * ys + y * xs + x

    cz, cy, cx = np.meshgrid(np.arange(zs - 1), np.arange(ys - 1), np.arange(xs - 1), indexing="ij")
    cx, cy, cz = cx.ravel(), cy.ravel(), cz.ravel()
    hexv = np.stack([idx(cx, cy, cz), idx(cx + 1, cy, cz), idx(cx + 1, cy + 1, cz), idx(cx, cy + 1, cz),
                     idx(cx, cy, cz + 1), idx(cx + 1, cy, cz + 1), idx(cx + 1, cy + 1, cz + 1),
                     idx(cx, cy + 1, cz + 1)], axis=1)
    cells = hexv[:, HEX_TO_TET_TABLE].reshape(-1, 4).astype(np.int32)
    return pos, cells


# Face f of a tet uses its vertices TET_FACE_TABLE[f] (Common/TetFaceTable.glsl:29-34, TetMesh.cpp:1175-1181).
TET_FACE_TABLE = np.array([[0, 1, 2], [1, 0, 3], [0, 2, 3], [2, 1, 3]], np.int64)
_TET_FACES = TET_FACE_TABLE


def boundary_vertex_mask(cells, num_vertices):
    """Vertices on faces used by exactly one tet (buildFacesSlim, TetMesh.cpp:1218-1268)."""
    cells = np.asarray(cells, np.int64)
    faces = np.sort(cells[:, _TET_FACES].reshape(-1, 3), axis=1)
    _, inv, cnt = np.unique(faces, axis=0, return_inverse=True, return_counts=True)
    bfaces = faces[cnt[inv.ravel()] == 1]
    mask = np.zeros(num_vertices, bool)
    mask[bfaces.ravel()] = True
    return mask


def tet_orientation(positions, cells):
    p = np.asarray(positions, np.float64)[np.asarray(cells, np.int64)]
    return np.einsum("ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0])


def jitter_interior(positions, cells, cell_size, amount=0.2, seed=0):
    """Moves interior vertices by U(-amount, amount) * cell_size per axis (boundary vertices fixed)."""
    rng = np.random.default_rng(seed)
    pos = np.array(positions, np.float32, copy=True)
    interior = ~boundary_vertex_mask(cells, pos.shape[0])
    d = rng.uniform(-amount, amount, size=(int(interior.sum()), 3)) * cell_size
    pos[interior] += d.astype(np.float32)
    return pos


def blob_colors(positions, seed=0, n_blobs=4, density_max=20.0):
    """Per-vertex RGBA: RGB = clip(sum of n_blobs seeded Gaussian blobs, 0, 1); alpha = density in
    [0, density_max] from the same blobs' summed weight."""
    rng = np.random.default_rng(seed)
    p = np.asarray(positions, np.float64)
    centers = rng.uniform(-0.35, 0.35, size=(n_blobs, 3))
    sigmas = rng.uniform(0.12, 0.3, size=n_blobs)
    rgb = rng.uniform(0.1, 1.0, size=(n_blobs, 3))
    w = np.exp(-np.sum((p[:, None, :] - centers[None]) ** 2, axis=2) / (2.0 * sigmas[None] ** 2))
    col = np.clip(w @ rgb, 0.0, 1.0)
    dens = density_max * np.clip(w.sum(axis=1) / max(w.sum(axis=1).max(), 1e-12), 0.0, 1.0)
    return np.concatenate([col, dens[:, None]], axis=1).astype(np.float32)


def value_noise(positions, seed=0, octaves=3, lattice=8, vmax=40.0):
    """Seeded 3-octave trilinear value noise on [-1,1]^3 scaled to [0, vmax]."""
    rng = np.random.default_rng(seed)
    p = (np.asarray(positions, np.float64) + 1.0) * 0.5
    out = np.zeros(p.shape[0])
    amp, total = 1.0, 0.0
    for o in range(octaves):
        n = lattice * (2 ** o)
        grid = rng.uniform(0.0, 1.0, size=(n + 2, n + 2, n + 2))
        q = np.clip(p, 0.0, 1.0) * n
        i0 = np.floor(q).astype(np.int64)
        f = q - i0
        val = 0.0
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    wgt = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * \
                          (f[:, 2] if dz else 1 - f[:, 2])
                    val = val + wgt * grid[i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz]
        out += amp * val
        total += amp
        amp *= 0.5
    return (vmax * out / total).astype(np.float32)


def delaunay_ball(n_points, seed=0):
    """Delaunay tetrahedralisation of n seeded uniform points in the unit ball, reference orientation."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n_points, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.uniform(0.0, 1.0, size=n_points) ** (1.0 / 3.0)
    pts = (d * r[:, None]).astype(np.float32)
    tri = Delaunay(pts.astype(np.float64))
    cells = tri.simplices.astype(np.int32)
    o = tet_orientation(pts, cells)
    flip = o > 0
    cells[flip] = cells[flip][:, [0, 2, 1, 3]]
    keep = np.abs(o) > 1e-12
    return pts, np.ascontiguousarray(cells[keep])


def single_tetrahedron():
    """TestCase.SINGLE_TETRAHEDRON (TetMesh.cpp:1016-1061)."""
    cells = np.array([[0, 1, 2, 3]], np.int32)
    pos = np.array([[-0.1, -0.1, 0.1], [0.1, -0.1, 0.1], [0.0, 0.3, 0.0], [0.1, -0.1, -0.1]], np.float32)
    col = np.array([[1, 1, 1, 0.1], [1, 1, 1, 0.1], [0, 1, 0, 0.1], [1, 1, 1, 0.8]], np.float32)
    return pos, cells, col


def cube_central_gradient():
    """TestCase.CUBE_CENTRAL_GRADIENT (TetMesh.cpp:969-1014, 1062-1075): 4 hexes, 24 tets, scaled by 1/16."""
    width = 4.0
    ml, mr = -width * 0.5, width * 0.5
    P = [[-4, -4, 4], [ml, -4, 4], [0, -4, 4], [mr, -4, 4], [4, -4, 4],
         [4, -4, -4], [mr, -4, -4], [0, -4, -4], [ml, -4, -4], [-4, -4, -4],
         [-4, 4, 4], [ml, 4, 4], [0, 4, 4], [mr, 4, 4], [4, 4, 4],
         [4, 4, -4], [mr, 4, -4], [0, 4, -4], [ml, 4, -4], [-4, 4, -4]]
    pos = np.array(P, np.float32) * np.float32(1.0 / 16.0)
    a = 0.25
    c0 = [0.5, 0.5, 0.5, 0.0625 * a]
    c1 = [1.0, 0.0, 0.0, 0.125 * a]
    col = np.array([c0, c0, c1, c0, c0] * 4, np.float32)
    hexes = np.array([[9, 8, 18, 19, 0, 1, 11, 10], [8, 7, 17, 18, 1, 2, 12, 11],
                      [7, 6, 16, 17, 2, 3, 13, 12], [6, 5, 15, 16, 3, 4, 14, 13]], np.int64)
    cells = hexes[:, HEX_TO_TET_TABLE].reshape(-1, 4).astype(np.int32)
    return pos, cells, col


def look_at_view_matrix(eye, up=(0.0, 1.0, 0.0)):
    """Column-major view matrix (16 floats) of a camera at `eye` looking at the origin
    (pytests/datasets/sample_view.py:114-130)."""
    eye = np.asarray(eye, np.float64)
    fwd = eye / np.linalg.norm(eye)
    right = np.cross(np.asarray(up, np.float64), fwd)
    right /= np.linalg.norm(right)
    cup = np.cross(fwd, right)
    cup /= np.linalg.norm(cup)
    rot = np.eye(4)
    rot[:3, 0], rot[:3, 1], rot[:3, 2] = right, cup, fwd
    trans = np.eye(4)
    trans[:3, 3] = eye
    view = np.linalg.inv(trans @ rot)
    return view.T.reshape(-1).astype(np.float32)  # element [i*4+j] = view[j, i]


def fibonacci_sphere(n, radius):
    """n points on a sphere (golden-angle spiral, half-step offsets so no point sits on the +-y poles)."""
    i = np.arange(n, dtype=np.float64) + 0.5
    y = 1.0 - 2.0 * i / n
    r = np.sqrt(1.0 - y * y)
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    return radius * np.stack([r * np.cos(phi), y, r * np.sin(phi)], axis=1)


def grid_workload(n, seed=0, jitter=0.2):
    """The plain-method workload: n^3 cube grid over [-0.5,0.5]^3 x 6 tets, interior vertices jittered by
    U(+-jitter*cell), blob colours with density in [0,20].  Returns dict(pos_gt, pos, cells, col)."""
    pos_gt, cells = hex_grid((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), n + 1, n + 1, n + 1)
    col = blob_colors(pos_gt, seed=seed + 1)
    pos = jitter_interior(pos_gt, cells, 1.0 / n, jitter, seed=seed + 2)
    return {"pos_gt": pos_gt, "pos": pos, "cells": cells, "col": col}


def unlink_tets(cells, positions, colors):
    """Every tet gets its own four vertices (TetMesh::unlinkTets, TetMesh.cpp:1950-1989): vertex 4t + i is the
    i-th corner of tet t, with that corner's position and colour.  Returns (cells [T,4], positions [4T,3],
    colors [4T,4], or None when colors is None: per-cell colours stay as they are)."""
    cells = np.asarray(cells, np.int64).reshape(-1, 4)
    flat = cells.reshape(-1)
    return (np.arange(flat.size, dtype=np.int32).reshape(-1, 4), np.asarray(positions)[flat],
            None if colors is None else np.asarray(colors)[flat])


def remove_transparent_tets(cells, positions, colors, alpha_threshold, cell_colors=None):
    """Keeps the tets with a vertex (or, with cell_colors, the cell) whose alpha is >= alpha_threshold and the
    vertices they use, in their original order (TetMesh::removeTransparentTets, TetMesh.cpp:1991-2079).  Returns
    (cells, positions, colors, cell_colors) of the kept part (colors / cell_colors None where not given)."""
    cells = np.asarray(cells, np.int64).reshape(-1, 4)
    positions = np.asarray(positions)
    if cell_colors is not None:
        keep = np.asarray(cell_colors)[:, 3] >= alpha_threshold
    else:
        keep = (np.asarray(colors)[cells, 3] >= alpha_threshold).any(axis=1)
    used = np.zeros(positions.shape[0], bool)
    used[cells[keep].reshape(-1)] = True
    new_index = np.cumsum(used) - 1
    return (new_index[cells[keep]].astype(np.int32), positions[used],
            None if colors is None else np.asarray(colors)[used],
            None if cell_colors is None else np.asarray(cell_colors)[keep])
