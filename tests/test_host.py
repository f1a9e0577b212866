"""Host-side checks that need no GPU: mesh generators, .bintet I/O, the C ABI's exported symbols, and the
module's fail-loudly behaviour."""
import ctypes
import os
import re

import numpy as np
import pytest

from difftetvr_amd import bintet, meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("n", [1, 3, 16])
def test_hex_grid_counts_orientation_boundary(n):
    pos, cells = meshgen.hex_grid((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), n + 1, n + 1, n + 1)
    assert pos.shape == ((n + 1) ** 3, 3) and cells.shape == (6 * n ** 3, 4)
    assert (meshgen.tet_orientation(pos, cells) < 0).all()  # TetMesh.cpp:532-541 convention
    # the 6 tets of a cell fill it exactly
    vol = -meshgen.tet_orientation(pos, cells) / 6.0
    assert abs(vol.sum() - 1.0) < 1e-6
    faces = np.sort(cells[:, [[0, 1, 2], [1, 0, 3], [0, 2, 3], [2, 1, 3]]].reshape(-1, 3), axis=1)
    uniq, cnt = np.unique(faces, axis=0, return_counts=True)
    assert uniq.shape[0] == 2 * cells.shape[0] + 6 * n * n    # F = 2T + 6n^2
    assert (cnt == 1).sum() == 12 * n * n                      # boundary faces
    mask = meshgen.boundary_vertex_mask(cells, pos.shape[0])
    on_hull = (np.abs(np.abs(pos) - 0.5) < 1e-7).any(axis=1)
    assert np.array_equal(mask, on_hull)


def test_jitter_keeps_boundary_and_orientation():
    w = meshgen.grid_workload(8, seed=0)
    mask = meshgen.boundary_vertex_mask(w["cells"], w["pos"].shape[0])
    assert np.array_equal(w["pos"][mask], w["pos_gt"][mask])
    d = np.abs(w["pos"] - w["pos_gt"])[~mask]
    assert d.max() <= 0.2 / 8 + 1e-7 and d.max() > 0.1 / 8
    assert (meshgen.tet_orientation(w["pos"], w["cells"]) < 0).all()
    assert w["col"][:, 3].min() >= 0 and w["col"][:, 3].max() <= 20.0 + 1e-5
    assert (w["col"][:, :3] >= 0).all() and (w["col"][:, :3] <= 1).all()


def test_builtin_test_meshes():
    pos, cells, col = meshgen.single_tetrahedron()
    assert cells.shape == (1, 4) and pos.shape == (4, 3) and col.shape == (4, 4)
    pos, cells, col = meshgen.cube_central_gradient()
    assert cells.shape == (24, 4) and pos.shape == (20, 3)
    assert np.allclose(np.abs(pos).max(), 0.25)


def test_delaunay_ball():
    pos, cells = meshgen.delaunay_ball(500, seed=1)
    assert (meshgen.tet_orientation(pos, cells) < 0).all()
    assert (np.linalg.norm(pos, axis=1) <= 1.0 + 1e-6).all()
    assert cells.shape[0] > 5 * 500 * 0.8


def test_look_at_matches_sample_view_construction():
    eye = np.array([0.4, -0.3, 1.7])
    v = meshgen.look_at_view_matrix(eye).reshape(4, 4).T.astype(np.float64)
    assert np.allclose(v @ np.append(eye, 1.0), [0, 0, 0, 1], atol=1e-6)
    fwd = v @ np.array([0.0, 0.0, 0.0, 1.0])   # the origin lies on the camera's -z axis
    assert abs(fwd[0]) < 1e-6 and abs(fwd[1]) < 1e-6 and fwd[2] < 0
    assert np.allclose(v[:3, :3] @ v[:3, :3].T, np.eye(3), atol=1e-6)


def test_fibonacci_sphere():
    p = meshgen.fibonacci_sphere(16, 2.0)
    assert np.allclose(np.linalg.norm(p, axis=1), 2.0)
    assert np.abs(p[:, 1]).max() < 2.0 * 0.97


def test_value_noise_range():
    pts = np.random.default_rng(0).uniform(-1, 1, (1000, 3))
    n = meshgen.value_noise(pts, seed=3)
    assert n.min() >= 0 and n.max() <= 40.0 and n.std() > 1.0


def test_bintet_roundtrip(tmp_path):
    pos, cells = meshgen.hex_grid((0, 0, 0), (1, 2, 3), 3, 4, 5)
    col = np.random.default_rng(0).uniform(size=(pos.shape[0], 4)).astype(np.float32)
    p = str(tmp_path / "m.bintet")
    bintet.write(p, cells, pos, col)
    c2, p2, col2, cc2 = bintet.read(p)
    assert np.array_equal(c2, cells) and np.array_equal(p2, pos) and np.array_equal(col2, col) and cc2 is None
    raw = open(p, "rb").read()
    assert raw[:4] == (1).to_bytes(4, "little")
    assert raw[4:8] == (cells.size).to_bytes(4, "little")
    # version 2: cell colours after an empty vertex colour array
    ccol = np.random.default_rng(1).uniform(size=(cells.shape[0], 4)).astype(np.float32)
    p_2 = str(tmp_path / "m2.bintet")
    bintet.write(p_2, cells, pos, None, ccol)
    raw2 = open(p_2, "rb").read()
    assert raw2[:4] == (2).to_bytes(4, "little")
    tail = 4 + 4 + cells.size * 4 + 4 + pos.size * 4
    assert raw2[tail:tail + 4] == (0).to_bytes(4, "little")
    assert raw2[tail + 4:tail + 8] == (cells.shape[0]).to_bytes(4, "little")
    assert len(raw2) == tail + 8 + ccol.size * 4
    c3, p3, col3, cc3 = bintet.read(p_2)
    assert np.array_equal(c3, cells) and np.array_equal(p3, pos) and col3 is None and np.array_equal(cc3, ccol)
    bad = str(tmp_path / "bad.bintet")
    open(bad, "wb").write((7).to_bytes(4, "little") + raw[4:])
    with pytest.raises(RuntimeError, match="Invalid version"):
        bintet.read(bad)
    open(bad, "wb").write(raw[:40])
    with pytest.raises(RuntimeError, match="truncated"):
        bintet.read(bad)


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "difftetvr_amd.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(dtv_\w+)\s*\(", txt, flags=re.M)))


def test_c_abi_exports_every_declared_symbol():
    from difftetvr_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    declared = _header_functions()
    assert len(declared) >= 14
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(n for n, _, _ in _lib.SIGNATURES) == declared
    out = os.popen(f"nm -D --defined-only {_lib.LIB_PATH}").read()
    exported = sorted(set(re.findall(r"\bT (dtv_\w+)", out)))
    assert exported == declared


def test_library_loads_and_reports_version():
    from difftetvr_amd import _lib
    lib = _lib.load()
    assert lib.dtv_get_version() == 1
    cam = _lib.DtvCamera()
    for i in range(16):
        cam.view[i] = 1.0 if i % 5 == 0 else 0.0
    cam.fovy, cam.znear, cam.zfar, cam.width, cam.height = 0.8, 0.01, 100.0, 64, 32
    out = (ctypes.c_float * 68)()
    assert lib.dtv_camera_uniforms(ctypes.byref(cam), out) == 0   # host-only entry point
    assert lib.dtv_render(None, None, None, None, None, None) == -1
    assert b"NULL" in lib.dtv_last_error()


def test_module_fails_loudly_without_gpu():
    import torch

    import difftetvr_amd as d
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError):
        d.set_device_type("cpu")
    with pytest.raises(d.DtvError, match="GPU"):
        d.TetMesh()


def test_render_params_layout_matches_header():
    """dtv_render_params in include/difftetvr_amd.h and its ctypes mirror list the same fields in order (the
    renderer selector last), and the selector values agree."""
    from difftetvr_amd import _lib
    txt = open(os.path.join(ROOT, "include", "difftetvr_amd.h")).read()
    body = txt[txt.index("typedef struct dtv_render_params {"):txt.index("} dtv_render_params;")]
    names = re.findall(r"^\s+(?:float|int32_t)\s+(\w+)(?:\[\d+\])?;", body, flags=re.M)
    assert names == [f[0] for f in _lib.DtvRenderParams._fields_]
    assert re.search(r"#define DTV_RENDERER_PROJECTION 1\b", txt) and _lib.DTV_RENDERER_PROJECTION == 1
    assert re.search(r"#define DTV_TILE_LAUNCH_QUEUE 1\b", txt) and _lib.DTV_TILE_LAUNCH_QUEUE == 1


def test_unlink_tets():
    from difftetvr_amd import meshgen
    w = meshgen.grid_workload(3, seed=4)
    cells, pos, col = meshgen.unlink_tets(w["cells"], w["pos"], w["col"])
    assert cells.shape == w["cells"].shape and pos.shape == (cells.size, 3) and col.shape == (cells.size, 4)
    assert np.array_equal(cells.ravel(), np.arange(cells.size))
    # every tet keeps its corner positions and colours, in order
    assert np.array_equal(pos[cells], w["pos"][w["cells"]])
    assert np.array_equal(col[cells], w["col"][w["cells"]])


def test_remove_transparent_tets():
    from difftetvr_amd import meshgen
    w = meshgen.grid_workload(4, seed=6)
    col = w["col"].copy()
    thr = float(np.quantile(col[:, 3], 0.7))
    cells, pos, col2, _ = meshgen.remove_transparent_tets(w["cells"], w["pos"], col, thr)
    keep = [t for t in range(w["cells"].shape[0]) if (col[w["cells"][t], 3] >= thr).any()]
    assert 0 < len(keep) < w["cells"].shape[0]
    assert cells.shape == (len(keep), 4)
    # the kept tets in order, with their corners' positions and colours
    assert np.array_equal(pos[cells], w["pos"][w["cells"][keep]])
    assert np.array_equal(col2[cells], col[w["cells"][keep]])
    # every kept vertex is used, in the original vertex order
    used = np.unique(w["cells"][keep])
    assert np.array_equal(pos, w["pos"][used])
    # cell colours decide by the cell's own alpha
    cc = np.random.default_rng(1).uniform(size=(w["cells"].shape[0], 4)).astype(np.float32)
    c3, _, _, cc2 = meshgen.remove_transparent_tets(w["cells"], w["pos"], None, 0.5, cell_colors=cc)
    assert np.array_equal(cc2, cc[cc[:, 3] >= 0.5]) and c3.shape[0] == cc2.shape[0]
