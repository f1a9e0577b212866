"""Reference API calls outside the render/adjoint core: the PPLL linked-list export (Module.cpp:960-981) and
TetMesh.set_triangle_mesh_data (TetMesh.cpp:369-478)."""
import numpy as np
import pytest
import torch

import oracle
from difftetvr_amd import meshgen
from tests.helpers import FOVY, ZFAR, ZNEAR, view_from

pytestmark = pytest.mark.gpu


def _ppll(d, w, W, H, eye, renderer_type=None):
    mesh = d.TetMesh()
    mesh.set_tet_mesh_data(w["cells"], w["pos"], w["col"])
    r = d.Renderer(renderer_type=d.RendererType.PPLL if renderer_type is None else renderer_type)
    r.set_tet_mesh(mesh)
    r.set_attenuation(1.0)
    r.set_viewport_size(W, H)
    r.set_camera_fovy(FOVY)
    r.set_view_matrix(view_from(eye))
    return mesh, r


def test_linked_list_export_matches_fragment_lists(gpu):
    """Each pixel's list, walked from its start offset, holds the renderer's fragments front to back: face bits
    (fragment id << 2 | boundary << 1 | 1), the face's window depth at the pixel centre (float64 restatement),
    and the boundary bit of that face."""
    import difftetvr_amd as d
    W, H, eye = 48, 40, (0.7, 0.5, 1.6)
    w = meshgen.grid_workload(6, seed=1)
    mesh, r = _ppll(d, w, W, H, eye)
    with pytest.raises(RuntimeError, match="exportLinkedListData is set to false"):
        r.render()
        r.get_fragment_buffer()
    r.set_export_linked_list_data()
    r.render()
    buf = r.get_fragment_buffer().cpu()
    start = r.get_start_offset_buffer().cpu().numpy()
    ids = r.get_fragment_ids(64).cpu().numpy()
    bits = buf[:, 0].contiguous().view(torch.int32).numpy().astype(np.int64) & 0xFFFFFFFF
    depth = buf[:, 1].numpy().astype(np.float64)
    nxt = buf[:, 2].contiguous().view(torch.int32).numpy().astype(np.int64) & 0xFFFFFFFF
    assert start.shape == (H, W) and buf.shape[1] == 3
    _, ud = oracle.camera_uniforms(view_from(eye), FOVY, ZNEAR, ZFAR, W, H)
    vp = ud[:16].reshape(4, 4).T
    cells, pos = w["cells"].astype(np.int64), w["pos"].astype(np.float64)
    faces = np.sort(cells[:, meshgen.TET_FACE_TABLE].reshape(-1, 3), axis=1)
    _, inv, cnt = np.unique(faces, axis=0, return_inverse=True, return_counts=True)
    bnd = (cnt[inv.reshape(-1)] == 1)
    nodes = 0
    for y in range(H):
        for x in range(W):
            lst = []
            k = int(start[y, x])
            while k != -1 and k != 0xFFFFFFFF:
                lst.append(k)
                k = int(nxt[k])
                assert len(lst) <= 64
            want = [int(i) for i in ids[y, x] if i >= 0]
            assert [int(bits[k] >> 2) for k in lst] == want
            for k, fid in zip(lst, want):
                assert bits[k] & 1 == 1 and ((bits[k] >> 1) & 1) == int(bnd[fid])
                v = cells[fid >> 2, meshgen.TET_FACE_TABLE[fid & 3]]
                c = (vp @ np.c_[pos[v], np.ones(3)].T).T
                q = c[:, :3] / c[:, 3:]
                px, py = 2 * (x + 0.5) / W - 1, 2 * (y + 0.5) / H - 1
                A = np.array([[q[0, 0], q[1, 0], q[2, 0]], [q[0, 1], q[1, 1], q[2, 1]], [1, 1, 1]])
                lam = np.linalg.solve(A, np.array([px, py, 1.0]))
                assert abs(depth[k] - lam @ q[:, 2]) < 1e-5
            nodes += len(lst)
    assert nodes == buf.shape[0] > W * H
    _, ri = _ppll(d, w, W, H, eye, renderer_type=d.RendererType.INTERSECTION)
    with pytest.raises(RuntimeError, match="must be of type RendererType.PPLL"):
        ri.set_export_linked_list_data()


def test_set_triangle_mesh_data(gpu):
    import difftetvr_amd as d
    m = d.TetMesh()
    tri = torch.tensor([[0, 1, 2], [0, 2, 3]], dtype=torch.int32)
    pos = torch.rand(4, 3)
    col = torch.rand(4, 4)
    with pytest.raises(RuntimeError, match="triangleIndicesTensor.dtype"):
        m.set_triangle_mesh_data(tri.long(), pos, col)
    with pytest.raises(RuntimeError, match="vertexPositionsTensor.size\\(0\\) != vertexColorsTensor.size\\(0\\)"):
        m.set_triangle_mesh_data(tri, pos, col[:3])
    with pytest.raises(RuntimeError, match="triangleIndicesTensor.size\\(1\\) != 3"):
        m.set_triangle_mesh_data(torch.zeros(2, 4, dtype=torch.int32), pos, col)
    m.set_triangle_mesh_data(tri, pos, col)
    assert m.get_has_triangle_mesh_data() and m.get_num_cells() == 2 and m.get_num_vertices() == 4
    assert torch.equal(m.get_vertex_positions().cpu(), pos)
    r = d.Renderer(renderer_type=d.RendererType.INTERSECTION)
    r.set_tet_mesh(m)
    r.set_viewport_size(16, 16)
    with pytest.raises(RuntimeError, match="no tets"):
        r.render()
