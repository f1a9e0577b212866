"""Tile launch order (dtv_render_params.tile_order): the heaviest-first order must render every tile exactly once.
Images and fragment counts equal the row-major launch's bit for bit (a tile left out would keep the NaN
written into the image storage beforehand); the adjoint equals the row-major one within FP32 atomic reordering
(relative L2 <= 1e-5).  Viewports span 1 tile, a partial last row/column, more than 1024 tiles, and more than
16384 tiles (more than one load step per pass of the ordering workgroup)."""
import numpy as np
import pytest
import torch

from tests.helpers import FOVY, grid_case, make_renderer, rel_l2, view_from

pytestmark = pytest.mark.gpu


def _render_into(d, r, img):
    """dtv_render into a caller-owned image (here pre-filled with NaN)."""
    import ctypes

    from difftetvr_amd import _lib
    desc, cam, prm = r._mesh._desc(), r._camera(), r._params()
    _lib.check(r._lib.dtv_render(r._view, ctypes.byref(desc), ctypes.byref(cam), ctypes.byref(prm),
                                 ctypes.c_void_p(img.data_ptr()), d._stream_ptr()), "dtv_render")
    r._rendered_mesh = r._mesh
    return img


@pytest.mark.parametrize("W,H,use_ert", [(400, 300, True), (61, 45, False), (8, 8, False), (1100, 1100, True)])
def test_cost_order_equals_row_major(gpu, W, H, use_ert):
    import difftetvr_amd as d
    w = grid_case(n=14, seed=7)
    gen = torch.Generator(device="cuda").manual_seed(1)
    dimg = torch.randn((H, W, 4), device="cuda", generator=gen)
    out = {}
    for order in ("row-major", "cost"):
        print(f"tile order {order}, {W}x{H}, ert={use_ert}", flush=True)
        mesh, r = make_renderer(d, w["cells"], w["pos"], w["col"], W, H, eye=(0.4, 0.5, 1.8), use_ert=use_ert)
        r.set_tile_order(order)
        img = _render_into(d, r, torch.full((H, W, 4), float("nan"), device="cuda"))
        cnt = r.get_fragment_counts()
        r.render_adjoint(dimg)
        torch.cuda.synchronize()
        out[order] = (img.clone(), cnt.clone(), mesh.get_vertex_positions().grad.cpu().numpy(),
                      mesh.get_vertex_colors().grad.cpu().numpy())
    a, b = out["row-major"], out["cost"]
    assert not torch.isnan(b[0]).any()
    assert torch.equal(a[0], b[0])
    assert torch.equal(a[1], b[1])
    print(f"adjoint rel_l2 pos={rel_l2(b[2], a[2]):.2e} col={rel_l2(b[3], a[3]):.2e}", flush=True)
    assert rel_l2(b[2], a[2]) <= 1e-5
    assert rel_l2(b[3], a[3]) <= 1e-5


def test_cost_order_batch(gpu):
    import difftetvr_amd as d
    wa, wb = grid_case(n=12, seed=3), grid_case(n=7, seed=4, jitter=0.3)
    meshes = []
    for w in (wa, wb):
        m = d.TetMesh()
        m.set_use_gradients(True)
        m.set_tet_mesh_data(w["cells"], w["pos"], w["col"])
        meshes.append(m)
    items = [(meshes[0], view_from((0.3, 0.4, 2.0)), 64, 48), (meshes[1], view_from((-1.7, 0.6, 1.1)), 37, 53),
             (meshes[0], view_from((0.05, 2.1, 0.3)), 200, 120)]
    res = {}
    for order in ("row-major", "cost"):
        print(f"batch tile order {order}", flush=True)
        br = d.BatchRenderer()
        br.set_attenuation(1.0)
        br.set_camera_fovy(FOVY)
        br.set_tile_order(order)
        br.set_items(items)
        imgs = br.render()
        torch.cuda.synchronize()
        res[order] = [i.clone() for i in imgs]
    for a, b in zip(res["row-major"], res["cost"]):
        assert torch.equal(a, b)


def test_tile_order_rejects_unknown():
    import difftetvr_amd as d
    r = d.Renderer(renderer_type=d.RendererType.INTERSECTION)
    with pytest.raises(RuntimeError):
        r.set_tile_order("diagonal")
    np.testing.assert_equal(r._params().tile_order, 0)
