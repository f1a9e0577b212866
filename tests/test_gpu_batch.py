"""Batched rendering (dtv_render_batch / dtv_render_adjoint_batch): every item of a batch must give exactly what
it gives alone (images and fragment counts bit-identical: the item's records keep their order in the batch's
sort), and the batch adjoint must equal the sum of the single-item adjoints (FP32 atomics reorder the sums:
relative L2 <= 1e-5).  Items mix meshes, viewports and cameras; one item sees nothing."""
import numpy as np
import pytest
import torch

import oracle
from tests.helpers import FOVY, ZFAR, ZNEAR, adjoint_references, check_gradients, grid_case, rel_l2, view_from

pytestmark = pytest.mark.gpu

EYES = [(0.3, 0.4, 2.0), (-1.7, 0.6, 1.1), (0.9, -1.2, -1.4), (0.05, 2.1, 0.3)]


def _mesh(d, w):
    m = d.TetMesh()
    m.set_use_gradients(True)
    m.set_tet_mesh_data(w["cells"], w["pos"], w["col"])
    return m


def test_batch_images_equal_single_renders(gpu):
    import difftetvr_amd as d
    wa, wb = grid_case(n=12, seed=3), grid_case(n=7, seed=4, jitter=0.3)
    ma, mb = _mesh(d, wa), _mesh(d, wb)
    items = [(ma, view_from(EYES[0]), 64, 48), (mb, view_from(EYES[1]), 37, 53), (ma, view_from(EYES[2]), 80, 80),
             (mb, view_from((0.0, 0.0, 60.0)), 16, 16)]  # last: the mesh is a speck far away
    br = d.BatchRenderer()
    br.set_attenuation(1.0)
    br.set_camera_fovy(FOVY)
    br.set_use_early_ray_termination(True)
    br.set_items(items)
    imgs = br.render()
    cnts = br.get_fragment_counts()
    torch.cuda.synchronize()
    assert isinstance(imgs, list) and len(imgs) == 4
    for (m, vm, W, H), img, cnt, w in zip(items, imgs, cnts, [wa, wb, wa, wb]):
        r = d.Renderer(renderer_type=d.RendererType.INTERSECTION)
        r.set_tet_mesh(m)
        r.set_attenuation(1.0)
        r.set_camera_fovy(FOVY)
        r.set_viewport_size(W, H)
        r.set_view_matrix(vm)
        r.set_use_early_ray_termination(True)
        ref = r.render()
        refc = r.get_fragment_counts()
        torch.cuda.synchronize()
        assert torch.equal(img, ref)
        assert torch.equal(cnt, refc)
    st = br.get_stats()
    assert st["num_items"] == 4 and st["num_pixels"] == 64 * 48 + 37 * 53 + 80 * 80 + 16 * 16


def test_batch_matches_oracle(gpu):
    import difftetvr_amd as d
    w = grid_case(n=10, seed=5)
    m = _mesh(d, w)
    W, H = 56, 40
    br = d.BatchRenderer()
    br.set_attenuation(1.0)
    br.set_camera_fovy(FOVY)
    br.set_use_early_ray_termination(False)
    br.set_views(m, [view_from(e) for e in EYES], W, H)
    imgs = br.render()
    cnts = br.get_fragment_counts()
    torch.cuda.synchronize()
    assert imgs.shape == (4, H, W, 4)
    for i, e in enumerate(EYES):
        uf, _ = oracle.camera_uniforms(view_from(e), FOVY, ZNEAR, ZFAR, W, H)
        fr = oracle.Fragments(w["cells"], w["pos"], uf, ZNEAR, W, H)
        ref, _, cnt = oracle.forward(fr, w["pos"], w["col"], uf, 1.0)
        assert np.abs(imgs[i].cpu().numpy() - ref).max() <= 1e-5
        assert np.array_equal(cnts[i].cpu().numpy(), cnt)


def test_batch_adjoint_is_sum_of_single_adjoints(gpu):
    import difftetvr_amd as d
    w = grid_case(n=10, seed=6)
    W, H = 48, 44
    gen = torch.Generator(device="cuda").manual_seed(0)
    dimgs = torch.randn((len(EYES), H, W, 4), device="cuda", generator=gen)
    # single renders, gradients accumulated over the views
    mesh1 = _mesh(d, w)
    for i, e in enumerate(EYES):
        r = d.Renderer(renderer_type=d.RendererType.INTERSECTION)
        r.set_tet_mesh(mesh1)
        r.set_attenuation(1.0)
        r.set_camera_fovy(FOVY)
        r.set_viewport_size(W, H)
        r.set_view_matrix(view_from(e))
        r.render()
        r.render_adjoint(dimgs[i].contiguous())
    # one batch
    mesh2 = _mesh(d, w)
    br = d.BatchRenderer()
    br.set_attenuation(1.0)
    br.set_camera_fovy(FOVY)
    br.set_views(mesh2, [view_from(e) for e in EYES], W, H)
    br.render()
    br.render_adjoint(dimgs)
    torch.cuda.synchronize()
    g1p, g1c = mesh1.get_vertex_positions().grad.cpu().numpy(), mesh1.get_vertex_colors().grad.cpu().numpy()
    g2p, g2c = mesh2.get_vertex_positions().grad.cpu().numpy(), mesh2.get_vertex_colors().grad.cpu().numpy()
    print(f"batch adjoint rel_l2 pos={rel_l2(g2p, g1p):.2e} col={rel_l2(g2c, g1c):.2e}", flush=True)
    assert rel_l2(g2p, g1p) <= 1e-5
    assert rel_l2(g2c, g1c) <= 1e-5


def test_batch_two_meshes_adjoint_targets(gpu):
    import difftetvr_amd as d
    wa, wb = grid_case(n=8, seed=7), grid_case(n=9, seed=8)
    W, H = 40, 40
    ma, mb = _mesh(d, wa), _mesh(d, wb)
    br = d.BatchRenderer()
    br.set_attenuation(1.0)
    br.set_camera_fovy(FOVY)
    br.set_items([(ma, view_from(EYES[0]), W, H), (mb, view_from(EYES[1]), W, H)])
    imgs = br.render()
    dimgs = torch.ones_like(imgs)
    br.render_adjoint(dimgs)
    torch.cuda.synchronize()
    ones = np.ones((H, W, 4), np.float32)
    for m, w, e in ((ma, wa, EYES[0]), (mb, wb, EYES[1])):
        uf, ud = oracle.camera_uniforms(view_from(e), FOVY, ZNEAR, ZFAR, W, H)
        fr = oracle.Fragments(w["cells"], w["pos"], uf, ZNEAR, W, H)
        # meshes with gradients render with early ray termination (the renderer's default)
        check_gradients("batch item", m.get_vertex_positions().grad.cpu().numpy(),
                        m.get_vertex_colors().grad.cpu().numpy(),
                        adjoint_references(oracle, fr, w["pos"], w["col"], uf, ud, 1.0, ones, use_ert=True))


def test_batch_validation(gpu):
    import difftetvr_amd as d
    w = grid_case(n=4, seed=1)
    m = _mesh(d, w)
    br = d.BatchRenderer()
    with pytest.raises(RuntimeError):
        br.render()
    with pytest.raises(RuntimeError):
        br.set_items([])
    br.set_views(m, [view_from(EYES[0]), view_from(EYES[1])], 16, 16)
    with pytest.raises(RuntimeError):
        br.render_adjoint(torch.zeros((2, 16, 16, 4), device="cuda"))
    br.render()
    with pytest.raises(RuntimeError):
        br.render_adjoint(torch.zeros((3, 16, 16, 4), device="cuda"))
    with pytest.raises(RuntimeError):
        br.render_adjoint(torch.zeros((2, 16, 15, 4), device="cuda"))


def test_batch_adjoint_rejects_changed_mesh(gpu):
    """The batched adjoint reads the meshes the render recorded: an item whose mesh was replaced since (new
    buffers or sizes) is refused before any work, through the C ABI (DTV_ERR_INVALID_ARGUMENT)."""
    import ctypes

    import difftetvr_amd as d
    from difftetvr_amd import _lib
    wa, wb = grid_case(n=5, seed=1), grid_case(n=6, seed=2)
    m = _mesh(d, wa)
    br = d.BatchRenderer()
    br.set_views(m, [view_from(EYES[0]), view_from(EYES[1])], 16, 16)
    br.render()
    keep = torch.zeros((2, 16, 16, 4), device="cuda")
    lib = _lib.load()

    def call(descs):
        arr = (_lib.DtvAdjointItem * 2)()
        for i in range(2):
            arr[i].mesh = descs[i]
            arr[i].dimage = keep[i].data_ptr()
            arr[i].grad_positions = m.get_vertex_positions().grad.data_ptr()
            arr[i].grad_colors = m.get_vertex_colors().grad.data_ptr()
        return lib.dtv_render_adjoint_batch(br._view, arr, 2, 0, 0, ctypes.c_void_p(0))

    assert call([m._desc(), m._desc()]) == 0
    torch.cuda.synchronize()
    m.set_tet_mesh_data(wb["cells"], wb["pos"], wb["col"])  # new buffers and sizes
    assert call([m._desc(), m._desc()]) == -1
    assert b"mesh differs" in lib.dtv_last_error()
    with pytest.raises(RuntimeError, match="mesh differs"):
        br.render_adjoint(keep)


def test_mse_loss_grad_views_matches_torch(gpu):
    from difftetvr_amd.optim import mse_loss_grad_views
    g = torch.Generator(device="cuda").manual_seed(3)
    imgs = torch.rand((5, 17, 23, 4), device="cuda", generator=g).requires_grad_(True)
    tgts = torch.rand((5, 17, 23, 4), device="cuda", generator=g)
    loss, dimg = mse_loss_grad_views(imgs.detach(), tgts)
    ref = sum(torch.nn.functional.mse_loss(imgs[i], tgts[i]) for i in range(5))
    ref.backward()
    torch.cuda.synchronize()
    r = float(ref.detach())
    assert abs(float(loss) - r) <= 1e-5 * r
    assert torch.allclose(dimg, imgs.grad, rtol=1e-6, atol=1e-9)
