"""Parity of the HIP path against the CPU oracle on the same seeded inputs (through the C ABI via the module).

Tolerances (written here, from the target spec): per-pixel fragment counts and per-pixel fragment order
bit-exact; images <= 1e-5 max-abs in FP32; gradients <= 1e-4 relative L2 against the oracle's exact adjoint.
"""
import math

import numpy as np
import pytest
import torch

import oracle
from tests.helpers import (FOVY, ZFAR, ZNEAR, adjoint_references, check_gradients, grid_case, make_renderer,
                           rel_l2, view_from)

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-5
GRAD_TOL = 1e-4


def _oracle_frags(cells, pos, eye, W, H, fovy=FOVY):
    uf, ud = oracle.camera_uniforms(view_from(eye), fovy, ZNEAR, ZFAR, W, H)
    return oracle.Fragments(cells, pos, uf, ZNEAR, W, H), uf, ud


def _check_forward(d, cells, pos, col, W, H, eye, att, clear=(0, 0, 0, 0), use_ert=False):
    mesh, r = make_renderer(d, cells, pos, col, W, H, eye=eye, attenuation=att, clear=clear, use_ert=use_ert)
    img = r.render()
    torch.cuda.synchronize()
    fr, uf, _ = _oracle_frags(cells, pos, eye, W, H)
    oimg, oC, ocnt = oracle.forward(fr, pos, col, uf, att, clear=clear, use_ert=use_ert)
    gimg = img.cpu().numpy()
    err = float(np.abs(gimg - oimg).max())
    counts = r.get_fragment_counts().cpu().numpy()
    return mesh, r, gimg, oimg, err, counts, ocnt, fr, uf


def test_camera_uniforms_bit_identical(gpu):
    import difftetvr_amd as d
    w = grid_case(4)
    _, r = make_renderer(d, w["cells"], w["pos"], w["col"], 200, 120, eye=(0.3, 0.5, 2.0))
    u = r.camera_uniforms()
    uf, _ = oracle.camera_uniforms(view_from((0.3, 0.5, 2.0)), FOVY, ZNEAR, ZFAR, 200, 120)
    assert np.array_equal(u, uf)


@pytest.mark.parametrize("eye", [(0.0, 0.0, 2.0), (1.2, 0.7, 1.4)])
def test_grid16_forward_parity(gpu, eye):
    import difftetvr_amd as d
    w = grid_case(16, seed=0)
    _, r, gimg, oimg, err, counts, ocnt, fr, _ = _check_forward(d, w["cells"], w["pos"], w["col"], 256, 256, eye, 1.0)
    print(f"grid16 eye={eye}: max|img diff|={err:.3e}, frags={fr.total}")
    assert np.array_equal(counts, ocnt), "per-pixel fragment counts differ"
    assert err <= IMG_TOL


def test_grid16_fragment_order_bit_exact(gpu):
    import difftetvr_amd as d
    w = grid_case(16, seed=3)
    eye = (0.9, -0.6, 1.7)
    mesh, r = make_renderer(d, w["cells"], w["pos"], w["col"], 128, 96, eye=eye)
    r.render()
    fr, _, _ = _oracle_frags(w["cells"], w["pos"], eye, 128, 96)
    off, ids = fr.lists()
    maxn = int(np.diff(off).max())
    g = r.get_fragment_ids(maxn).cpu().numpy().reshape(-1, maxn)
    o = np.full_like(g, -1)
    for p in range(128 * 96):
        n = off[p + 1] - off[p]
        o[p, :n] = ids[off[p]:off[p + 1]]
    assert np.array_equal(g, o)


@pytest.mark.parametrize("W,H", [(61, 37), (8, 8), (1, 1), (130, 9)])
def test_ragged_viewports(gpu, W, H):
    import difftetvr_amd as d
    w = grid_case(6, seed=1)
    _, _, _, _, err, counts, ocnt, _, _ = _check_forward(d, w["cells"], w["pos"], w["col"], W, H, (0.4, 0.3, 1.8), 3.0)
    assert np.array_equal(counts, ocnt)
    assert err <= IMG_TOL


def test_clear_colour_and_attenuation(gpu):
    import difftetvr_amd as d
    w = grid_case(8, seed=2)
    _, _, gimg, _, err, counts, ocnt, _, _ = _check_forward(d, w["cells"], w["pos"], w["col"], 96, 80,
                                                            (0.0, 0.5, 2.2), 7.5, clear=(0.2, 0.4, 0.6, 0.8))
    assert np.array_equal(counts, ocnt)
    assert err <= IMG_TOL


def test_builtin_test_meshes(gpu):
    import difftetvr_amd as d
    from difftetvr_amd import meshgen
    for pos, cells, col in (meshgen.single_tetrahedron(), meshgen.cube_central_gradient()):
        _, _, gimg, _, err, counts, ocnt, _, _ = _check_forward(d, cells, pos, col, 128, 128, (0.3, 0.2, 1.0), 100.0)
        assert np.array_equal(counts, ocnt)
        assert err <= IMG_TOL
        assert gimg[..., 3].max() > 0.0


def test_empty_view_and_culled_tets(gpu):
    import difftetvr_amd as d
    w = grid_case(4)
    # camera looking away from the mesh: nothing visible
    mesh, r = make_renderer(d, w["cells"], w["pos"] + np.float32([0, 0, 5.0]), w["col"], 64, 64,
                            clear=(0.1, 0.2, 0.3, 0.4))
    img = r.render().cpu().numpy()
    assert np.allclose(img, np.float32([0.1, 0.2, 0.3, 0.4]))
    assert int(r.get_fragment_counts().sum()) == 0
    # mesh straddling the camera: tets behind the near plane are culled, the rest renders like the oracle
    pos = w["pos"] * np.float32(4.0) + np.float32([0, 0, 1.5])
    _, _, _, _, err, counts, ocnt, _, _ = _check_forward(d, w["cells"], pos, w["col"], 64, 64, (0.0, 0.0, 2.0), 1.0)
    assert np.array_equal(counts, ocnt)
    assert err <= IMG_TOL


def test_early_ray_termination(gpu):
    import difftetvr_amd as d
    w = grid_case(12, seed=4)
    col = w["col"].copy()
    col[:, 3] *= 5.0
    _, r, gimg, oimg, err, counts, ocnt, _, _ = _check_forward(d, w["cells"], w["pos"], col, 128, 128,
                                                               (0.0, 0.0, 2.0), 1.0, use_ert=True)
    mismatch = int((counts != ocnt).sum())
    print(f"ERT: max|diff|={err:.3e}, count mismatches={mismatch}")
    assert mismatch <= 4  # alpha crossing the threshold within an exp() ulp of the CPU
    assert err <= 2e-4    # the skipped remainder is bounded by the 1e-4 transmittance threshold


def test_grid16_adjoint_parity(gpu):
    import difftetvr_amd as d
    w = grid_case(16, seed=0)
    eye = (0.5, 0.4, 1.9)
    W = H = 192
    mesh, r = make_renderer(d, w["cells"], w["pos"], w["col"], W, H, eye=eye)
    img = r.render()
    dimg = np.random.default_rng(7).normal(size=(H, W, 4)).astype(np.float32)
    r.render_adjoint(torch.from_numpy(dimg).cuda(), False)
    torch.cuda.synchronize()
    gp = mesh.get_vertex_positions().grad.cpu().numpy()
    gc = mesh.get_vertex_colors().grad.cpu().numpy()
    fr, uf, ud = _oracle_frags(w["cells"], w["pos"], eye, W, H)
    check_gradients("grid16 192^2 attenuation 1", gp, gc,
                    adjoint_references(oracle, fr, w["pos"], w["col"], uf, ud, 1.0, dimg))


def test_adjoint_accumulates_and_abs(gpu):
    import difftetvr_amd as d
    w = grid_case(8, seed=5)
    eye = (0.2, 0.9, 1.8)
    mesh, r = make_renderer(d, w["cells"], w["pos"], w["col"], 64, 64, eye=eye, attenuation=4.0)
    r.render()
    dimg = torch.from_numpy(np.random.default_rng(1).normal(size=(64, 64, 4)).astype(np.float32)).cuda()
    r.render_adjoint(dimg, False)
    r.render_adjoint(dimg, False)
    torch.cuda.synchronize()
    g2 = mesh.get_vertex_colors().grad.cpu().numpy().copy()
    fr, uf, _ = _oracle_frags(w["cells"], w["pos"], eye, 64, 64)
    _, oc = oracle.adjoint_exact(fr, w["pos"], w["col"], uf, 4.0, dimg.cpu().numpy())
    assert rel_l2(g2, 2 * oc) <= GRAD_TOL
    mesh.get_vertex_colors().grad.zero_()
    mesh.get_vertex_positions().grad.zero_()
    r.render_adjoint(dimg, True)
    torch.cuda.synchronize()
    op, oc = oracle.adjoint_exact(fr, w["pos"], w["col"], uf, 4.0, dimg.cpu().numpy(), use_abs_grad=True)
    assert rel_l2(mesh.get_vertex_colors().grad.cpu().numpy(), oc) <= GRAD_TOL
    assert rel_l2(mesh.get_vertex_positions().grad.cpu().numpy(), op) <= GRAD_TOL


def test_render_adjoint_input_validation(gpu):
    import difftetvr_amd as d
    w = grid_case(4)
    mesh, r = make_renderer(d, w["cells"], w["pos"], w["col"], 32, 32)
    r.render()
    with pytest.raises(RuntimeError, match="sizes"):
        r.render_adjoint(torch.zeros(32, 32 * 4, device="cuda"), False)
    with pytest.raises(RuntimeError, match="channels"):
        r.render_adjoint(torch.zeros(32, 32, 3, device="cuda"), False)
    with pytest.raises(RuntimeError, match="32-bit float"):
        r.render_adjoint(torch.zeros(32, 32, 4, device="cuda", dtype=torch.float64), False)
    with pytest.raises(RuntimeError, match="device"):
        r.render_adjoint(torch.zeros(32, 32, 4), False)


@pytest.mark.parametrize("use_abs", [False, True])
def test_reference_adjoint_mode(gpu, use_abs):
    """set_adjoint_mode('reference'): the reference's own adjoint (sub-segment 0 skipped, fixed-barycentric
    hit points), against the oracle's restatement of it (FP32; the inversion per sub-segment carries both
    implementations' rounding along each ray, hence 1e-4 relative L2 for colours and positions alike)."""
    import difftetvr_amd as d
    w = grid_case(12, seed=3)
    eye = (0.5, 0.4, 1.9)
    W, H = 96, 80
    mesh, r = make_renderer(d, w["cells"], w["pos"], w["col"], W, H, eye=eye, attenuation=1.0)
    r.set_adjoint_mode("reference")
    img = r.render()
    dimg = np.random.default_rng(9).normal(size=(H, W, 4)).astype(np.float32)
    r.render_adjoint(torch.from_numpy(dimg).cuda(), use_abs)
    torch.cuda.synchronize()
    gp = mesh.get_vertex_positions().grad.cpu().numpy()
    gc = mesh.get_vertex_colors().grad.cpu().numpy()
    fr, uf, _ = _oracle_frags(w["cells"], w["pos"], eye, W, H)
    # the reference adjoint starts from the rendered image (it inverts the blended output): the GPU's own
    op, oc = oracle.adjoint_reference(fr, w["pos"], w["col"], uf, 1.0, img.cpu().numpy(), dimg, use_abs_grad=use_abs)
    _, ec = oracle.adjoint_exact(fr, w["pos"], w["col"], uf, 1.0, dimg, use_abs_grad=use_abs)
    print(f"reference adjoint rel L2: pos {rel_l2(gp, op):.3e} col {rel_l2(gc, oc):.3e}; "
          f"reference vs exact colour {rel_l2(oc, ec):.3e}", flush=True)
    assert rel_l2(gc, oc) <= GRAD_TOL
    assert rel_l2(gp, op) <= GRAD_TOL
    assert rel_l2(oc, ec) > 0.05          # the quirks are really there


@pytest.mark.parametrize("att,interval", [(100.0, 64), (12.0, 3)])
def test_adjoint_dense_field_no_early_termination(gpu, att, interval):
    """Attenuation 100 (the reference renderer's default) without early ray termination: rays saturate (alpha
    exactly 1 in float) within a few tets and many fragments are opaque, so the adjoint cannot recover those
    fragments' input states by inverting 'over'; and a checkpoint interval of 3 asks for more periodic
    checkpoints than there are slots.  The forward's checkpoint rules must cover both: the gradients stay finite
    and match the oracle's exact adjoint (which stores every fragment's state)."""
    import difftetvr_amd as d
    w = grid_case(12, seed=6)
    col = w["col"].copy()
    col[:, 3] = col[:, 3] * 2.0 + 5.0  # density 5..45
    eye = (0.3, 0.5, 1.9)
    W = H = 96
    mesh, r = make_renderer(d, w["cells"], w["pos"], col, W, H, eye=eye, attenuation=att)
    r.set_checkpoint_interval(interval)
    img = r.render()
    dimg = np.random.default_rng(23).normal(size=(H, W, 4)).astype(np.float32)
    r.render_adjoint(torch.from_numpy(dimg).cuda(), False)
    torch.cuda.synchronize()
    gp = mesh.get_vertex_positions().grad.cpu().numpy()
    gc = mesh.get_vertex_colors().grad.cpu().numpy()
    fr, uf, ud = _oracle_frags(w["cells"], w["pos"], eye, W, H)
    oimg, _, ocnt = oracle.forward(fr, w["pos"], col, uf, att)
    a = img.cpu().numpy()[..., 3]
    print(f"att={att} K={interval}: saturated pixels {(a == 1.0).sum()}, max fragments {ocnt.max()}", flush=True)
    assert (a == 1.0).sum() > 100
    assert float(np.abs(img.cpu().numpy() - oimg).max()) <= IMG_TOL
    check_gradients(f"dense field att={att} K={interval}", gp, gc,
                    adjoint_references(oracle, fr, w["pos"], col, uf, ud, att, dimg))


def test_ppll_renderer_linear_depth_parity(gpu):
    """Renderer() defaults to RendererType.PPLL (Module.cpp:861), whose segment lengths are differences of linear
    eye depths (PPLL/Forward.glsl:53,188): image and exact adjoint against the oracle's linear-depth mode; off
    axis it differs from the INTERSECTION renderer's Euclidean lengths; the reference adjoint mode (the
    INTERSECTION renderer's) is refused for it."""
    import difftetvr_amd as d
    w = grid_case(12, seed=8)
    eye = (1.1, 0.6, 1.3)
    W, H = 128, 112
    mesh = d.TetMesh()
    mesh.set_use_gradients(True)
    mesh.set_tet_mesh_data(w["cells"], w["pos"], w["col"])
    r = d.Renderer()
    assert r.get_renderer_type() == d.RendererType.PPLL
    r.set_tet_mesh(mesh)
    r.set_attenuation(2.0)
    r.set_viewport_size(W, H)
    r.set_camera_fovy(FOVY)
    r.set_view_matrix(view_from(eye))
    r.set_use_early_ray_termination(False)
    img = r.render()
    dimg = np.random.default_rng(12).normal(size=(H, W, 4)).astype(np.float32)
    r.render_adjoint(torch.from_numpy(dimg).cuda(), False)
    torch.cuda.synchronize()
    fr, uf, ud = _oracle_frags(w["cells"], w["pos"], eye, W, H)
    oimg, _, ocnt = oracle.forward(fr, w["pos"], w["col"], uf, 2.0, linear_depth=True)
    gp = mesh.get_vertex_positions().grad.cpu().numpy()
    gc = mesh.get_vertex_colors().grad.cpu().numpy()
    err = float(np.abs(img.cpu().numpy() - oimg).max())
    iimg, _, _ = oracle.forward(fr, w["pos"], w["col"], uf, 2.0)
    print(f"PPLL: max|image - oracle| {err:.2e} (vs the INTERSECTION image {np.abs(iimg - oimg).max():.2e})",
          flush=True)
    assert np.array_equal(r.get_fragment_counts().cpu().numpy(), ocnt)
    assert err <= IMG_TOL
    assert float(np.abs(iimg - oimg).max()) > 1e-3
    check_gradients("PPLL grid12", gp, gc,
                    adjoint_references(oracle, fr, w["pos"], w["col"], uf, ud, 2.0, dimg, linear_depth=True))
    r.set_adjoint_mode("reference")
    with pytest.raises(RuntimeError, match="Euclidean"):
        r.render_adjoint(torch.from_numpy(dimg).cuda(), False)
