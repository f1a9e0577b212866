"""The batched and serving workloads at their real sizes.

Multi-view (16 cameras on a Fibonacci sphere of radius 2, 512 x 512 each, the 64^3 x 6 = 1,572,864-tet grid,
one batch):
  * constant colour c and density s: every view's pixels are c * (1 - exp(-s * chord)) against the exact cube
    chord, and the batched adjoint's colour and density gradient sums equal their closed forms (the identities of
    tests/test_gpu_fullsize.py, summed over the views);
  * the seeded blob field: sampled rows of four views against the oracle (images 1e-5; the adjoint with the image
    adjoint nonzero on those rows only, 1e-4 relative L2).
Uneven batch (8 Delaunay meshes of 2k..512k seeded points in the unit ball, 3-octave value-noise density in
[0, 40], attenuation 1, one 768 x 768 view each, early ray termination off: rays cross hundreds of tets and
saturate): every item's sampled rows against the oracle, images and adjoint.
Serving (a subset of the 32-mesh mix: grids of 16^3, 48^3 and 80^3 cells and Delaunay meshes of 10k, 100k and
500k points, viewports of 256^2, 512^2 and 1024^2, some requests differentiated): every result equals the
request rendered alone, served one at a time and mixed.
"""
import math

import numpy as np
import pytest
import torch

import oracle
from difftetvr_amd import meshgen
from tests.helpers import FOVY, ZFAR, ZNEAR, adjoint_references, check_gradients, rel_l2, view_from

pytestmark = pytest.mark.gpu

COLOR = np.array([0.7, 0.3, 0.2], np.float32)
SIGMA = 2.0


def _cube_chord(eye, d):
    with np.errstate(divide="ignore", invalid="ignore"):
        t1 = (-0.5 - eye) / d
        t2 = (0.5 - eye) / d
    return max(0.0, float(np.min(np.maximum(t1, t2)) - np.max(np.minimum(t1, t2))))


def _batch(d, items, att=1.0, ert=False):
    br = d.BatchRenderer()
    br.set_attenuation(att)
    br.set_camera_fovy(FOVY)
    br.set_use_early_ray_termination(ert)
    br.set_items(items)
    return br


def _mesh(d, cells, pos, col):
    m = d.TetMesh()
    m.set_use_gradients(True)
    m.set_tet_mesh_data(cells, pos, col)
    return m


# ---- multi-view ------------------------------------------------------------------------------------------
MV_RES = 512
MV_EYES = [tuple(e) for e in meshgen.fibonacci_sphere(16, 2.0)]


@pytest.fixture(scope="module")
def grid64():
    return meshgen.grid_workload(64, seed=0)


def test_multiview_constant_field(gpu, grid64):
    import difftetvr_amd as d
    w = grid64
    col = np.tile(np.append(COLOR, np.float32(SIGMA)), (w["pos"].shape[0], 1)).astype(np.float32)
    m = _mesh(d, w["cells"], w["pos"], col)
    br = _batch(d, [(m, view_from(e), MV_RES, MV_RES) for e in MV_EYES])
    imgs = br.render()
    torch.cuda.synchronize()
    im = imgs.cpu().numpy()
    worst_frac, nchk = 1.0, 0
    for v, e in enumerate(MV_EYES):
        uf, _ = oracle.camera_uniforms(view_from(e), FOVY, ZNEAR, ZFAR, MV_RES, MV_RES)
        eye = np.asarray(e, np.float64)
        errs = []
        for y in range(5, MV_RES, 24):
            for x in range(3, MV_RES, 24):
                chord = _cube_chord(eye, oracle.ray_dir(uf, MV_RES, MV_RES, x, y).astype(np.float64))
                a = 1.0 - math.exp(-SIGMA * chord)
                errs.append(max(abs(im[v, y, x, 3] - a), float(np.abs(im[v, y, x, :3] - COLOR * a).max())))
        errs = np.array(errs)
        nchk += errs.size
        worst_frac = min(worst_frac, float((errs <= 2e-5).mean()))
        assert errs.max() <= 1e-2, (v, float(errs.max()))
    print(f"multiview constant field: {nchk} pixels over 16 views, worst view within 2e-5: {worst_frac:.4f}",
          flush=True)
    assert worst_frac >= 0.99
    # the batched adjoint's gradient sums
    gen = torch.Generator(device="cuda").manual_seed(9)
    dimg = torch.rand(imgs.shape, device="cuda", generator=gen) + 0.5
    br.render_adjoint(dimg)
    torch.cuda.synchronize()
    gcol = m.get_vertex_colors().grad.to(torch.float64).sum(dim=0).cpu().numpy()
    a = imgs[..., 3].to(torch.float64)
    di = dimg.to(torch.float64)
    expect_rgb = (di[..., :3] * a[..., None]).sum(dim=(0, 1, 2)).cpu().numpy()
    one_minus = (1.0 - a).clamp_min(1e-30)
    chord = -torch.log(one_minus) / SIGMA
    c = torch.tensor(COLOR, dtype=torch.float64, device="cuda")
    expect_s = float(((di[..., 3] + (di[..., :3] * c).sum(dim=-1)) * one_minus * chord).sum())
    rel_rgb = np.abs(gcol[:3] - expect_rgb) / np.abs(expect_rgb)
    rel_s = abs(gcol[3] - expect_s) / abs(expect_s)
    print(f"multiview gradient sums: colour rel {rel_rgb}, density rel {rel_s:.2e}", flush=True)
    assert (rel_rgb <= 1e-4).all()
    assert rel_s <= 1e-3


def test_multiview_blob_field_rows(gpu, grid64):
    import difftetvr_amd as d
    w = grid64
    m = _mesh(d, w["cells"], w["pos"], w["col"])
    br = _batch(d, [(m, view_from(e), MV_RES, MV_RES) for e in MV_EYES])
    imgs = br.render()
    torch.cuda.synchronize()
    im = imgs.cpu().numpy()
    views, rows = (0, 5, 10, 15), (131, 389)
    rng = np.random.default_rng(3)
    dimg = np.zeros(im.shape, np.float32)
    refs = None
    err_img = 0.0
    for v in views:
        uf, ud = oracle.camera_uniforms(view_from(MV_EYES[v]), FOVY, ZNEAR, ZFAR, MV_RES, MV_RES)
        fr = oracle.Fragments(w["cells"], w["pos"], uf, ZNEAR, MV_RES, MV_RES, rows[0], rows[1] + 1,
                              rows[1] - rows[0])
        oimg, _, _ = oracle.forward(fr, w["pos"], w["col"], uf, 1.0)
        for y in rows:
            err_img = max(err_img, float(np.abs(im[v, y] - oimg[y]).max()))
            dimg[v, y] = rng.normal(size=(MV_RES, 4))
        r_ = adjoint_references(oracle, fr, w["pos"], w["col"], uf, ud, 1.0, dimg[v])
        refs = r_ if refs is None else tuple(a + b for a, b in zip(refs, r_))
    br.render_adjoint(torch.from_numpy(dimg).cuda())
    torch.cuda.synchronize()
    gp = m.get_vertex_positions().grad.cpu().numpy()
    gc = m.get_vertex_colors().grad.cpu().numpy()
    print(f"multiview blob rows: image max diff {err_img:.2e}", flush=True)
    assert err_img <= 1e-5
    check_gradients("multiview blob rows", gp, gc, refs)


# ---- uneven Delaunay batch ---------------------------------------------------------------------------------
UNEVEN_NS = (2_000, 8_000, 32_000, 64_000, 128_000, 256_000, 384_000, 512_000)
UNEVEN_RES = 768


@pytest.fixture(scope="module")
def delaunay8():
    out = []
    rng = np.random.default_rng(1)
    for i, n in enumerate(UNEVEN_NS):
        pts, cells = meshgen.delaunay_ball(n, seed=500 + i)
        col = meshgen.blob_colors(pts, seed=600 + i)
        col[:, 3] = meshgen.value_noise(pts, seed=700 + i, vmax=40.0)
        e = rng.normal(size=3)
        out.append((pts, cells, col, tuple(e / np.linalg.norm(e) * 2.5)))
    return out


def test_uneven_delaunay_batch_rows(gpu, delaunay8):
    import difftetvr_amd as d
    meshes = [_mesh(d, cells, pts, col) for pts, cells, col, _ in delaunay8]
    br = _batch(d, [(m, view_from(e), UNEVEN_RES, UNEVEN_RES) for m, (_, _, _, e) in zip(meshes, delaunay8)])
    imgs = br.render()
    counts = br.get_fragment_counts()
    torch.cuda.synchronize()
    im = imgs.cpu().numpy()
    maxc = counts.view(len(meshes), -1).max(dim=1).values.cpu().numpy()
    rows = (300, 384)
    rng = np.random.default_rng(4)
    dimg = np.zeros(im.shape, np.float32)
    orcs = []
    for i, (pts, cells, col, e) in enumerate(delaunay8):
        uf, ud = oracle.camera_uniforms(view_from(e), FOVY, ZNEAR, ZFAR, UNEVEN_RES, UNEVEN_RES)
        fr = oracle.Fragments(cells, pts, uf, ZNEAR, UNEVEN_RES, UNEVEN_RES, rows[0], rows[1] + 1, rows[1] - rows[0])
        oimg, _, ocnt = oracle.forward(fr, pts, col, uf, 1.0)
        err = max(float(np.abs(im[i, y] - oimg[y]).max()) for y in rows)
        assert np.array_equal(counts[i].cpu().numpy()[list(rows)], ocnt[list(rows)]), i
        for y in rows:
            dimg[i, y] = rng.normal(size=(UNEVEN_RES, 4))
        orcs.append((err, adjoint_references(oracle, fr, pts, col, uf, ud, 1.0, dimg[i])))
    br.render_adjoint(torch.from_numpy(dimg).cuda())
    torch.cuda.synchronize()
    for i, (m, (err, refs)) in enumerate(zip(meshes, orcs)):
        gp = m.get_vertex_positions().grad.cpu().numpy()
        gc = m.get_vertex_colors().grad.cpu().numpy()
        a = im[i, ..., 3]
        print(f"Delaunay {UNEVEN_NS[i]} points ({m.get_num_cells()} tets): max fragments per pixel {maxc[i]}, "
              f"saturated pixels {(a == 1.0).sum()}, rows image diff {err:.2e}", flush=True)
        assert err <= 1e-5
        check_gradients(f"Delaunay {UNEVEN_NS[i]}", gp, gc, refs)
    assert maxc.max() >= 200  # depth complexity in the hundreds


# ---- serving mix -------------------------------------------------------------------------------------------
def test_serving_mix_equal_alone_and_mixed(gpu):
    import difftetvr_amd as d
    from difftetvr_amd.serve import Server
    meshes = []
    for i, n in enumerate((16, 48, 80)):
        w = meshgen.grid_workload(n, seed=i)
        meshes.append(_mesh(d, w["cells"], w["pos"], w["col"]))
    for i, n in enumerate((10_000, 100_000, 500_000)):
        pts, cells = meshgen.delaunay_ball(n, seed=100 + i)
        col = meshgen.blob_colors(pts, seed=200 + i)
        col[:, 3] = meshgen.value_noise(pts, seed=300 + i, vmax=40.0)
        meshes.append(_mesh(d, cells, pts, col))
    rng = np.random.default_rng(0)
    reqs = []
    for k in range(18):
        e = rng.normal(size=3)
        e = e / np.linalg.norm(e) * rng.uniform(1.5, 3.0)
        res = (256, 512, 1024)[k % 3]
        reqs.append((k % len(meshes), view_from(tuple(e)), res, k % 6 == 5))
    srv = Server(num_workers=4, max_batch=8, memory_budget=16 << 30, attenuation=1.0, fovy=FOVY,
                 use_early_ray_termination=True)
    ids = [srv.add_mesh(m) for m in meshes]
    targets = {s: torch.zeros((s, s, 4), device="cuda") for s in (256, 512, 1024)}

    def grads_zero():
        for m in meshes:
            m.get_vertex_positions().grad.zero_()
            m.get_vertex_colors().grad.zero_()
        torch.cuda.synchronize()

    def submit(mi, vm, res, train):
        if train:
            return srv.submit(ids[mi], vm, res, res, target=targets[res])
        return srv.submit(ids[mi], vm, res, res)

    grads_zero()
    alone, g_alone = [], []
    for mi, vm, res, train in reqs:
        img, _, _ = srv.wait(submit(mi, vm, res, train))
        alone.append(img.clone())
    torch.cuda.synchronize()
    g_alone = [(m.get_vertex_positions().grad.clone(), m.get_vertex_colors().grad.clone()) for m in meshes]
    grads_zero()
    tickets = [submit(mi, vm, res, train) for mi, vm, res, train in reqs]
    mixed = [srv.wait(t)[0] for t in tickets]
    torch.cuda.synchronize()
    st = srv.stats()
    srv.close()
    print(f"serving mix: {st}", flush=True)
    assert st["max_batch_seen"] > 1
    for k, ((mi, vm, res, train), a, b) in enumerate(zip(reqs, alone, mixed)):
        assert torch.equal(a, b), k
    for m, (gp, gc) in zip(meshes, g_alone):
        if float(gc.abs().sum()) == 0.0:
            continue
        assert rel_l2(m.get_vertex_colors().grad.cpu().numpy(), gc.cpu().numpy()) <= 1e-5
        assert rel_l2(m.get_vertex_positions().grad.cpu().numpy(), gp.cpu().numpy()) <= 1e-5
    # and alone equals a plain single-view render of the request
    for (mi, vm, res, train), a in list(zip(reqs, alone))[:6]:
        r = d.Renderer(renderer_type=d.RendererType.INTERSECTION)
        r.set_tet_mesh(meshes[mi])
        r.set_attenuation(1.0)
        r.set_camera_fovy(FOVY)
        r.set_viewport_size(res, res)
        r.set_view_matrix(vm)
        r.set_use_early_ray_termination(True)
        assert torch.equal(r.render(), a)
