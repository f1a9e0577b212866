"""Binning tiles of 16 and 32 pixels (rendered as 8x8 sub-tiles) and 64/256 segments per pass: the same pixels
see the same fragments in the same order (records are sorted by the same keys, stable for equal keys), so
images and fragment lists are bit-identical to 8x8 tiles with 128 segments per pass, and gradients equal up to
the order of their float atomics.  The autotuner times every setting and picks one."""
import numpy as np
import pytest
import torch

from difftetvr_amd import meshgen
from tests.helpers import FOVY, grid_case, make_renderer, rel_l2, view_from

pytestmark = pytest.mark.gpu


def _grads(mesh):
    return mesh.get_vertex_positions().grad.clone(), mesh.get_vertex_colors().grad.clone()


def _zero(mesh):
    mesh.get_vertex_positions().grad.zero_()
    mesh.get_vertex_colors().grad.zero_()


@pytest.mark.parametrize("use_ert", [False, True])
def test_tile_sizes_and_segments_per_pass_agree(gpu, use_ert):
    import difftetvr_amd as d
    w = meshgen.grid_workload(16, seed=0)
    W, H = 203, 171  # not a multiple of any tile size
    mesh, r = make_renderer(d, w["cells"], w["pos"], w["col"], W, H, eye=(0.9, 0.6, 1.5), use_ert=use_ert)
    dimg = torch.randn(H, W, 4, device="cuda")
    ref_img = r.render().clone()
    ref_cnt = r.get_fragment_counts().clone()
    ref_ids = r.get_fragment_ids(96).clone()
    _zero(mesh)
    r.render_adjoint(dimg)
    ref_gp, ref_gc = _grads(mesh)
    for tile in (8, 16, 32):
        for spp in (64, 128, 256):
            r.set_tile_size(tile)
            r.set_segments_per_pass(spp)
            img = r.render()
            assert torch.equal(img, ref_img), (tile, spp)
            assert torch.equal(r.get_fragment_counts(), ref_cnt), (tile, spp)
            assert torch.equal(r.get_fragment_ids(96), ref_ids), (tile, spp)
            _zero(mesh)
            r.render_adjoint(dimg)
            gp, gc = _grads(mesh)
            ep, ec = rel_l2(gp.cpu().numpy(), ref_gp.cpu().numpy()), rel_l2(gc.cpu().numpy(), ref_gc.cpu().numpy())
            print(f"tile {tile} spp {spp}: gradients vs 8x8/128 rel L2 pos {ep:.1e} col {ec:.1e}", flush=True)
            assert ep < 1e-5 and ec < 1e-6, (tile, spp, ep, ec)


def test_batch_with_large_tiles(gpu):
    """A batch of views of different sizes binned in 32-pixel tiles equals its single 8x8 renders."""
    import difftetvr_amd as d
    w = meshgen.grid_workload(10, seed=3)
    mesh = d.TetMesh()
    mesh.set_tet_mesh_data(w["cells"], w["pos"], w["col"])
    sizes = [(64, 48), (100, 70), (33, 90)]
    eyes = [(0.0, 0.3, 2.0), (1.3, 0.4, 1.2), (-0.8, -0.9, 1.4)]
    singles = []
    for (W, H), e in zip(sizes, eyes):
        _, r = make_renderer(d, w["cells"], w["pos"], w["col"], W, H, eye=e, use_gradients=False)
        singles.append(r.render().clone())
    br = d.BatchRenderer()
    br.set_attenuation(1.0)
    br.set_camera_fovy(FOVY)
    br.set_tile_size(32)
    br.set_segments_per_pass(64)
    br.set_items([(mesh, view_from(e), W, H) for (W, H), e in zip(sizes, eyes)])
    for a, b in zip(br.render(), singles):
        assert torch.equal(a, b)


def test_autotune_picks_a_measured_setting(gpu):
    import difftetvr_amd as d
    w = meshgen.grid_workload(24, seed=1)
    mesh, r = make_renderer(d, w["cells"], w["pos"], w["col"], 256, 256, eye=(0.9, 0.6, 1.5))
    ref = r.render().clone()
    r.set_autotune(True, with_adjoint=True, repeats=2)
    img = r.render()
    assert torch.equal(img, ref)
    cache = d.autotune_cache()
    assert len(cache) >= 1
    choice = list(cache.values())[-1]
    assert len(choice["candidates"]) == 9 and choice["ms"] > 0 and choice["ms_default"] > 0
    assert choice["ms"] == min(c[2] for c in choice["candidates"])
    assert r.get_tuning() == (choice["tile_size"], choice["segments_per_pass"])
    print(f"autotune 24^3 grid at 256^2: {choice['tile_size']} px tiles, {choice['segments_per_pass']} segments per "
          f"pass, {choice['ms']:.3f} ms vs {choice['ms_default']:.3f} ms at 8/128", flush=True)


def test_gpu_queue_launch_matches_grid(gpu):
    """Persistent tile kernels taking tiles from a device work queue: images, fragment counts and fragment lists
    identical to one workgroup per tile, gradients equal up to the order of the float atomics, for a single view
    and for a batch of two meshes at different viewports; repeated adjoints of one render each claim every tile."""
    import difftetvr_amd as d
    w = grid_case(n=12, seed=31)
    res = {}
    dimg = torch.rand((72, 80, 4), generator=torch.Generator().manual_seed(5)).to(gpu)
    for launch in ("grid", "queue"):
        mesh, r = make_renderer(d, w["cells"], w["pos"], w["col"], 80, 72, eye=(0.6, 0.4, 1.8))
        r.set_tile_launch(launch)
        img = r.render()
        r.render_adjoint(dimg)
        r.render_adjoint(dimg)  # a second adjoint of the same render accumulates the same gradients again
        torch.cuda.synchronize()
        res[launch] = (img.cpu(), r.get_fragment_counts().cpu(), r.get_fragment_ids(48).cpu(),
                       mesh.get_vertex_positions().grad.cpu().numpy(), mesh.get_vertex_colors().grad.cpu().numpy())
    g, q = res["grid"], res["queue"]
    assert torch.equal(g[0], q[0]) and torch.equal(g[1], q[1]) and torch.equal(g[2], q[2])
    print(f"queue vs grid gradients: positions {rel_l2(q[3], g[3]):.2e}, colours {rel_l2(q[4], g[4]):.2e}",
          flush=True)
    assert rel_l2(q[3], g[3]) <= 1e-5 and rel_l2(q[4], g[4]) <= 1e-5

    w2 = grid_case(n=8, seed=32)
    outs = {}
    for launch in ("grid", "queue"):
        m1 = d.TetMesh()
        m1.set_use_gradients(True)
        m1.set_tet_mesh_data(w["cells"], w["pos"], w["col"])
        m2 = d.TetMesh()
        m2.set_use_gradients(True)
        m2.set_tet_mesh_data(w2["cells"], w2["pos"], w2["col"])
        br = d.BatchRenderer()
        br.set_attenuation(2.0)
        br.set_camera_fovy(FOVY)
        br.set_items([(m1, view_from((0.5, 0.2, 1.9)), 64, 48), (m2, view_from((-0.4, 0.6, 1.7)), 40, 56)])
        br.set_tile_launch(launch)
        imgs = br.render()
        br.render_adjoint([torch.ones_like(imgs[0]), torch.ones_like(imgs[1])])
        torch.cuda.synchronize()
        outs[launch] = ([i.cpu() for i in imgs], m1.get_vertex_positions().grad.cpu().numpy(),
                        m2.get_vertex_colors().grad.cpu().numpy())
    (ig, pg, cg), (iq, pq, cq) = outs["grid"], outs["queue"]
    assert all(torch.equal(a, b) for a, b in zip(ig, iq))
    assert rel_l2(pq, pg) <= 1e-5 and rel_l2(cq, cg) <= 1e-5
