"""Render server (dtv_server_*): every result must equal the request rendered alone (images bit-identical: a batch
renders each item exactly as a single render), whether requests are served one at a time or mixed across
meshes and resolutions in concurrent batches; differentiated requests accumulate the same gradients (FP32
atomics reorder the sums: relative L2 <= 1e-5)."""
import math

import numpy as np
import pytest
import torch

from difftetvr_amd import meshgen
from tests.helpers import FOVY, grid_case, rel_l2, view_from

pytestmark = pytest.mark.gpu


def _meshes(d):
    ws = [grid_case(n=8, seed=11), grid_case(n=12, seed=12, jitter=0.3)]
    pts, cells = meshgen.delaunay_ball(1500, seed=13)
    ws.append({"pos": pts, "cells": cells, "col": meshgen.blob_colors(pts, seed=14)})
    out = []
    for w in ws:
        m = d.TetMesh()
        m.set_use_gradients(True)
        m.set_tet_mesh_data(w["cells"], w["pos"], w["col"])
        out.append(m)
    return out


def _requests():
    rng = np.random.default_rng(5)
    sizes = [(64, 64), (96, 80), (128, 128), (40, 56)]
    reqs = []
    for i in range(14):
        eye = rng.normal(size=3)
        eye = tuple(eye / np.linalg.norm(eye) * rng.uniform(1.6, 2.6))
        reqs.append((i % 3, view_from(eye), *sizes[i % len(sizes)]))
    return reqs


def _single(d, mesh, vm, W, H):
    r = d.Renderer(renderer_type=d.RendererType.INTERSECTION)
    r.set_tet_mesh(mesh)
    r.set_attenuation(1.0)
    r.set_camera_fovy(FOVY)
    r.set_viewport_size(W, H)
    r.set_view_matrix(vm)
    r.set_use_early_ray_termination(False)
    img = r.render()
    torch.cuda.synchronize()
    return img


def test_server_results_equal_alone_and_in_the_mix(gpu):
    import difftetvr_amd as d
    from difftetvr_amd.serve import Server
    meshes = _meshes(d)
    reqs = _requests()
    srv = Server(num_workers=2, max_batch=8, fovy=FOVY)
    ids = [srv.add_mesh(m) for m in meshes]
    alone = []
    for mi, vm, W, H in reqs:
        img, lat, _ = srv.wait(srv.submit(ids[mi], vm, W, H))
        alone.append(img.clone())
        assert lat > 0.0
    tickets = [srv.submit(ids[mi], vm, W, H) for mi, vm, W, H in reqs]
    mixed = [srv.wait(t)[0] for t in tickets]
    st = srv.stats()
    srv.close()
    print(f"server stats {st}", flush=True)
    assert st["completed"] == 2 * len(reqs)
    assert st["batches"] < 2 * len(reqs)      # the mixed requests were batched
    for (mi, vm, W, H), a, b in zip(reqs, alone, mixed):
        assert torch.equal(a, b)
        assert torch.equal(a, _single(d, meshes[mi], vm, W, H))


def test_server_differentiated_requests(gpu):
    import difftetvr_amd as d
    from difftetvr_amd.serve import Server
    meshes = _meshes(d)
    reqs = _requests()[:6]
    gen = torch.Generator(device="cuda").manual_seed(1)
    targets = [torch.rand((H, W, 4), device="cuda", generator=gen) for _, _, W, H in reqs]
    torch.cuda.synchronize()
    srv = Server(num_workers=2, max_batch=4, fovy=FOVY)
    ids = [srv.add_mesh(m) for m in meshes]
    # one at a time
    for (mi, vm, W, H), t in zip(reqs, targets):
        srv.wait(srv.submit(ids[mi], vm, W, H, target=t))
    g1 = [(m.get_vertex_positions().grad.clone(), m.get_vertex_colors().grad.clone()) for m in meshes]
    for m in meshes:
        m.get_vertex_positions().grad.zero_()
        m.get_vertex_colors().grad.zero_()
    torch.cuda.synchronize()
    # mixed with forward-only requests (released when the batch's forward is done, before its adjoint)
    tickets = []
    for (mi, vm, W, H), t in zip(reqs, targets):
        tickets.append((srv.submit(ids[mi], vm, W, H, target=t), None))
        tickets.append((srv.submit(ids[(mi + 1) % 3], vm, W, H), ((mi + 1) % 3, vm, W, H)))
    fwd_only = [(srv.wait(t)[0], spec) for t, spec in tickets]
    srv.close()
    for img, spec in fwd_only:
        if spec is not None:
            mi, vm, W, H = spec
            assert torch.equal(img, _single(d, meshes[mi], vm, W, H))
    for (gp, gc), m in zip(g1, meshes):
        assert float(gc.abs().sum()) > 0.0
        assert rel_l2(m.get_vertex_positions().grad.cpu().numpy(), gp.cpu().numpy()) <= 1e-5
        assert rel_l2(m.get_vertex_colors().grad.cpu().numpy(), gc.cpu().numpy()) <= 1e-5


def test_server_validation(gpu):
    import difftetvr_amd as d
    from difftetvr_amd.serve import Server
    srv = Server(num_workers=1, max_batch=2, fovy=FOVY)
    with pytest.raises(d.DtvError):
        srv.submit(0, view_from((0, 0, 2)), 16, 16)   # no mesh registered
    m = _meshes(d)[0]
    mid = srv.add_mesh(m)
    img, _, _ = srv.wait(srv.submit(mid, view_from((0.3, 0.2, 2.0)), 16, 16))
    assert img.shape == (16, 16, 4)
    srv.close()
    with pytest.raises(d.DtvError):
        Server(num_workers=0)
