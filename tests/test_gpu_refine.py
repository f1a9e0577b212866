"""Coarse-to-fine step on the GPU path: accumulate |gradients|, split_by_largest_gradient_magnitudes
(Module.cpp:820-832), render again.  Refinement inserts edge midpoints with interpolated colours, so a field of
constant colour and density renders the same image before and after (alpha = 1 - exp(-sigma * chord) whatever
the tessellation); only pixels whose centres sit on a silhouette edge may change, because the new vertices snap
to the 1/256-pixel grid."""
import numpy as np
import pytest
import torch

from tests.helpers import grid_case, make_renderer

pytestmark = pytest.mark.gpu


def test_split_keeps_a_constant_field_image_and_training_continues(gpu):
    import difftetvr_amd as d
    from difftetvr_amd.optim import FusedAdam, mse_loss_grad
    w = grid_case(n=8, seed=31)
    col = np.tile(np.array([[0.7, 0.3, 0.2, 3.0]], np.float32), (w["pos"].shape[0], 1))
    mesh, r = make_renderer(d, w["cells"], w["pos"], col, 96, 80, eye=(0.6, 0.5, 1.8))
    before = r.render().clone()
    dimg = torch.randn_like(before)
    r.render_adjoint(dimg, True)                     # |gradients|, as train.py:393-404 accumulates them
    t0, v0 = mesh.get_num_cells(), mesh.get_num_vertices()
    old_pos = mesh.get_vertex_positions()
    mesh.split_by_largest_gradient_magnitudes(r, d.SplitGradientType.ABS_COLOR, 0.1)
    assert mesh.get_num_cells() > t0 and mesh.get_num_vertices() == int(mesh.get_vertex_positions().shape[0]) > v0
    assert mesh.get_vertex_positions() is not old_pos
    assert torch.equal(mesh.get_vertex_positions()[:v0].detach(), old_pos[:v0].detach())
    after = r.render().clone()
    torch.cuda.synchronize()
    diff = (after - before).abs().cpu().numpy().max(axis=2)
    frac = float((diff <= 2e-5).mean())
    print(f"split {t0}->{mesh.get_num_cells()} tets; pixels within 2e-5: {frac:.4f}, max diff {diff.max():.2e}",
          flush=True)
    assert frac >= 0.995
    assert float(diff.mean()) <= 1e-4
    # the optimiser takes the new tensors and a step runs on the refined mesh
    opt = FusedAdam([{"params": [mesh.get_vertex_colors()], "lr": 1e-3, "clamp_nonneg": True},
                     {"params": [mesh.get_vertex_positions()], "lr": 1e-4}], zero_grad_after=True)
    img = r.render()
    loss, dl = mse_loss_grad(img, before)
    r.render_adjoint(dl)
    g = mesh.get_vertex_colors().grad.clone()
    opt.step()
    torch.cuda.synchronize()
    assert g.shape == (mesh.get_num_vertices(), 4) and float(g.abs().sum()) > 0.0
