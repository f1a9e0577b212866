"""Planned (synchronisation-free) rendering and HIP-graph capture of a whole training step.

Planned renders must equal ordinary renders bit for bit (same records, same order: all depth digits are sorted
instead of the varying ones only); a capacity below the record count must be reported, not crash; and a
captured render + loss + adjoint + Adam step replayed N times must track N eager steps (FP32 atomics reorder the
gradient sums: parameters within 1e-5 relative L2)."""
import numpy as np
import pytest
import torch

from tests.helpers import FOVY, grid_case, make_renderer, rel_l2, view_from

pytestmark = pytest.mark.gpu


def test_planned_render_equals_ordinary(gpu):
    import difftetvr_amd as d
    w = grid_case(n=12, seed=21)
    _, r1 = make_renderer(d, w["cells"], w["pos"], w["col"], 72, 56, eye=(0.4, -0.3, 2.0))
    _, r2 = make_renderer(d, w["cells"], w["pos"], w["col"], 72, 56, eye=(0.4, -0.3, 2.0))
    a = r1.render()
    k = r1.get_stats()["num_pairs"]
    r2.set_record_capacity(k + 100)
    b = r2.render()
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    assert torch.equal(r1.get_fragment_counts(), r2.get_fragment_counts())
    assert r2.records_overflowed() == (False, k)
    assert r2.get_stats()["num_pairs"] == k


def test_planned_capacity_overflow_is_reported(gpu):
    import difftetvr_amd as d
    w = grid_case(n=10, seed=22)
    _, r = make_renderer(d, w["cells"], w["pos"], w["col"], 64, 64, eye=(0.2, 0.5, 1.9))
    r.render()
    k = r.get_stats()["num_pairs"]
    r.set_record_capacity(k // 3)
    r.render()
    over, k2 = r.records_overflowed()
    assert over and k2 == k
    r.set_record_capacity(0)
    r.render()
    assert r.records_overflowed() == (False, k)


def _train_setup(d, w, planned, device_step):
    from difftetvr_amd.optim import FusedAdam
    mesh, r = make_renderer(d, w["cells"], w["pos"], w["col"], 64, 48, eye=(0.5, 0.3, 1.9))
    if planned:
        r.set_record_capacity(200_000)
    opt = FusedAdam([{"params": [mesh.get_vertex_colors()], "lr": 1e-3, "clamp_nonneg": True},
                     {"params": [mesh.get_vertex_positions()], "lr": 1e-4}], zero_grad_after=True,
                    device_step=device_step)
    return mesh, r, opt


def test_graph_captured_training_step(gpu):
    import difftetvr_amd as d
    from difftetvr_amd.optim import mse_loss_grad
    w = grid_case(n=10, seed=23)
    tgt_mesh, tgt_r = make_renderer(d, w["cells"], w["pos_gt"], w["col"], 64, 48, eye=(0.5, 0.3, 1.9))
    target = tgt_r.render().clone()
    steps = 4

    mesh_a, r_a, opt_a = _train_setup(d, w, planned=False, device_step=False)
    for _ in range(steps):
        img = r_a.render()
        _, dimg = mse_loss_grad(img, target)
        r_a.render_adjoint(dimg)
        opt_a.step()

    mesh_b, r_b, opt_b = _train_setup(d, w, planned=True, device_step=True)

    def step():
        img = r_b.render()
        _, dimg = mse_loss_grad(img, target)
        r_b.render_adjoint(dimg)
        opt_b.step()

    step()  # sizes every buffer (and is step 1)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step()
    for _ in range(steps - 1):
        g.replay()
        torch.cuda.synchronize()
    assert r_b.records_overflowed()[0] is False
    pa = mesh_a.get_vertex_positions().detach().cpu().numpy()
    pb = mesh_b.get_vertex_positions().detach().cpu().numpy()
    ca = mesh_a.get_vertex_colors().detach().cpu().numpy()
    cb = mesh_b.get_vertex_colors().detach().cpu().numpy()
    moved = rel_l2(ca, w["col"])
    print(f"graph vs eager: positions {rel_l2(pb, pa):.2e}, colours {rel_l2(cb, ca):.2e} (colours moved {moved:.2e})",
          flush=True)
    assert moved > 1e-5
    assert rel_l2(pb, pa) <= 1e-5
    assert rel_l2(cb, ca) <= 1e-5


def test_graph_captured_batched_step(gpu):
    """The batched multi-view step (one batched render, the summed per-view MSE, one batched adjoint, Adam) in
    planned mode, captured once and replayed, tracks the same steps run eagerly without planning."""
    import difftetvr_amd as d
    from difftetvr_amd import meshgen
    from difftetvr_amd.optim import FusedAdam, mse_loss_grad_views
    w = grid_case(n=10, seed=24)
    views = [meshgen.look_at_view_matrix(tuple(e)) for e in meshgen.fibonacci_sphere(4, 2.0)]
    steps = 4

    def setup(pos, grads, planned):
        m = d.TetMesh()
        m.set_use_gradients(grads)
        m.set_tet_mesh_data(w["cells"], pos, w["col"])
        br = d.BatchRenderer()
        br.set_attenuation(2.0)
        br.set_camera_fovy(FOVY)
        br.set_views(m, views, 56, 48)
        if planned:
            br.set_record_capacity(400_000)
        opt = FusedAdam([{"params": [m.get_vertex_colors()], "lr": 1e-3, "clamp_nonneg": True},
                         {"params": [m.get_vertex_positions()], "lr": 1e-4}], zero_grad_after=True,
                        device_step=planned)
        return m, br, opt

    _, tbr, _ = setup(w["pos_gt"], False, False)
    targets = tbr.render().clone()

    def stepper(br, opt):
        def step():
            imgs = br.render()
            _, dimg = mse_loss_grad_views(imgs, targets)
            br.render_adjoint(dimg)
            opt.step()
        return step

    mesh_a, br_a, opt_a = setup(w["pos"], True, False)
    step_a = stepper(br_a, opt_a)
    for _ in range(steps):
        step_a()
    mesh_b, br_b, opt_b = setup(w["pos"], True, True)
    step_b = stepper(br_b, opt_b)
    step_b()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step_b()
    for _ in range(steps - 1):
        g.replay()
        torch.cuda.synchronize()
    assert br_b.records_overflowed()[0] is False
    pa = mesh_a.get_vertex_positions().detach().cpu().numpy()
    pb = mesh_b.get_vertex_positions().detach().cpu().numpy()
    ca = mesh_a.get_vertex_colors().detach().cpu().numpy()
    cb = mesh_b.get_vertex_colors().detach().cpu().numpy()
    moved = rel_l2(ca, w["col"])
    print(f"batched graph vs eager: positions {rel_l2(pb, pa):.2e}, colours {rel_l2(cb, ca):.2e} "
          f"(colours moved {moved:.2e})", flush=True)
    assert moved > 1e-5
    assert rel_l2(pb, pa) <= 1e-5
    assert rel_l2(cb, ca) <= 1e-5
