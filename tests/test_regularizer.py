"""Tet-quality regulariser (Data/Shaders/Optimizer/TetRegularizer.glsl:108-166).

The oracle's gradient follows the reference's explicit dq/dp expressions; it is pinned here by the closed form
of the regular tetrahedron (q = 1, stationary), the invariances of q (translation, rotation, scale), FP64
central differences of the loss written from the definition of q, and the reference's S^{5/2} >= 1e-5 clamp.
The GPU kernel is then held to the oracle: relative L2 <= 1e-5 against FP64 (FP32 atomics reorder the sums).
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from difftetvr_amd import meshgen
from tests.helpers import grid_case, rel_l2

LAM, BETA = 0.7, 3.0


def _regular_tet():
    # reference orientation: dot(cross(p1 - p0, p2 - p0), p3 - p0) < 0
    pos = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], np.float64)
    return np.array([[0, 1, 2, 3]], np.int32), pos


def _random_tet(rng, scale=1.0):
    p = rng.normal(size=(4, 3))
    if np.dot(np.cross(p[1] - p[0], p[2] - p[0]), p[3] - p[0]) > 0:
        p[[1, 2]] = p[[2, 1]]
    return np.array([[0, 1, 2, 3]], np.int32), p * scale


def _fd_grad(cells, pos, h):
    g = np.zeros_like(pos)
    for v in range(pos.shape[0]):
        for k in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[v, k] += h
            pm[v, k] -= h
            g[v, k] = (oracle.tet_regularizer_loss(cells, pp, LAM, BETA)
                       - oracle.tet_regularizer_loss(cells, pm, LAM, BETA)) / (2 * h)
    return g


def test_regular_tet_closed_form():
    cells, pos = _regular_tet()
    # q = 1 for the regular tetrahedron, so the loss is (lambda / beta) log(1 + e^-beta)
    assert oracle.tet_regularizer_loss(cells, pos, LAM, BETA) == pytest.approx(LAM / BETA * np.log1p(np.exp(-BETA)),
                                                                                rel=1e-12)
    # q is maximal there: the gradient vanishes
    g = oracle.tet_regularizer_grad(cells, pos, LAM, BETA, dtype=np.float64)
    assert np.abs(g).max() < 1e-12
    # a flat tet has q = 0: loss (lambda / beta) log 2
    flat = pos.copy()
    flat[3] = (flat[0] + flat[1] + flat[2]) / 3
    assert oracle.tet_regularizer_loss(cells, flat, LAM, BETA) == pytest.approx(LAM / BETA * np.log(2.0), rel=1e-12)


def test_invariances_of_quality():
    rng = np.random.default_rng(3)
    for _ in range(8):
        cells, p = _random_tet(rng)
        g = oracle.tet_regularizer_grad(cells, p, LAM, BETA, dtype=np.float64)
        norm = np.abs(g).max()
        assert norm > 1e-6
        assert np.abs(g.sum(0)).max() < 1e-12 * max(norm, 1)                 # translation
        assert np.abs(np.cross(p, g).sum(0)).max() < 1e-11 * max(norm, 1)    # rotation
        assert abs(float((p * g).sum())) < 1e-11 * max(norm, 1)              # scale


def test_gradient_matches_fp64_central_differences():
    w = grid_case(n=3, seed=5, jitter=0.25)
    cells, pos = w["cells"], w["pos"].astype(np.float64)
    g = oracle.tet_regularizer_grad(cells, pos, LAM, BETA, dtype=np.float64)
    fd = _fd_grad(cells, pos, 1e-6)
    assert rel_l2(g, fd) < 1e-7


def test_small_tet_denominator_clamp():
    # S^{5/2} < 1e-5: the reference divides by 1e-5 instead, which scales the exact gradient by S^{5/2} / 1e-5
    rng = np.random.default_rng(11)
    cells, p = _random_tet(rng, scale=0.01)
    S = sum(float(np.sum((p[i] - p[j]) ** 2)) for i in range(4) for j in range(i + 1, 4))
    assert S ** 2.5 < 1e-5
    g = oracle.tet_regularizer_grad(cells, p, LAM, BETA, dtype=np.float64)
    fd = _fd_grad(cells, p, 1e-9)
    assert rel_l2(g, fd * (S ** 2.5 / 1e-5)) < 1e-6


def test_float_matches_double():
    w = grid_case(n=8, seed=2, jitter=0.3)
    gf = oracle.tet_regularizer_grad(w["cells"], w["pos"], LAM, BETA, dtype=np.float32)
    gd = oracle.tet_regularizer_grad(w["cells"], w["pos"].astype(np.float64), LAM, BETA, dtype=np.float64)
    assert rel_l2(gf, gd) < 1e-5


# ---- GPU ---------------------------------------------------------------------------------------------------

def _gpu_mesh(d, cells, pos, col):
    mesh = d.TetMesh()
    mesh.set_use_gradients(True)
    mesh.set_tet_mesh_data(cells, pos, col)
    return mesh


@pytest.mark.gpu
@pytest.mark.parametrize("scale", [1.0, 0.004])
def test_gpu_regularizer_parity(gpu, scale):
    import difftetvr_amd as d
    w = grid_case(n=16, seed=7, jitter=0.3)
    pos = (w["pos"] * scale).astype(np.float32)
    mesh = _gpu_mesh(d, w["cells"], pos, w["col"])
    reg = d.TetRegularizer(mesh, LAM, BETA)
    reg.compute_grad()
    got = mesh.get_vertex_positions().grad.cpu().numpy()
    want = oracle.tet_regularizer_grad(w["cells"], pos.astype(np.float64), LAM, BETA, dtype=np.float64)
    print(f"regularizer scale={scale} rel_l2={rel_l2(got, want):.3e}", flush=True)
    assert rel_l2(got, want) < 1e-5
    # accumulates (+=) into .grad
    reg.compute_grad()
    assert rel_l2(mesh.get_vertex_positions().grad.cpu().numpy(), 2 * want) < 1e-5


@pytest.mark.gpu
def test_gpu_regularizer_in_differentiable_renderer(gpu):
    import difftetvr_amd as d
    from tests.helpers import make_renderer
    w = grid_case(n=8, seed=1, jitter=0.2)
    mesh, r = make_renderer(d, w["cells"], w["pos"], w["col"], 48, 40, eye=(0.4, 0.3, 2.1))
    dr = d.DifferentiableRenderer(r, True, LAM, BETA)
    img = dr()
    (img * torch.linspace(0, 1, img.numel(), device=img.device).reshape(img.shape)).sum().backward()
    both = mesh.get_vertex_positions().grad.cpu().numpy().copy()
    # the same without the regulariser
    mesh2, r2 = make_renderer(d, w["cells"], w["pos"], w["col"], 48, 40, eye=(0.4, 0.3, 2.1))
    img2 = d.DifferentiableRenderer(r2, False, 0.0, 1.0)()
    (img2 * torch.linspace(0, 1, img2.numel(), device=img2.device).reshape(img2.shape)).sum().backward()
    render_only = mesh2.get_vertex_positions().grad.cpu().numpy()
    want = oracle.tet_regularizer_grad(w["cells"], w["pos"].astype(np.float64), LAM, BETA, dtype=np.float64)
    assert rel_l2(both - render_only, want) < 1e-4


@pytest.mark.gpu
def test_gpu_regularizer_empty_mesh(gpu):
    from difftetvr_amd import _lib
    lib = _lib.load()
    g = torch.zeros(8, 3, device="cuda")
    desc = _lib.DtvMesh(0, 0, 0, 0, 0)
    _lib.check(lib.dtv_tet_regularizer_grad(ctypes.byref(desc), LAM, BETA, ctypes.c_void_p(g.data_ptr()), None))
    torch.cuda.synchronize()
    assert float(g.abs().sum()) == 0.0
    assert lib.dtv_tet_regularizer_grad(None, LAM, BETA, ctypes.c_void_p(g.data_ptr()), None) != 0
