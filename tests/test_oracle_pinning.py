"""Pins the CPU oracle (oracle/oracle.c) with checks that do not retype its formulas: closed forms, numpy
float64 restatements of the camera, brute force, invariants, scipy quadrature and FP64 finite differences."""
import math

import numpy as np
import pytest

import oracle
from difftetvr_amd import meshgen
from tests.helpers import FOVY, ZFAR, ZNEAR, view_from


def _persp_np(fovy, aspect, n, f):
    """glm::perspectiveRH_ZO with the top-left-origin y flip, as a row-major float64 matrix."""
    t = math.tan(fovy / 2)
    P = np.zeros((4, 4))
    P[0, 0] = 1 / (aspect * t)
    P[1, 1] = -1 / t
    P[2, 2] = f / (n - f)
    P[2, 3] = -(f * n) / (f - n)
    P[3, 2] = -1.0
    return P


def test_camera_uniforms_closed_form():
    eye = (0.7, -0.4, 1.9)
    view = view_from(eye)
    uf, ud = oracle.camera_uniforms(view, FOVY, ZNEAR, ZFAR, 320, 200)
    V = np.asarray(view, np.float64).reshape(4, 4).T  # column-major -> row-major
    P = _persp_np(float(np.float32(FOVY)), 320 / 200, float(np.float32(ZNEAR)), float(np.float32(ZFAR)))
    VP = (P @ V).T.reshape(-1)
    assert np.allclose(ud[0:16], VP, rtol=0, atol=1e-12)
    assert np.allclose(ud[32:48].reshape(4, 4).T @ P, np.eye(4), atol=1e-9)
    assert np.allclose(ud[48:64].reshape(4, 4).T @ V, np.eye(4), atol=1e-9)
    assert np.allclose(ud[64:67], eye, atol=1e-6)
    assert np.array_equal(uf, ud.astype(np.float32))


def test_ray_direction_points_through_pixel_centre():
    W, H = 64, 48
    eye = np.array([0.3, 0.2, 2.0])
    uf, ud = oracle.camera_uniforms(view_from(eye), FOVY, ZNEAR, ZFAR, W, H)
    VP = ud[0:16].reshape(4, 4).T
    for (x, y) in [(0, 0), (63, 47), (31, 20), (10, 40)]:
        d = oracle.ray_dir(uf, W, H, x, y).astype(np.float64)
        assert abs(np.linalg.norm(d) - 1) < 1e-6
        q = VP @ np.append(eye + 1.5 * d, 1.0)
        ndc = q[:2] / q[3]
        px = (ndc + 1) * 0.5 * np.array([W, H])
        assert np.allclose(px, [x + 0.5, y + 0.5], atol=2e-3)


def test_front_faces_and_keys_numpy():
    w = meshgen.grid_workload(4, seed=1)
    eye = (0.9, 0.6, 1.6)
    uf, _ = oracle.camera_uniforms(view_from(eye), FOVY, ZNEAR, ZFAR, 64, 64)
    fr = oracle.Fragments(w["cells"], w["pos"], uf, ZNEAR, 64, 64)
    keys, ids = fr.triangles()
    # numpy float32 projection of every tet, face visibility by the camera's side of each face plane
    FT = np.array([[0, 1, 2, 3], [1, 0, 3, 2], [0, 2, 3, 1], [2, 1, 3, 0]])
    p = w["pos"][w["cells"]].astype(np.float64)
    VP = uf[0:16].astype(np.float64).reshape(4, 4).T
    clip = np.einsum("ij,tvj->tvi", VP, np.concatenate([p, np.ones(p.shape[:2] + (1,))], axis=2))
    ndcz = clip[..., 2] / clip[..., 3]
    exp = {}
    for t in range(p.shape[0]):
        nvis = 0
        for f in range(4):
            a, b, c, o = (p[t, FT[f, k]] for k in range(4))
            n = np.cross(b - a, c - a)
            if np.sign(np.dot(np.asarray(eye) - a, n)) != np.sign(np.dot(o - a, n)):
                exp[t * 4 + f] = ndcz[t, FT[f, :3]].mean()
                nvis += 1
        assert 1 <= nvis <= 3
    got = dict(zip(ids.tolist(), keys.view(np.float32).astype(np.float64).tolist()))
    assert set(got) == set(exp)
    for k in exp:
        assert abs(got[k] - exp[k]) < 1e-6


def test_sort_is_stable_ascending():
    w = meshgen.grid_workload(6, seed=2)
    uf, _ = oracle.camera_uniforms(view_from((0.0, 0.0, 2.0)), FOVY, ZNEAR, ZFAR, 64, 64)
    fr = oracle.Fragments(w["cells"], w["pos"], uf, ZNEAR, 64, 64)
    keys, ids = fr.triangles()
    order = np.lexsort((ids, keys))
    assert np.array_equal(order, np.arange(keys.size))
    assert (np.diff(keys.astype(np.int64)) >= 0).all()


def _cube_chord(eye, d):
    """Length of the ray eye + t d inside [-0.5, 0.5]^3 (slab test, float64)."""
    with np.errstate(divide="ignore"):
        t1 = (-0.5 - eye) / d
        t2 = (0.5 - eye) / d
    tmin = np.max(np.minimum(t1, t2))
    tmax = np.min(np.maximum(t1, t2))
    return max(0.0, tmax - tmin), tmin, tmax


@pytest.mark.parametrize("linear_depth", [False, True])
@pytest.mark.parametrize("eye", [(0.05, 0.03, 2.0), (0.8, 0.5, 1.6), (-1.1, 0.3, -1.2)])
def test_constant_field_closed_form_and_watertight(eye, linear_depth):
    """Constant colour c and density s: pixel = c (1 - exp(-s * chord)) exactly (accumulateLinearConst with
    c0 = c1 composes through 'over' to a single exponential).  Missing or doubled fragments would change the
    chord, so this pins coverage, ray-tet intersection and compositing together.  With the PPLL renderer's
    linear-depth segment lengths the chord is measured along the view axis: chord * (d . axis), the axis being
    the direction from the eye to the origin the camera looks at.  Generic cameras only: a ray lying exactly in
    a face plane (e.g. an on-axis eye and the image diagonal) is degenerate for the reference's ray-tet test."""
    w = meshgen.grid_workload(5, seed=3)
    col = np.tile(np.float32([0.3, 0.6, 0.9, 2.5]), (w["pos"].shape[0], 1))
    W, H = 50, 46
    uf, ud = oracle.camera_uniforms(view_from(eye), FOVY, ZNEAR, ZFAR, W, H)
    fr = oracle.Fragments(w["cells"], w["pos"], uf, ZNEAR, W, H)
    img, _, cnt = oracle.forward(fr, w["pos"].astype(np.float64), col.astype(np.float64), ud, 1.0, dtype=np.float64,
                                 linear_depth=linear_depth)
    e = np.asarray(eye, np.float64)
    axis = -e / np.linalg.norm(e)
    checked = 0
    errs = []
    for y in range(H):
        for x in range(W):
            d = oracle.ray_dir(uf, W, H, x, y).astype(np.float64)
            chord, tmin, tmax = _cube_chord(e, d)
            # skip the one-pixel band around the silhouette, where the fixed-point raster decides
            near_edge = False
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    if 0 <= x + dx < W and 0 <= y + dy < H:
                        c2, _, _ = _cube_chord(e, oracle.ray_dir(uf, W, H, x + dx, y + dy).astype(np.float64))
                        near_edge |= (c2 > 0) != (chord > 0)
            if near_edge:
                continue
            if linear_depth:
                chord *= float(d @ axis) / float(np.linalg.norm(d))
            a = 1 - math.exp(-2.5 * chord)
            errs.append(max(abs(img[y, x, 3] - a), float(np.abs(img[y, x, :3] - np.array([0.3, 0.6, 0.9]) * a).max())))
            checked += 1
    errs = np.array(errs)
    assert checked > W * H // 4
    # Exact except where the 8-bit sub-pixel raster misses (or adds) a sliver of a tet far thinner than a pixel;
    # a missing or doubled tet would be an error of ~0.05.
    assert (errs < 2e-6).mean() >= 0.95, (errs >= 2e-6).sum()
    assert errs.max() < 5e-4, errs.max()


def test_accumulate_linear_const_against_quadrature():
    """accumulateLinearConst (ForwardCommon.glsl:42-49) equals the emission integral of the colour ramp
    c(s) = c0 + (c1 - c0) * s (slope per unit ray length, s in [0, t]) under constant extinction a."""
    from scipy.integrate import quad
    rng = np.random.default_rng(0)
    for _ in range(20):
        t = rng.uniform(0.01, 0.3)
        a = rng.uniform(0.1, 30.0)
        c0, c1 = rng.uniform(0, 1, 3), rng.uniform(0, 1, 3)
        out = oracle.accumulate_linear_const(t, c0, c1, a)
        for ch in range(3):
            ref = quad(lambda s: (c0[ch] + (c1[ch] - c0[ch]) * s) * a * math.exp(-a * s), 0, t)[0]
            assert abs(out[ch] - ref) < 1e-9
        assert abs(out[3] - (1 - math.exp(-a * t))) < 1e-12
    assert np.all(oracle.accumulate_linear_const(0.1, [1, 1, 1], [0, 0, 0], 1e-7) == 0)  # a < 1e-6 branch


def test_ray_tet_brute_force():
    """Entry/exit distances against an exact plane-clipping computation in Fraction arithmetic."""
    from fractions import Fraction as Fr
    rng = np.random.default_rng(5)
    for _ in range(30):
        v = rng.uniform(-1, 1, (4, 3))
        if abs(np.linalg.det(v[1:] - v[0])) < 0.05:
            continue
        target = v.mean(axis=0) + rng.uniform(-0.05, 0.05, 3)
        ro = rng.uniform(-1, 1, 3)
        ro = ro / np.linalg.norm(ro) * 4.0
        rd = (target - ro) / np.linalg.norm(target - ro)
        hit, f0, f1, t0, t1 = oracle.ray_tet(ro, rd, v)
        # exact: intersect the ray with the 4 half-spaces
        V = [[Fr(float(c)) for c in p] for p in v]
        O = [Fr(float(c)) for c in ro]
        D = [Fr(float(c)) for c in rd]
        lo, hi = Fr(-10 ** 9), Fr(10 ** 9)
        cen = [sum(V[i][k] for i in range(4)) / 4 for k in range(3)]
        for f in ([0, 1, 2], [1, 0, 3], [0, 2, 3], [2, 1, 3]):
            a, b, c = (V[i] for i in f)
            e1 = [b[k] - a[k] for k in range(3)]
            e2 = [c[k] - a[k] for k in range(3)]
            n = [e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]]
            if sum(n[k] * (cen[k] - a[k]) for k in range(3)) > 0:
                n = [-x for x in n]
            nd = sum(n[k] * D[k] for k in range(3))
            no = sum(n[k] * (O[k] - a[k]) for k in range(3))
            t = -no / nd
            if nd < 0:
                lo = max(lo, t)
            else:
                hi = min(hi, t)
        assert hit
        assert abs(t0 - float(lo)) < 1e-9 and abs(t1 - float(hi)) < 1e-9


def _fd_case():
    w = meshgen.grid_workload(3, seed=11, jitter=0.25)
    col = w["col"].copy()
    col[:, 3] = 1.0 + 3.0 * col[:, 3] / 20.0
    return w, col


@pytest.mark.parametrize("linear_depth", [False, True])
def test_exact_adjoint_matches_fp64_central_differences(linear_depth):
    """FP64 central differences of sum(w * image) on seeded vertices, restricted to pixels whose fragment lists
    are identical at p - h, p and p + h (visibility changes are not differentiable).  h = 1e-6: at 1e-4 a step
    often crosses a switch of a ray's entry or exit face inside the same fragment list (a kink), and the
    difference quotient then averages two one-sided slopes.  Both segment lengths: Euclidean and linear depth."""
    w, col = _fd_case()
    pos = w["pos"].astype(np.float64)
    colors = col.astype(np.float64)
    W = H = 40
    eye = (0.6, 0.45, 1.7)
    uf, ud = oracle.camera_uniforms(view_from(eye), FOVY, ZNEAR, ZFAR, W, H)
    rng = np.random.default_rng(3)
    wimg = rng.normal(size=(H, W, 4))
    interior = np.nonzero(~meshgen.boundary_vertex_mask(w["cells"], pos.shape[0]))[0]
    verts = rng.choice(pos.shape[0], 10, replace=False).tolist() + interior[:4].tolist()
    h = 1e-6
    fr0 = oracle.Fragments(w["cells"], pos.astype(np.float32), uf, ZNEAR, W, H)
    off0, ids0 = fr0.lists()

    def lists_equal_mask(fr):
        off, ids = fr.lists()
        m = np.zeros(W * H, bool)
        for p in range(W * H):
            m[p] = np.array_equal(ids[off[p]:off[p + 1]], ids0[off0[p]:off0[p + 1]])
        return m

    errs_p, errs_c = [], []
    for v in verts:
        for comp in range(3):
            mask = np.ones(W * H, bool)
            vals = []
            for s in (+1, -1):
                pp = pos.copy()
                pp[v, comp] += s * h
                frs = oracle.Fragments(w["cells"], pp.astype(np.float32), uf, ZNEAR, W, H)
                mask &= lists_equal_mask(frs)
                vals.append((frs, pp))
            mw = wimg * mask.reshape(H, W, 1)
            fplus = oracle.forward(vals[0][0], vals[0][1], colors, ud, 1.0, dtype=np.float64, linear_depth=linear_depth)[0]
            fminus = oracle.forward(vals[1][0], vals[1][1], colors, ud, 1.0, dtype=np.float64, linear_depth=linear_depth)[0]
            fd = float(np.sum(mw * (fplus - fminus)) / (2 * h))
            gp, _ = oracle.adjoint_exact(fr0, pos, colors, ud, 1.0, mw, dtype=np.float64, linear_depth=linear_depth)
            errs_p.append((fd, gp[v, comp]))
        for comp in range(4):
            cc = colors.copy()
            cc[v, comp] += h
            fp = oracle.forward(fr0, pos, cc, ud, 1.0, dtype=np.float64, linear_depth=linear_depth)[0]
            cc[v, comp] -= 2 * h
            fm = oracle.forward(fr0, pos, cc, ud, 1.0, dtype=np.float64, linear_depth=linear_depth)[0]
            fd = float(np.sum(wimg * (fp - fm)) / (2 * h))
            _, gc = oracle.adjoint_exact(fr0, pos, colors, ud, 1.0, wimg, dtype=np.float64, linear_depth=linear_depth)
            errs_c.append((fd, gc[v, comp]))
    for name, errs in (("pos", errs_p), ("col", errs_c)):
        fd = np.array([e[0] for e in errs])
        an = np.array([e[1] for e in errs])
        assert np.linalg.norm(fd) > 1e-3, name
        rel = np.linalg.norm(fd - an) / np.linalg.norm(fd)
        assert rel < 1e-5, (name, rel)


def test_exact_adjoint_float_matches_double():
    w, col = _fd_case()
    W = H = 40
    uf, ud = oracle.camera_uniforms(view_from((0.2, 0.7, 1.8)), FOVY, ZNEAR, ZFAR, W, H)
    fr = oracle.Fragments(w["cells"], w["pos"], uf, ZNEAR, W, H)
    dimg = np.random.default_rng(1).normal(size=(H, W, 4))
    gp32, gc32 = oracle.adjoint_exact(fr, w["pos"], col, uf, 1.0, dimg.astype(np.float32))
    gp64, gc64 = oracle.adjoint_exact(fr, w["pos"].astype(np.float64), col.astype(np.float64), ud, 1.0, dimg,
                                      dtype=np.float64)
    assert np.linalg.norm(gp32 - gp64) / np.linalg.norm(gp64) < 1e-4
    assert np.linalg.norm(gc32 - gc64) / np.linalg.norm(gc64) < 1e-5


def test_early_ray_termination_semantics():
    """ERT stops before the first fragment reached with alpha > threshold (PPLL/Forward.glsl:168-173)."""
    w = meshgen.grid_workload(6, seed=4)
    col = w["col"].copy()
    col[:, 3] = 60.0
    W = H = 32
    uf, _ = oracle.camera_uniforms(view_from((0.0, 0.1, 2.0)), FOVY, ZNEAR, ZFAR, W, H)
    fr = oracle.Fragments(w["cells"], w["pos"], uf, ZNEAR, W, H)
    img_full, C_full, n_full = oracle.forward(fr, w["pos"], col, uf, 1.0)
    img_ert, C_ert, n_ert = oracle.forward(fr, w["pos"], col, uf, 1.0, use_ert=True, ert_alpha=0.99)
    assert (n_ert <= n_full).all() and (n_ert < n_full).any()
    stopped = n_ert < n_full
    assert (C_ert[stopped, 3] > 0.99).all()
    assert np.abs(img_full - img_ert).max() <= 0.01 + 1e-6


def test_adam_matches_torch_golden():
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "adam_golden.npz"))
    for case in ("pos", "col"):
        p = g[f"{case}_p0"].copy()
        m = np.zeros_like(p)
        v = np.zeros_like(p)
        lr = float(g[f"{case}_lr"])
        clamp = bool(g[f"{case}_clamp"])
        for step in range(g[f"{case}_grads"].shape[0]):
            oracle.adam(p, g[f"{case}_grads"][step], m, v, lr, step + 1, clamp_nonneg=clamp)
            ref = g[f"{case}_p"][step]
            assert np.allclose(p, ref, rtol=2e-6, atol=1e-7), (case, step, np.abs(p - ref).max())


def test_mse_grad_matches_torch():
    import torch
    rng = np.random.default_rng(0)
    a = rng.normal(size=(5, 7, 4)).astype(np.float32)
    b = rng.normal(size=(5, 7, 4)).astype(np.float32)
    loss, d = oracle.mse_grad(a, b)
    ta = torch.from_numpy(a).requires_grad_(True)
    tl = torch.nn.MSELoss()(ta, torch.from_numpy(b))
    tl.backward()
    assert abs(loss - tl.item()) < 1e-6
    assert np.allclose(d, ta.grad.numpy(), atol=1e-7)
