"""Projection renderer (projected tetrahedra): pinning of the oracle's restatement, then the GPU against it.

Oracle (CPU):
  * single tets, against independent float64 geometry: the projected corners of the 3 or 4 triangles tile the
    convex hull of the tet's projection; in class 2 the extra vertex is the crossing of the two edges'
    projections and its thickness the distance between the two edge points seen there; in class 1 the thick
    vertex's thickness is the distance from it to the opposite face along its eye ray; the colours there are
    the reference's screen-space averages (GenerateTrianglesVTK.glsl:280-357);
  * a constant colour and density seen almost orthographically: thicknesses interpolated linearly over the
    triangles sum to the chord through the cube, so every pixel is c (1 - exp(-s * chord)) (coverage, order and
    compositing together; a missing or doubled tet would show).
GPU (marked gpu): fragment counts bit-identical and images within 1e-5 of the FP32 oracle (per-vertex and
per-cell colours), the adjoint against FP64 central differences of the oracle's forward with its fragment
lists and decisions frozen, batches of items against their single renders, and at the benchmark's size
(1.57M tets, 1024^2) the colour, alpha and position gradients through three scaling identities.
"""
import math

import numpy as np
import pytest

import oracle
from difftetvr_amd import meshgen
from tests.helpers import FOVY, ZFAR, ZNEAR, rel_l2, view_from


def _proj(ud, p):
    vp = ud[:16].reshape(4, 4).T
    c = vp @ np.append(p, 1.0)
    return c[:3] / c[3]


def _cross2(a, b):
    return a[0] * b[1] - a[1] * b[0]


def _seg_cross(p1, p2, p3, p4):
    """Parameters (s, t) of the crossing of 2D lines p1 + s (p2 - p1) and p3 + t (p4 - p3)."""
    A, B, C = p2 - p1, p4 - p3, p3 - p1
    den = _cross2(A, B)
    return _cross2(C, B) / den, _cross2(C, A) / den


def _eye_ray_point(ud, ndc_xy, z=0.5):
    """World point on the eye ray through NDC (x, y) (unprojection of depth z)."""
    ip = ud[32:48].reshape(4, 4).T
    iv = ud[48:64].reshape(4, 4).T
    e = ip @ np.array([ndc_xy[0], ndc_xy[1], z, 1.0])
    return (iv @ np.append(e[:3] / e[3], 1.0))[:3]


def _eye_dist(ud, p, q):
    """Distance between world points p and q measured in eye space (the float-rounded view matrix is orthonormal
    only to float precision; the reference measures thickness in eye space)."""
    v = ud[16:32].reshape(4, 4).T
    return float(np.linalg.norm(v[:3, :3] @ (p - q)))


def _line_hit(o, d, a, b):
    """Point on segment line a + u (b - a) closest to the ray o + t d (they meet for a projected crossing)."""
    w = b - a
    M = np.array([[d @ d, -(d @ w)], [d @ w, -(w @ w)]])
    rhs = np.array([d @ (a - o), w @ (a - o)])
    t, u = np.linalg.solve(M, rhs)
    return a + u * w


def test_projected_tetrahedra_single_tet_closed_forms():
    rng = np.random.default_rng(7)
    W = H = 64
    eye = np.array([0.3, 0.4, 2.2])
    uf, ud = oracle.camera_uniforms(view_from(tuple(eye)), FOVY, ZNEAR, ZFAR, W, H)
    cam = ud[64:67]
    seen = {True: 0, False: 0}
    for _ in range(200):
        P = rng.uniform(-0.3, 0.3, (4, 3))
        if abs(np.linalg.det(np.stack([P[1] - P[0], P[2] - P[0], P[3] - P[0]]))) < 1e-3:
            continue
        C = rng.uniform(0, 1, (4, 4))
        r = oracle.proj_tet(P, C, ud, uf, ZNEAR, 2.0)
        assert r["ntri"] in (3, 4)
        ndc = np.array([_proj(ud, p) for p in P])
        assert np.allclose(r["ndc"][:4], ndc, atol=1e-12)
        # the triangles tile the convex hull of the projected vertices
        pts = np.vstack([ndc[:, :2], r["ndc"][4:5, :2]])
        tri_area = sum(abs(_cross2(pts[c[1]] - pts[c[0]], pts[c[2]] - pts[c[0]])) / 2
                       for c in r["corners"][:r["ntri"]])
        from scipy.spatial import ConvexHull
        assert abs(tri_area - ConvexHull(ndc[:, :2]).volume) < 1e-9 * max(1.0, tri_area)
        seen[r["cls2"]] += 1
        if r["cls2"]:
            # the crossing of the two edges' projections, and the eye-space distance between the edge points
            found = False
            for (a0, a1), (b0, b1) in (((0, 1), (2, 3)), ((0, 2), (1, 3)), ((0, 3), (1, 2))):
                s, t = _seg_cross(ndc[a0, :2], ndc[a1, :2], ndc[b0, :2], ndc[b1, :2])
                if 0 <= s <= 1 and 0 <= t <= 1:
                    x = ndc[a0, :2] + s * (ndc[a1, :2] - ndc[a0, :2])
                    assert np.allclose(r["ndc"][4, :2], x, atol=1e-9)
                    d = _eye_ray_point(ud, x) - cam
                    q1, q2 = _line_hit(cam, d, P[a0], P[a1]), _line_hit(cam, d, P[b0], P[b1])
                    assert abs(r["dep"][4] - 2.0 * _eye_dist(ud, q1, q2)) < 1e-9
                    col = 0.5 * (C[a0] + s * (C[a1] - C[a0]) + C[b0] + t * (C[b1] - C[b0]))
                    assert np.allclose(r["col"][4], col, atol=1e-9)
                    found = True
            assert found
        else:
            # the thick vertex projects inside the opposite triangle; its thickness: along its eye ray to the
            # opposite face's plane
            thick = [v for v in range(4) if r["dep"][v] != 0]
            assert len(thick) == 1
            v = thick[0]
            o = [i for i in range(4) if i != v]
            n = np.cross(P[o[1]] - P[o[0]], P[o[2]] - P[o[0]])
            d = P[v] - cam
            hit = cam + d * ((P[o[0]] - cam) @ n) / (d @ n)
            assert abs(r["dep"][v] - 2.0 * _eye_dist(ud, hit, P[v])) < 1e-9 * max(1.0, r["dep"][v])
            # colour: the average of the vertex's and the opposite face's (screen-space barycentrics)
            a, b, c = ndc[o[0], :2], ndc[o[1], :2], ndc[o[2], :2]
            lam = np.linalg.solve(np.array([[a[0], b[0], c[0]], [a[1], b[1], c[1]], [1, 1, 1]]),
                                  np.append(ndc[v, :2], 1.0))
            facec = lam[0] * C[o[0]] + lam[1] * C[o[1]] + lam[2] * C[o[2]]
            assert np.allclose(r["col"][v], 0.5 * (facec + C[v]), atol=1e-9)
    assert seen[True] > 20 and seen[False] > 20


def test_projected_tetrahedra_constant_field_near_orthographic():
    """Seen from afar through a narrow field of view, the interpolated thicknesses sum to the chord."""
    w = meshgen.grid_workload(4, seed=3)
    col = np.tile(np.float32([0.3, 0.6, 0.9, 2.5]), (w["pos"].shape[0], 1))
    W, H = 48, 40
    dist = 400.0
    d = np.array([0.35, 0.5, 1.0])
    eye = tuple(dist * d / np.linalg.norm(d))
    fovy = 2.0 * math.atan(0.9 / dist)
    uf, ud = oracle.camera_uniforms(view_from(eye), fovy, dist - 2.0, dist + 2.0, W, H)
    pf = oracle.ProjFragments(w["cells"], w["pos"], col, uf, dist - 2.0, W, H)
    img, _ = oracle.proj_forward(pf, w["pos"].astype(np.float64), col.astype(np.float64), ud, 1.0, dtype=np.float64)
    e = np.asarray(eye)
    errs = []
    for y in range(H):
        for x in range(W):
            rd = oracle.ray_dir(uf, W, H, x, y).astype(np.float64)
            with np.errstate(divide="ignore"):
                t1, t2 = (-0.5 - e) / rd, (0.5 - e) / rd
            chord = max(0.0, np.min(np.maximum(t1, t2)) - np.max(np.minimum(t1, t2)))
            if chord < 0.05:  # away from the silhouette
                continue
            a = 1 - math.exp(-2.5 * chord)
            errs.append(max(abs(img[y, x, 3] - a), float(np.abs(img[y, x, :3] - np.array([0.3, 0.6, 0.9]) * a).max())))
    errs = np.array(errs)
    assert errs.size > W * H // 4
    # Coverage snaps vertices to 1/256 pixel while the barycentrics use the unsnapped positions, so near a
    # silhouette edge a covered pixel can lie just outside its float triangle (thickness slope x 1/512 pixel,
    # ~1e-4 here); a missing or doubled tet would be off by ~0.1.
    assert np.median(errs) < 2e-5, np.median(errs)
    assert np.percentile(errs, 99) < 2e-4 and errs.max() < 5e-4, (np.percentile(errs, 99), errs.max())


# ---------------------------------------------------------------------------------------------------------- GPU

def _renderer(d, cells, pos, col=None, cell_col=None, W=192, H=176, eye=(0.9, 0.6, 1.5), att=1.0):
    mesh = d.TetMesh()
    mesh.set_use_gradients(True)
    mesh.set_tet_mesh_data(cells, pos, col, cell_col)
    r = d.Renderer(renderer_type=d.RendererType.PROJECTION)
    r.set_tet_mesh(mesh)
    r.set_attenuation(att)
    r.set_clear_color(d.vec4(0, 0, 0, 0))
    r.set_viewport_size(W, H)
    r.set_camera_fovy(FOVY)
    r.set_view_matrix(view_from(eye))
    return mesh, r


@pytest.mark.gpu
@pytest.mark.parametrize("per_cell", [False, True])
def test_gpu_projection_forward_parity(gpu, per_cell):
    import difftetvr_amd as d
    W, H, eye = 192, 176, (0.9, 0.6, 1.5)
    w = meshgen.grid_workload(16, seed=0)
    ccol = None
    if per_cell:
        rng = np.random.default_rng(3)
        T = w["cells"].shape[0]
        ccol = np.concatenate([rng.uniform(0, 1, (T, 3)), rng.uniform(0, 8, (T, 1))], axis=1).astype(np.float32)
    mesh, r = _renderer(d, w["cells"], w["pos"], None if per_cell else w["col"], ccol, W, H, eye)
    img = r.render().cpu().numpy()
    cnt = r.get_fragment_counts().cpu().numpy()
    uf, _ = oracle.camera_uniforms(view_from(eye), FOVY, ZNEAR, ZFAR, W, H)
    pf = oracle.ProjFragments(w["cells"], w["pos"], None if per_cell else w["col"], uf, ZNEAR, W, H,
                              cell_colors=ccol)
    ref, _ = oracle.proj_forward(pf, w["pos"], None if per_cell else w["col"], uf, 1.0, cell_colors=ccol)
    err = float(np.abs(img - ref).max())
    print(f"projection grid16 {'per-cell' if per_cell else 'per-vertex'} {W}x{H}: fragments {int(cnt.sum())}, "
          f"max|img diff| {err:.3e}, alpha max {ref[..., 3].max():.3f}", flush=True)
    assert np.array_equal(cnt, pf.counts())
    assert ref[..., 3].max() > 0.5
    assert err <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("per_cell", [False, True])
def test_gpu_projection_adjoint_central_differences(gpu, per_cell):
    """GPU adjoint of sum(w * image) against FP64 central differences of the oracle's forward (lists, order and
    triangle decisions frozen at the rendered mesh: the function the adjoint differentiates)."""
    import torch
    import difftetvr_amd as d
    W = H = 64
    eye = (0.7, 0.5, 1.6)
    w = meshgen.grid_workload(4, seed=2)
    pos = w["pos"]
    col = w["col"].copy()
    col[:, 3] = 1.0 + 3.0 * col[:, 3] / 20.0
    rng = np.random.default_rng(5)
    T = w["cells"].shape[0]
    ccol = np.concatenate([rng.uniform(0, 1, (T, 3)), rng.uniform(0.5, 4, (T, 1))], axis=1).astype(np.float32)
    vcol = None if per_cell else col
    cc = ccol if per_cell else None
    mesh, r = _renderer(d, w["cells"], pos, vcol, cc, W, H, eye)
    r.render()
    wimg = rng.normal(size=(H, W, 4)).astype(np.float32)
    r.render_adjoint(torch.from_numpy(wimg).cuda())
    torch.cuda.synchronize()
    gp = mesh.get_vertex_positions().grad.cpu().numpy().astype(np.float64)
    gc = (mesh.get_cell_colors() if per_cell else mesh.get_vertex_colors()).grad.cpu().numpy().astype(np.float64)
    uf, ud = oracle.camera_uniforms(view_from(eye), FOVY, ZNEAR, ZFAR, W, H)
    pf = oracle.ProjFragments(w["cells"], pos, vcol, uf, ZNEAR, W, H, cell_colors=cc)
    p64 = pos.astype(np.float64)
    c64 = (ccol if per_cell else col).astype(np.float64)

    def loss(p, c):
        img, _ = oracle.proj_forward(pf, p, None if per_cell else c, ud, 1.0, dtype=np.float64,
                                     cell_colors=c if per_cell else None)
        return float(np.sum(wimg.astype(np.float64) * img))

    h = 1e-6
    fd_p, an_p, fd_c, an_c = [], [], [], []
    interior = np.nonzero(~meshgen.boundary_vertex_mask(w["cells"], pos.shape[0]))[0]
    for v in list(rng.choice(pos.shape[0], 10, replace=False)) + list(interior[:6]):
        for k in range(3):
            pp, pm = p64.copy(), p64.copy()
            pp[v, k] += h
            pm[v, k] -= h
            fd_p.append((loss(pp, c64) - loss(pm, c64)) / (2 * h))
            an_p.append(gp[v, k])
    for t in rng.choice(c64.shape[0], 12, replace=False):
        for k in range(4):
            cp, cm = c64.copy(), c64.copy()
            cp[t, k] += h
            cm[t, k] -= h
            fd_c.append((loss(p64, cp) - loss(p64, cm)) / (2 * h))
            an_c.append(gc[t, k])
    ep, ec = rel_l2(an_p, fd_p), rel_l2(an_c, fd_c)
    print(f"projection adjoint ({'per-cell' if per_cell else 'per-vertex'}) vs FP64 differences: positions "
          f"{ep:.2e}, colours {ec:.2e}", flush=True)
    assert np.linalg.norm(fd_p) > 1e-3 and np.linalg.norm(fd_c) > 1e-3
    assert ec <= 1e-4
    assert ep <= 1e-3
    # every triangle construction (class 2; class 1 with the thick vertex at either end of segment 1), each
    # through a directional derivative over all positions and colours of the vertices of that class's tets
    groups = {"class 2": [], "class 1": [], "class 1 flipped": []}
    for t in range(T):
        vs = w["cells"][t]
        cl = ccol[t][None].repeat(4, 0) if per_cell else col[vs]
        pt = oracle.proj_tet(pos[vs], cl, ud, uf, ZNEAR, 1.0, per_cell=per_cell)
        if pt["ntri"] == 0:
            continue
        # class 1: corner 0 of the first triangle is the thick vertex, tet vertex 0 when segment 1 is reversed
        name = "class 2" if pt["cls2"] else ("class 1 flipped" if pt["corners"][0, 0] == 0 else "class 1")
        groups[name].append(t)
    for name, tets in groups.items():
        assert len(tets) >= 5, f"{name}: only {len(tets)} tets in view"
        vmask = np.zeros(pos.shape[0], bool)
        vmask[w["cells"][tets].reshape(-1)] = True
        dp = rng.normal(size=p64.shape) * vmask[:, None]
        if per_cell:
            dc = np.zeros_like(c64)
            dc[tets] = rng.normal(size=(len(tets), 4))
        else:
            dc = rng.normal(size=c64.shape) * vmask[:, None]
        fd = (loss(p64 + h * dp, c64 + h * dc) - loss(p64 - h * dp, c64 - h * dc)) / (2 * h)
        an = float(np.sum(gp * dp) + np.sum(gc * dc))
        print(f"  {name} ({len(tets)} tets): directional {an:.6e} vs differences {fd:.6e}", flush=True)
        assert abs(an - fd) <= 1e-3 * abs(fd) + 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("per_cell", [False, True])
def test_gpu_projection_batch(gpu, per_cell):
    """A batch of Projection items (two meshes, one seen twice) gives each item's single render bit for bit, and
    its adjoint the sum of the single adjoints (FP32 atomics reorder the sums: relative L2 <= 1e-5)."""
    import torch
    import difftetvr_amd as d
    rng = np.random.default_rng(11)
    wa, wb = meshgen.grid_workload(6, seed=1), meshgen.grid_workload(5, seed=2)

    def mesh_of(w):
        m = d.TetMesh()
        m.set_use_gradients(True)
        if per_cell:
            T = w["cells"].shape[0]
            cc = np.concatenate([rng.uniform(0, 1, (T, 3)), rng.uniform(0.5, 4, (T, 1))], axis=1).astype(np.float32)
            m.set_tet_mesh_data(w["cells"], w["pos"], None, cc)
        else:
            m.set_tet_mesh_data(w["cells"], w["pos"], w["col"])
        return m

    ma, mb = mesh_of(wa), mesh_of(wb)
    eyes = [(0.9, 0.6, 1.5), (-1.2, 0.5, 1.3), (0.4, -1.4, -1.1)]
    W = H = 48
    items = [(ma, view_from(eyes[0]), W, H), (mb, view_from(eyes[1]), W, H), (ma, view_from(eyes[2]), W, H)]
    br = d.BatchRenderer(renderer_type=d.RendererType.PROJECTION)
    br.set_attenuation(1.0)
    br.set_clear_color(d.vec4(0, 0, 0, 0))
    br.set_camera_fovy(FOVY)
    br.set_items(items)
    imgs = br.render()
    cnts = br.get_fragment_counts()
    dimgs = torch.randn((3, H, W, 4), device="cuda", generator=torch.Generator(device="cuda").manual_seed(2))
    br.render_adjoint(dimgs)
    torch.cuda.synchronize()
    # the same meshes' gradients from single renders, into copies
    sa = d.TetMesh()
    sb = d.TetMesh()
    for src, dst in ((ma, sa), (mb, sb)):
        dst.set_use_gradients(True)
        cells = src.get_cell_indices().cpu().numpy()
        pos = src.get_vertex_positions().detach().cpu().numpy()
        if per_cell:
            dst.set_tet_mesh_data(cells, pos, None, src.get_cell_colors().detach().cpu().numpy())
        else:
            dst.set_tet_mesh_data(cells, pos, src.get_vertex_colors().detach().cpu().numpy())
    for i, (m, vm, w_, h_) in enumerate(items):
        r = d.Renderer(renderer_type=d.RendererType.PROJECTION)
        r.set_tet_mesh(sa if m is ma else sb)
        r.set_attenuation(1.0)
        r.set_clear_color(d.vec4(0, 0, 0, 0))
        r.set_camera_fovy(FOVY)
        r.set_viewport_size(w_, h_)
        r.set_view_matrix(vm)
        ref = r.render()
        refc = r.get_fragment_counts()
        r.render_adjoint(dimgs[i].contiguous())
        torch.cuda.synchronize()
        assert torch.equal(imgs[i], ref)
        assert torch.equal(cnts[i], refc)
        assert int(refc.sum()) > 0
    for m, s_ in ((ma, sa), (mb, sb)):
        gp, gp1 = m.get_vertex_positions().grad.cpu().numpy(), s_.get_vertex_positions().grad.cpu().numpy()
        col = (lambda x: x.get_cell_colors()) if per_cell else (lambda x: x.get_vertex_colors())
        gc, gc1 = col(m).grad.cpu().numpy(), col(s_).grad.cpu().numpy()
        print(f"projection batch ({'per-cell' if per_cell else 'per-vertex'}): positions {rel_l2(gp, gp1):.2e}, "
              f"colours {rel_l2(gc, gc1):.2e}", flush=True)
        assert np.linalg.norm(gp1) > 0 and np.linalg.norm(gc1) > 0
        assert rel_l2(gp, gp1) <= 1e-5 and rel_l2(gc, gc1) <= 1e-5


@pytest.mark.gpu
def test_gpu_projection_fullsize_scaling_identities(gpu):
    """The benchmark's size (64^3 x 6 tets, 1024 x 1024, its camera), through two identities the construction
    fixes: every corner colour (the crossing's and the thick vertex's too) is linear in the vertex colours and
    alpha depends on them only through c.a times the thickness, which is proportional to the attenuation. So
    for L = sum dI . I (clear colour 0) the RGB gradients satisfy sum_v c_rgb(v) . dL/dc_rgb(v) = sum dI_rgb .
    I_rgb (degree-one homogeneity), and the alpha gradients sum_v c_a(v) dL/dc_a(v) = att dL/d(att), taken by
    central differences of two more renders."""
    import torch
    import difftetvr_amd as d
    w = meshgen.grid_workload(64, seed=0)
    W = H = 1024
    eye = tuple(2.0 * np.array([0.3, 0.4, 1.0]) / np.linalg.norm([0.3, 0.4, 1.0]))
    mesh, r = _renderer(d, w["cells"], w["pos"], w["col"], None, W, H, eye, att=1.0)
    img = r.render().clone()
    dimg = torch.rand((H, W, 4), device="cuda", generator=torch.Generator(device="cuda").manual_seed(4)) + 0.5
    r.render_adjoint(dimg)
    torch.cuda.synchronize()
    g = mesh.get_vertex_colors().grad.cpu().numpy().astype(np.float64)
    c = w["col"].astype(np.float64)
    im, di = img.cpu().numpy().astype(np.float64), dimg.cpu().numpy().astype(np.float64)
    lhs, rhs = float(np.sum(c[:, :3] * g[:, :3])), float(np.sum(di[..., :3] * im[..., :3]))
    eps = 1e-3
    losses = []
    for att in (1.0 + eps, 1.0 - eps):
        r.set_attenuation(att)
        losses.append(float(torch.sum(dimg.double() * r.render().double())))
    fd = (losses[0] - losses[1]) / (2 * eps)
    lhs_a = float(np.sum(c[:, 3] * g[:, 3]))
    # scaling the mesh about the eye keeps every vertex's screen position and every triangle, and scales each
    # thickness (the distance between two points on one eye ray): sum_v (p_v - eye) . dL/dp_v = att dL/d(att)
    gp = mesh.get_vertex_positions().grad.cpu().numpy().astype(np.float64)
    tp = (w["pos"].astype(np.float64) - np.asarray(eye, np.float64)) * gp
    lhs_p = float(np.sum(tp))
    print(f"projection 1024^2: RGB {lhs:.9e} vs {rhs:.9e}; alpha {lhs_a:.6e}, positions {lhs_p:.6e} (terms "
          f"{float(np.sum(np.abs(tp))):.3e}) vs differences {fd:.6e}", flush=True)
    assert rhs > 1.0 and abs(lhs - rhs) <= 1e-5 * rhs
    assert abs(fd) > 1.0 and abs(lhs_a - fd) <= 1e-3 * abs(fd)
    assert abs(lhs_p - fd) <= 1e-3 * abs(fd) + 1e-5 * float(np.sum(np.abs(tp)))


@pytest.mark.gpu
def test_gpu_projection_refusals(gpu):
    import torch
    import difftetvr_amd as d
    w = meshgen.grid_workload(3, seed=1)
    mesh, r = _renderer(d, w["cells"], w["pos"], w["col"], None, 32, 32)
    r.render()
    with pytest.raises(RuntimeError, match="triangles"):
        r.get_fragment_ids(4)
    r.set_adjoint_mode("reference")
    with pytest.raises(RuntimeError, match="exact"):
        r.render_adjoint(torch.zeros(32, 32, 4, device="cuda"))
