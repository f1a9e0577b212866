"""Parity at the benchmark's full size: the 64^3 x 6 = 1,572,864-tet grid at 1024 x 1024, through properties the
mathematics fixes whatever the size.

With constant colour c and density s (attenuation 1, no early termination, clear colour 0):
  * every pixel is c * a, a = 1 - exp(-s * chord), chord = the ray's length inside the unit cube (the grid's
    boundary vertices are not jittered): checked against the exact slab-test chord on a lattice of pixels;
  * adding d to every vertex colour adds d * a to every pixel's RGB, so the colour gradient summed over all
    vertices equals sum over pixels of dL/dI_rgb * a;
  * adding d to every density multiplies each pixel's transmittance by exp(-d * chord), so the density gradient
    summed over all vertices equals sum over pixels of (dL/dI_a + dL/dI_rgb . c) * (1 - a) * chord, with
    chord = -ln(1 - a) / s read back from the image itself.
The gradient identities exercise every fragment of the exact adjoint (76M ray-tet segments) at once.  On the
benchmark's own field (seeded colours and densities): sampled rows' fragment lists, images and adjoint against
the oracle, and every pixel's colour and position gradients through RGB homogeneity, the density-attenuation
scaling and the scaling of the mesh about the eye."""
import math

import numpy as np
import pytest
import torch

import oracle
from difftetvr_amd import meshgen
from tests.helpers import FOVY, ZFAR, ZNEAR, adjoint_references, check_gradients, make_renderer, view_from

pytestmark = pytest.mark.gpu

W = H = 1024
EYE = tuple(2.0 * v / math.sqrt(0.3 ** 2 + 0.4 ** 2 + 1.0) for v in (0.3, 0.4, 1.0))
COLOR = np.array([0.7, 0.3, 0.2], np.float32)
SIGMA = 2.0


def _cube_chord(eye, d):
    with np.errstate(divide="ignore", invalid="ignore"):
        t1 = (-0.5 - eye) / d
        t2 = (0.5 - eye) / d
    tmin = np.max(np.minimum(t1, t2))
    tmax = np.min(np.maximum(t1, t2))
    return max(0.0, tmax - tmin)


@pytest.fixture(scope="module")
def fullsize(gpu):
    import difftetvr_amd as d
    w = meshgen.grid_workload(64, seed=0)
    col = np.tile(np.append(COLOR, np.float32(SIGMA)), (w["pos"].shape[0], 1)).astype(np.float32)
    mesh, r = make_renderer(d, w["cells"], w["pos"], col, W, H, eye=EYE, attenuation=1.0)
    img = r.render().clone()
    torch.cuda.synchronize()
    return mesh, r, img


def test_fullsize_constant_field_image(fullsize):
    """Closed form on a lattice of pixels; where a pixel departs from it (a ray through a tet edge that the
    reference's ray-tet test rejects, or a silhouette sliver the fixed-point raster decides), the oracle's
    render of that pixel's row shows the departure is the algorithm's, not the GPU's."""
    mesh, r, img = fullsize
    im = img.cpu().numpy()
    uf, _ = oracle.camera_uniforms(view_from(EYE), FOVY, ZNEAR, ZFAR, W, H)
    eye = np.asarray(EYE, np.float64)
    errs, where = [], []
    for y in range(3, H, 16):
        for x in range(5, W, 16):
            chord = _cube_chord(eye, oracle.ray_dir(uf, W, H, x, y).astype(np.float64))
            a = 1.0 - math.exp(-SIGMA * chord)
            errs.append(max(abs(im[y, x, 3] - a), float(np.abs(im[y, x, :3] - COLOR * a).max())))
            where.append((x, y))
    errs = np.array(errs)
    covered = float((im[..., 3] > 0).mean())
    print(f"full size: {errs.size} pixels checked, covered {covered:.3f}, within 2e-5: {(errs <= 2e-5).mean():.4f}, "
          f"max {errs.max():.2e}, fragments {int(r.get_fragment_counts().to(torch.int64).sum())}", flush=True)
    assert covered > 0.2
    assert (errs <= 2e-5).mean() >= 0.99
    # the worst departures: the GPU equals the oracle there (FP32, same row, same fragment lists)
    cells = mesh.get_cell_indices().cpu().numpy()
    pos = mesh.get_vertex_positions().detach().cpu().numpy()
    col = mesh.get_vertex_colors().detach().cpu().numpy()
    for k in np.argsort(errs)[::-1][:3]:
        x, y = where[k]
        fr = oracle.Fragments(cells, pos, uf, ZNEAR, W, H, y, y + 1, 1)
        oimg, _, _ = oracle.forward(fr, pos, col, uf, 1.0)
        d = float(np.abs(oimg[y, x] - im[y, x]).max())
        print(f"pixel ({x},{y}): closed-form error {errs[k]:.2e}, GPU vs oracle {d:.2e}", flush=True)
        assert d <= 1e-5
    assert errs.max() <= 1e-2      # a whole missing or doubled tet (edge ~1/64) would be ~0.03


def test_fullsize_gradient_identities(fullsize):
    mesh, r, img = fullsize
    gen = torch.Generator(device="cuda").manual_seed(5)
    dimg = torch.rand(img.shape, device="cuda", generator=gen) + 0.5     # positive: no cancellation in the sums
    mesh.get_vertex_colors().grad.zero_()
    mesh.get_vertex_positions().grad.zero_()
    r.set_timing(True)                     # also counts the adjoint's gradient atomics
    r.render_adjoint(dimg)
    r.set_timing(False)
    st = r.get_stats()
    # at most one atomic per (distinct vertex of a 32-record batch, value): 28 per record
    assert 0 < st["num_gradient_atomics"] <= 28 * st["num_pairs"]
    gcol = mesh.get_vertex_colors().grad.to(torch.float64).sum(dim=0).cpu().numpy()
    a = img[..., 3].to(torch.float64)
    di = dimg.to(torch.float64)
    expect_rgb = (di[..., :3] * a[..., None]).sum(dim=(0, 1)).cpu().numpy()
    one_minus = (1.0 - a).clamp_min(1e-30)
    chord = -torch.log(one_minus) / SIGMA
    c = torch.tensor(COLOR, dtype=torch.float64, device="cuda")
    expect_s = float(((di[..., 3] + (di[..., :3] * c).sum(dim=-1)) * one_minus * chord).sum())
    rel_rgb = np.abs(gcol[:3] - expect_rgb) / np.abs(expect_rgb)
    rel_s = abs(gcol[3] - expect_s) / abs(expect_s)
    print(f"colour-gradient sums {gcol[:3]} vs {expect_rgb} (rel {rel_rgb}); density {gcol[3]:.6e} vs "
          f"{expect_s:.6e} (rel {rel_s:.2e})", flush=True)
    assert (rel_rgb <= 1e-4).all()
    assert rel_s <= 1e-3


# ---- the benchmark's own field: seeded blob colours, density up to 20, jittered interior vertices ----------
ROWS = tuple(range(37, H, 128))  # 8 rows across the image


@pytest.fixture(scope="module")
def fullsize_blob(gpu):
    import difftetvr_amd as d
    w = meshgen.grid_workload(64, seed=0)
    mesh, r = make_renderer(d, w["cells"], w["pos"], w["col"], W, H, eye=EYE, attenuation=1.0)
    img = r.render().clone()
    torch.cuda.synchronize()
    uf, _ = oracle.camera_uniforms(view_from(EYE), FOVY, ZNEAR, ZFAR, W, H)
    fr = oracle.Fragments(w["cells"], w["pos"], uf, ZNEAR, W, H, ROWS[0], ROWS[-1] + 1, ROWS[1] - ROWS[0])
    return w, mesh, r, img, uf, fr


def test_fullsize_fragment_order_rows(fullsize_blob):
    """The 9M-record, multi-pass radix sort at full size: each sampled row's per-pixel ordered fragment lists
    equal the oracle's (a stable sort of the same depth keys) bit for bit."""
    w, mesh, r, img, uf, fr = fullsize_blob
    counts = r.get_fragment_counts()
    maxn = int(counts.max())
    ids = r.get_fragment_ids(maxn)
    off, oids = fr.lists()
    nfr = 0
    for y in ROWS:
        g = ids[y].cpu().numpy()
        o = np.full_like(g, -1)
        for x in range(W):
            p = y * W + x
            n = off[p + 1] - off[p]
            o[x, :n] = oids[off[p]:off[p + 1]]
            nfr += n
        bad = int((g != o).any(axis=1).sum())
        assert bad == 0, f"row {y}: {bad} pixels with a different fragment list"
    st = r.get_stats()
    print(f"full size: {len(ROWS)} rows, {nfr} fragments in the same order as the oracle; "
          f"{st['num_pairs']} records, {st['sort_passes']} sort passes", flush=True)
    assert nfr > 100_000


def test_fullsize_blob_field_rows(fullsize_blob):
    """Images of the sampled rows within 1e-5 of the oracle (FP32)."""
    w, mesh, r, img, uf, fr = fullsize_blob
    oimg, _, _ = oracle.forward(fr, w["pos"], w["col"], uf, 1.0)
    im = img.cpu().numpy()
    err = max(float(np.abs(im[y] - oimg[y]).max()) for y in ROWS)
    print(f"full size blob field: max |image - oracle| over {len(ROWS)} rows = {err:.2e}", flush=True)
    assert err <= 1e-5


def test_fullsize_adjoint_rows(fullsize_blob):
    """The exact adjoint at full size with the image adjoint nonzero on the sampled rows only: the gradients then
    sum exactly those rows' fragments, which the oracle's adjoint of the same rows gives (check_gradients)."""
    w, mesh, r, img, uf, fr = fullsize_blob
    rng = np.random.default_rng(17)
    dimg = np.zeros((H, W, 4), np.float32)
    for y in ROWS:
        dimg[y] = rng.normal(size=(W, 4))
    mesh.get_vertex_colors().grad.zero_()
    mesh.get_vertex_positions().grad.zero_()
    r.render_adjoint(torch.from_numpy(dimg).cuda(), False)
    torch.cuda.synchronize()
    gp = mesh.get_vertex_positions().grad.cpu().numpy()
    gc = mesh.get_vertex_colors().grad.cpu().numpy()
    _, ud = oracle.camera_uniforms(view_from(EYE), FOVY, ZNEAR, ZFAR, W, H)
    check_gradients(f"full size adjoint over {len(ROWS)} rows", gp, gc,
                    adjoint_references(oracle, fr, w["pos"], w["col"], uf, ud, 1.0, dimg))


def test_fullsize_blob_scaling_identities(fullsize_blob):
    """Every pixel's gradients at once, through two identities of the exact forward: the RGB of each sub-segment
    is linear in the interpolated vertex RGB and its opacity depends on the density only through density x
    length x attenuation. So for L = sum dI . I (clear colour 0, no early termination) sum_v c_rgb(v) .
    dL/dc_rgb(v) = sum dI_rgb . I_rgb, and sum_v c_a(v) dL/dc_a(v) = att dL/d(att), taken by central differences
    of two more renders."""
    w, mesh, r, img, uf, fr = fullsize_blob
    dimg = torch.rand((H, W, 4), device="cuda", generator=torch.Generator(device="cuda").manual_seed(5)) + 0.5
    mesh.get_vertex_colors().grad.zero_()
    mesh.get_vertex_positions().grad.zero_()
    r.render()
    r.render_adjoint(dimg, False)
    torch.cuda.synchronize()
    g = mesh.get_vertex_colors().grad.cpu().numpy().astype(np.float64)
    c = w["col"].astype(np.float64)
    im, di = img.cpu().numpy().astype(np.float64), dimg.cpu().numpy().astype(np.float64)
    lhs, rhs = float(np.sum(c[:, :3] * g[:, :3])), float(np.sum(di[..., :3] * im[..., :3]))
    eps = 1e-3
    losses = []
    try:
        for att in (1.0 + eps, 1.0 - eps):
            r.set_attenuation(att)
            losses.append(float(torch.sum(dimg.double() * r.render().double())))
    finally:
        r.set_attenuation(1.0)
    fd = (losses[0] - losses[1]) / (2 * eps)
    lhs_a = float(np.sum(c[:, 3] * g[:, 3]))
    # scaling the mesh about the eye keeps every ray's faces and barycentrics and scales every segment length:
    # sum_v (p_v - eye) . dL/dp_v = att dL/d(att) as well
    gp = mesh.get_vertex_positions().grad.cpu().numpy().astype(np.float64)
    tp = (w["pos"].astype(np.float64) - np.asarray(EYE, np.float64)) * gp
    lhs_p = float(np.sum(tp))
    print(f"full size blob field: RGB {lhs:.9e} vs {rhs:.9e}; alpha {lhs_a:.6e}, positions {lhs_p:.6e} (terms "
          f"{float(np.sum(np.abs(tp))):.3e}) vs differences {fd:.6e}", flush=True)
    assert rhs > 1.0 and abs(lhs - rhs) <= 1e-5 * rhs
    assert abs(fd) > 1.0 and abs(lhs_a - fd) <= 1e-3 * abs(fd)
    assert abs(lhs_p - fd) <= 1e-3 * abs(fd) + 1e-5 * float(np.sum(np.abs(tp)))
