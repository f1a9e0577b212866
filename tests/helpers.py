"""Shared workload builders for the tests (seeded, small)."""
import math

import numpy as np

from difftetvr_amd import meshgen

FOVY = math.pi / 4.0
ZNEAR, ZFAR = 0.01, 100.0


def grid_case(n=16, seed=0, jitter=0.2):
    return meshgen.grid_workload(n, seed=seed, jitter=jitter)


def view_from(eye):
    return meshgen.look_at_view_matrix(eye)


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def make_renderer(d, cells, pos, col, width, height, eye=(0.0, 0.0, 2.0), attenuation=1.0, clear=(0, 0, 0, 0),
                  use_gradients=True, use_ert=False, fovy=FOVY):
    mesh = d.TetMesh()
    mesh.set_use_gradients(use_gradients)
    mesh.set_tet_mesh_data(cells, pos, col)
    r = d.Renderer(renderer_type=d.RendererType.INTERSECTION)
    r.set_tet_mesh(mesh)
    r.set_attenuation(attenuation)
    r.set_clear_color(d.vec4(*clear))
    r.set_viewport_size(width, height)
    r.set_camera_fovy(fovy)
    r.set_view_matrix(view_from(eye))
    r.set_use_early_ray_termination(use_ert)
    return mesh, r


def adjoint_references(oracle, frags, pos, col, cam_f, cam_d, att, dimg, **kw):
    """The oracle's exact adjoint in FP32, and in FP64 with the FP32 evaluation's ray-tet decisions (the FP64
    counterpart of the same function: no edge-grazing ray decides differently)."""
    op, oc = oracle.adjoint_exact(frags, pos, col, cam_f, att, dimg, **kw)
    kw64 = dict(kw)
    if kw64.get("cell_colors") is not None:  # per-cell colours (col None)
        kw64["cell_colors"] = np.asarray(kw64["cell_colors"], np.float64)
    mp, mc = oracle.adjoint_exact(frags, np.asarray(pos, np.float64), None if col is None else np.asarray(col, np.float64),
                                  cam_d, att, np.asarray(dimg, np.float64), dtype=np.float64, f32_decisions=True, **kw64)
    return op, oc, mp, mc


def check_gradients(tag, gp, gc, refs, tol=1e-4):
    """Colours within tol relative L2 of the FP32 oracle's exact adjoint (the same FP32 function: same fragments,
    same ray-tet decisions).  Positions: a vertex's position gradient sums large per-fragment terms that cancel
    across the faces neighbouring tets share, so every FP32 evaluation sits at some distance from the FP64
    gradient with the same decisions, the FP32 floor (the oracle's own is measured here: 1e-5 to 5e-2 over the
    test workloads).  Two bars: the GPU within 1.25 times the floor of FP64 (tol when the floor is below it); and
    the GPU within tol of the FP32 oracle, or within a quarter of the floor where that is larger.  Where the floor
    is large (deep rays through fine Delaunay meshes) the oracle's own error dominates that distance: it forms
    each fragment's input transmittance as 1 - alpha, as the reference does, which keeps only absolute
    precision, so vertices whose exact gradient is ~1e-14 get ~1e-5 from it and ~1e-14 from the GPU, whose
    transmittance is a product.  Returns the measured distances."""
    op, oc, mp, mc = refs
    d = {"col_vs_fp32": rel_l2(gc, oc), "col_vs_fp64": rel_l2(gc, mc), "pos_vs_fp32": rel_l2(gp, op),
         "pos_vs_fp64": rel_l2(gp, mp), "fp32_pos_vs_fp64": rel_l2(op, mp), "fp32_col_vs_fp64": rel_l2(oc, mc)}
    print(f"{tag}: rel L2 vs FP32 oracle: colours {d['col_vs_fp32']:.2e}, positions {d['pos_vs_fp32']:.2e}; "
          f"vs FP64 (same decisions): colours {d['col_vs_fp64']:.2e}, positions {d['pos_vs_fp64']:.2e} "
          f"(FP32 oracle: {d['fp32_pos_vs_fp64']:.2e})", flush=True)
    assert np.isfinite(gp).all() and np.isfinite(gc).all()
    assert d["col_vs_fp32"] <= tol
    assert d["pos_vs_fp32"] <= max(tol, 0.25 * d["fp32_pos_vs_fp64"])
    assert d["pos_vs_fp64"] <= max(tol, 1.25 * d["fp32_pos_vs_fp64"])
    return d
