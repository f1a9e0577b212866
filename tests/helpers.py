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
