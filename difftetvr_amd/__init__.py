"""difftetvr_amd: MI355X (gfx950) differentiable tetrahedral direct volume renderer.

Drop-in for the renderer/optimiser path of the reference Python module ``difftetvr``
(src/Module/Module.cpp:695-1175, stub src/Module/difftetvr.pyi): the same class and method names, argument
meaning, tensor layouts and error behaviour.  The compute runs in hand-written HIP kernels behind the C ABI of
``include/difftetvr_amd.h``; PyTorch supplies device memory and the current stream.

Renderer types: INTERSECTION renders with that renderer's maths (Euclidean segment length through each tet, 4
linear sub-segments, face-centroid depth order); PPLL (the reference's default) with the same tile renderer and
PPLL's segment length, the difference of linear eye depths (PPLL/Forward.glsl:53,188), in the same face-centroid
order (PPLL sorts each pixel's faces by depth; the two orders agree wherever the centroid order is
visibility-correct along the ray); PROJECTION renders projected tetrahedra (3 or 4 screen triangles per tet with
a thickness at the thick vertex, blended in centroid-depth order) and has no early ray termination.  Gradients
are the exact derivative of the forward (``set_adjoint_mode("reference")`` selects the INTERSECTION renderer's
own adjoint), coarse-to-fine refinement chooses prism diagonals by vertex index, and only GPU tensors are
supported.
"""
import ctypes
import enum
import math

import numpy as np
import torch

from . import _lib
from . import meshgen
from ._lib import DtvError

__all__ = ["vec3", "vec4", "uvec3", "AABB3", "ColorStorage", "TestCase", "RendererType", "SplitGradientType",
           "TetMesh", "Renderer", "BatchRenderer", "TetRegularizer", "set_device_type", "DtvError", "meshgen",
           "autotune_cache"]


class vec3:
    def __init__(self, x=0.0, y=0.0, z=0.0):
        self.x, self.y, self.z = float(x), float(y), float(z)

    def __repr__(self):
        return f"vec3({self.x}, {self.y}, {self.z})"


class vec4:
    def __init__(self, x=0.0, y=0.0, z=0.0, w=0.0):
        self.x, self.y, self.z, self.w = float(x), float(y), float(z), float(w)

    def __repr__(self):
        return f"vec4({self.x}, {self.y}, {self.z}, {self.w})"


class uvec3:
    def __init__(self, x=0, y=0, z=0):
        self.x, self.y, self.z = int(x), int(y), int(z)


class AABB3:
    """Axis-aligned box (sgl::AABB3 as bound at Module.cpp:725-735)."""

    def __init__(self, min=None, max=None):  # noqa: A002 - reference argument names
        self.min = min if min is not None else vec3(float("inf"), float("inf"), float("inf"))
        self.max = max if max is not None else vec3(-float("inf"), -float("inf"), -float("inf"))

    def get_dimensions(self):
        return vec3(self.max.x - self.min.x, self.max.y - self.min.y, self.max.z - self.min.z)

    def get_extent(self):
        d = self.get_dimensions()
        return vec3(d.x * 0.5, d.y * 0.5, d.z * 0.5)

    def get_center(self):
        return vec3((self.max.x + self.min.x) * 0.5, (self.max.y + self.min.y) * 0.5, (self.max.z + self.min.z) * 0.5)

    def get_minimum(self):
        return self.min

    def get_maximum(self):
        return self.max


class ColorStorage(enum.IntEnum):
    PER_VERTEX = 0
    PER_CELL = 1


class TestCase(enum.IntEnum):
    SINGLE_TETRAHEDRON = 0
    CUBE_CENTRAL_GRADIENT = 1


class SplitGradientType(enum.IntEnum):
    """Which gradient ranks the vertices for coarse-to-fine splitting (OptimizerDefines.hpp:50-52).  The ABS_
    variants read the same buffers; the caller accumulates them with ``use_abs_grad`` (train.py:369-404)."""
    POSITION = 0
    COLOR = 1
    ABS_POSITION = 2
    ABS_COLOR = 3


class RendererType(enum.IntEnum):
    PPLL = 0
    PROJECTION = 1
    INTERSECTION = 2


_state = {"device_type": "cuda", "ctx": None, "objects_created": False}


def set_device_type(device_type):
    """Selects the torch device type before any object is created (Module.cpp:617-644).  Only GPU devices
    ('cuda', 'hip') are supported: there is no CPU path."""
    if _state["objects_created"]:
        raise RuntimeError("set_device_type must be called before any TetMesh or Renderer is created.")
    dt = str(device_type).lower()
    if dt in ("cuda", "hip"):
        _state["device_type"] = "cuda"
    else:
        raise RuntimeError(f"Unsupported device type '{device_type}': difftetvr_amd runs on the GPU only.")


def _ctx():
    if _state["ctx"] is None:
        if not torch.cuda.is_available():
            raise DtvError("difftetvr_amd requires a visible GPU (torch.cuda.is_available() is False).")
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(lib.dtv_ctx_create(torch.cuda.current_device(), ctypes.byref(h)), "dtv_ctx_create")
        _state["ctx"] = h
    _state["objects_created"] = True
    return _state["ctx"]


def _device():
    return torch.device("cuda", torch.cuda.current_device())


def _stream_ptr():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class TetMesh:
    """Tet mesh with per-vertex colours (TetMesh, src/Tet/TetMesh.hpp:89-286).

    get_vertex_positions()/get_vertex_colors() return the same device tensors on every call; when gradients
    are enabled their ``.grad`` is preallocated and ``Renderer.render_adjoint`` accumulates into it
    (TetMesh.cpp:236-367).
    """

    def __init__(self):
        _ctx()
        self._use_gradients = False
        self._cells = None
        self._positions = None
        self._colors = None       # [V,4] (PER_VERTEX) or [T,4] (PER_CELL)
        self._boundary = None
        self._color_storage = ColorStorage.PER_VERTEX

    # -- construction -------------------------------------------------------------------------------------
    def set_use_gradients(self, use_gradients=True):
        self._use_gradients = bool(use_gradients)
        if self._positions is not None:
            self._setup_grads()

    def get_use_gradients(self):
        return self._use_gradients

    def set_tet_mesh_data(self, cells, vertex_positions, vertex_colors=None, cell_colors=None):
        """Sets cells [T,4] (int), positions [V,3] and either per-vertex colours [V,4] (ColorStorage.PER_VERTEX)
        or per-cell colours [T,4] (ColorStorage.PER_CELL) (array-likes or tensors)."""
        dev = _device()
        if (vertex_colors is None) == (cell_colors is None):
            raise RuntimeError("set_tet_mesh_data: give exactly one of vertex_colors and cell_colors.")
        cells = torch.as_tensor(np.asarray(cells) if not torch.is_tensor(cells) else cells)
        cells = cells.to(device=dev, dtype=torch.int32).reshape(-1, 4).contiguous()
        pos = torch.as_tensor(vertex_positions if torch.is_tensor(vertex_positions) else np.asarray(vertex_positions))
        pos = pos.to(device=dev, dtype=torch.float32).reshape(-1, 3).contiguous()
        colors = vertex_colors if cell_colors is None else cell_colors
        col = torch.as_tensor(colors if torch.is_tensor(colors) else np.asarray(colors))
        col = col.to(device=dev, dtype=torch.float32).reshape(-1, 4).contiguous()
        if cell_colors is None and col.shape[0] != pos.shape[0]:
            raise RuntimeError("set_tet_mesh_data: vertex colours must be [V,4].")
        if cell_colors is not None and col.shape[0] != cells.shape[0]:
            raise RuntimeError("set_tet_mesh_data: cell colours must be [T,4].")
        if cells.numel() and (int(cells.min()) < 0 or int(cells.max()) >= pos.shape[0]):
            raise RuntimeError("set_tet_mesh_data: cell index out of range.")
        self._cells = cells
        self._positions = pos.detach().clone()
        self._colors = col.detach().clone()
        self._color_storage = ColorStorage.PER_VERTEX if cell_colors is None else ColorStorage.PER_CELL
        self._boundary = None
        self._setup_grads()

    def set_triangle_mesh_data(self, triangle_indices, vertex_positions, vertex_colors):
        """Stores a triangle surface mesh (TetMesh::setTriangleMeshData, TetMesh.cpp:369-478): int32 triangle
        indices [F,3], float32 positions [V,3] and colours [V,4], with the reference's checks and messages.  The
        tet renderers refuse such a mesh (it has no tets); the tensors and counts are available as for tets."""
        def _t(x):
            return x if torch.is_tensor(x) else torch.as_tensor(np.asarray(x))
        tri, pos, col = _t(triangle_indices), _t(vertex_positions), _t(vertex_colors)
        name = "Error in TetMesh::setTriangleMeshData: "
        if tri.dtype != torch.int32:
            raise RuntimeError(name + "triangleIndicesTensor.dtype() != torch::kInt32.")
        if pos.dtype != torch.float32:
            raise RuntimeError(name + "vertexPositionsTensor.dtype() != torch::kFloat32.")
        if col.dtype != torch.float32:
            raise RuntimeError(name + "vertexColorsTensor.dtype() != torch::kFloat32.")
        for t, what in ((tri, "triangleIndicesTensor"), (pos, "vertexPositionsTensor"), (col, "vertexColorsTensor")):
            if t.dim() != 2:
                raise RuntimeError(name + f"{what}.sizes().size() != 2.")
        if pos.shape[0] != col.shape[0]:
            raise RuntimeError(name + "vertexPositionsTensor.size(0) != vertexColorsTensor.size(0).")
        if tri.shape[1] != 3:
            raise RuntimeError(name + "triangleIndicesTensor.size(1) != 3.")
        if pos.shape[1] != 3:
            raise RuntimeError(name + "vertexPositionsTensor.size(1) != 3.")
        if col.shape[1] != 4:
            raise RuntimeError(name + "vertexColorsTensor.size(1) != 3.")
        dev = _device()
        self._cells = tri.to(device=dev).contiguous()
        self._positions = pos.to(device=dev).detach().clone().contiguous()
        self._colors = col.to(device=dev).detach().clone().contiguous()
        self._color_storage = ColorStorage.PER_VERTEX
        self._boundary = None
        self._setup_grads()

    def get_has_triangle_mesh_data(self):
        """True after set_triangle_mesh_data (TetMesh::getHasTriangleMeshData)."""
        return self._cells is not None and self._cells.dim() == 2 and self._cells.shape[1] == 3

    def _face_boundary_bits(self):
        """bool [T,4]: face f of tet t lies on the mesh boundary (no other tet shares it); cached per mesh."""
        key = (self._cells.data_ptr(), self._cells.shape[0])
        if getattr(self, "_fb_key", None) != key:
            c = self._cells.cpu().numpy().astype(np.int64)
            faces = np.sort(c[:, meshgen.TET_FACE_TABLE[:, :3]], axis=2).reshape(-1, 3)
            _, inv, cnt = np.unique(faces, axis=0, return_inverse=True, return_counts=True)
            self._fb = (cnt[inv.reshape(-1)] == 1).reshape(-1, 4)
            self._fb_key = key
        return self._fb

    def _set_like(self, cells, pos, col):
        """set_tet_mesh_data with this mesh's colour storage."""
        if self._color_storage == ColorStorage.PER_CELL:
            self.set_tet_mesh_data(cells, pos, cell_colors=col)
        else:
            self.set_tet_mesh_data(cells, pos, col)

    def _setup_grads(self):
        if self._use_gradients:
            self._positions.requires_grad_(True)
            self._colors.requires_grad_(True)
            if self._positions.grad is None:
                self._positions.grad = torch.zeros_like(self._positions)
            if self._colors.grad is None:
                self._colors.grad = torch.zeros_like(self._colors)

    def set_hex_mesh_const(self, aabb, xs, ys, zs, const_color, color_storage=ColorStorage.PER_VERTEX):
        """Tetrahedralised hex grid with xs*ys*zs vertices and a constant colour per vertex or per cell
        (TetMesh.cpp:1882-1936)."""
        pos, cells = meshgen.hex_grid((aabb.min.x, aabb.min.y, aabb.min.z), (aabb.max.x, aabb.max.y, aabb.max.z),
                                      int(xs), int(ys), int(zs))
        c = np.array([const_color.x, const_color.y, const_color.z, const_color.w], np.float32)
        if ColorStorage(color_storage) == ColorStorage.PER_CELL:
            self.set_tet_mesh_data(cells, pos, cell_colors=np.tile(c, (cells.shape[0], 1)))
        else:
            self.set_tet_mesh_data(cells, pos, np.tile(c, (pos.shape[0], 1)))

    def load_test_data(self, test_case):
        """Built-in meshes (TetMesh::loadTestData, TetMesh.cpp:1016-1077)."""
        if TestCase(test_case) == TestCase.SINGLE_TETRAHEDRON:
            pos, cells, col = meshgen.single_tetrahedron()
        else:
            pos, cells, col = meshgen.cube_central_gradient()
        self.set_tet_mesh_data(cells, pos, col)

    def load_from_file(self, file_path):
        from . import bintet
        cells, pos, col, cell_col = bintet.read(file_path)
        self.set_tet_mesh_data(cells, pos, col, cell_col)
        return True

    def save_to_file(self, file_path):
        from . import bintet
        col = self._colors.detach().cpu().numpy()
        per_cell = self._color_storage == ColorStorage.PER_CELL
        bintet.write(file_path, self._cells.cpu().numpy(), self._positions.detach().cpu().numpy(),
                     None if per_cell else col, col if per_cell else None)
        return True

    # -- queries ------------------------------------------------------------------------------------------
    def _require(self):
        if self._cells is None:
            raise RuntimeError("TetMesh is empty.")

    def get_num_cells(self):
        return 0 if self._cells is None else int(self._cells.shape[0])

    def get_num_vertices(self):
        return 0 if self._positions is None else int(self._positions.shape[0])

    def get_is_empty(self):
        return self.get_num_cells() == 0

    def get_vertex_positions(self):
        self._require()
        return self._positions

    def get_vertex_colors(self):
        """float32 [V,4] (TetMesh::getVertexColorTensor, TetMesh.cpp:271-308)."""
        self._require()
        if self._color_storage != ColorStorage.PER_VERTEX:
            raise RuntimeError("Error in TetMesh::getVertexColorTensor: No vertex color data available.")
        return self._colors

    def get_cell_colors(self):
        """float32 [T,4] (TetMesh::getCellColorTensor, TetMesh.cpp:310-345)."""
        self._require()
        if self._color_storage != ColorStorage.PER_CELL:
            raise RuntimeError("Error in TetMesh::getCellColorTensor: No cell color data available.")
        return self._colors

    def _color_tensor(self):
        """The colour tensor of either storage (its .grad is the adjoint's colour target)."""
        self._require()
        return self._colors

    def get_cell_indices(self):
        self._require()
        return self._cells

    def get_color_storage(self):
        return self._color_storage

    def get_vertex_boundary_bit_tensor(self):
        """int32 [V,1], 1 for vertices on boundary faces (TetMesh.cpp:1261-1267, Module.cpp:809-812)."""
        self._require()
        if self._boundary is None:
            mask = meshgen.boundary_vertex_mask(self._cells.cpu().numpy(), self.get_num_vertices())
            self._boundary = torch.as_tensor(mask.astype(np.int32)).reshape(-1, 1).to(_device())
        return self._boundary

    def get_bounding_box(self):
        self._require()
        p = self._positions.detach()
        mn, mx = p.min(dim=0).values.tolist(), p.max(dim=0).values.tolist()
        return AABB3(vec3(*mn), vec3(*mx))

    def check_is_any_tet_degenerate(self):
        """True if any tet has volume <= 0 under the reference orientation (TetMesh.cpp:544-558)."""
        self._require()
        p = self._positions.detach()[self._cells.long()]
        o = torch.sum(torch.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], dim=1) * (p[:, 3] - p[:, 0]), dim=1)
        return bool((o >= 0).any())

    def unlink_tets(self):
        """Gives every tet its own four vertices, undoing the shared-index representation (TetMesh.cpp:1950-1989).
        The mesh tensors are new afterwards."""
        self._require()
        per_cell = self._color_storage == ColorStorage.PER_CELL
        cells, pos, col = meshgen.unlink_tets(self._cells.cpu().numpy(), self._positions.detach().cpu().numpy(),
                                              None if per_cell else self._colors.detach().cpu().numpy())
        self._set_like(cells, pos, self._colors.detach().cpu().numpy() if per_cell else col)

    def remove_transparent_tets(self, alpha_threshold):
        """Removes the tets none of whose vertices has alpha >= alpha_threshold, and the vertices no remaining tet
        uses (TetMesh.cpp:1991-2079).  The mesh tensors are new afterwards."""
        self._require()
        per_cell = self._color_storage == ColorStorage.PER_CELL
        col = self._colors.detach().cpu().numpy()
        cells, pos, vcol, ccol = meshgen.remove_transparent_tets(
            self._cells.cpu().numpy(), self._positions.detach().cpu().numpy(), None if per_cell else col,
            float(alpha_threshold), cell_colors=col if per_cell else None)
        self._set_like(cells, pos, ccol if per_cell else vcol)

    def set_force_use_ovm_representation(self):
        """Accepted for compatibility (Module.cpp:794): refinement works on the plain cell array."""

    def split_by_largest_gradient_magnitudes(self, volume_renderer, split_gradient_type, splits_ratio):
        """Coarse-to-fine step (Module.cpp:820-832, TetMesh.cpp:1786-1879): subdivides the ceil(splits_ratio*V)
        vertices with the longest position (POSITION, ABS_POSITION) or colour (COLOR, ABS_COLOR) gradient
        rows, then re-creates the mesh tensors (new tensors: optimisers must be given them, as train.py:413-420
        does) and hands the mesh to ``volume_renderer``."""
        from . import refine
        self._require()
        if self._color_storage != ColorStorage.PER_VERTEX:
            raise RuntimeError("split_by_largest_gradient_magnitudes: subdivision interpolates vertex colours; "
                               "the mesh uses ColorStorage.PER_CELL.")
        kind = SplitGradientType(split_gradient_type)
        src = self._positions if kind in (SplitGradientType.POSITION, SplitGradientType.ABS_POSITION) else self._colors
        if src.grad is None:
            raise RuntimeError("split_by_largest_gradient_magnitudes: the mesh has no gradients "
                               "(set_use_gradients(True) and accumulate them first).")
        torch.cuda.synchronize()
        cells, pos, col = refine.split_by_largest_gradient_magnitudes(
            self._cells.cpu().numpy(), self._positions.detach().cpu().numpy(), self._colors.detach().cpu().numpy(),
            src.grad.detach().cpu().numpy(), float(splits_ratio))
        self.set_tet_mesh_data(cells, pos, col)
        if volume_renderer is not None:
            volume_renderer.set_tet_mesh(self)

    def set_vertices_changed(self, vertices_changed=True):
        """Device tensors are the render inputs, so there is nothing to upload (TetMesh.cpp:788-818)."""

    def on_zero_grad(self):
        """No-op on the GPU device (TetMesh.cpp:820-834)."""

    def _desc(self):
        self._require()
        if self.get_has_triangle_mesh_data():
            raise RuntimeError("The mesh holds triangle data (set_triangle_mesh_data): there are no tets to render.")
        per_cell = self._color_storage == ColorStorage.PER_CELL
        return _lib.DtvMesh(self._cells.data_ptr(), self._positions.data_ptr(),
                            None if per_cell else self._colors.data_ptr(), self._cells.shape[0],
                            self._positions.shape[0], self._colors.data_ptr() if per_cell else None)


_TUNE_CACHE = {}


def autotune_cache():
    """The autotuner's choices so far: {key: {tile_size, segments_per_pass, ms, ms_default, candidates}}."""
    return dict(_TUNE_CACHE)


class Renderer:
    """Tet-mesh volume renderer (TetMeshVolumeRenderer, src/Renderer/TetMeshVolumeRenderer.hpp:84-280)."""

    def __init__(self, renderer_type=RendererType.PPLL):
        ctx = _ctx()
        lib = _lib.load()
        self._renderer_type = RendererType(renderer_type)
        h = ctypes.c_void_p()
        _lib.check(lib.dtv_view_create(ctx, ctypes.byref(h)), "dtv_view_create")
        self._view = h
        self._lib = lib
        self._mesh = None
        self._attenuation = 100.0                   # TetMeshVolumeRenderer.hpp:237
        self._clear = (0.0, 0.0, 0.0, 0.0)
        self._use_ert = True                        # :243-244
        self._ert_thresh = 1e-4
        self._width, self._height = 1024, 1024
        self._fovy = 2.0 * math.atan(0.5)           # Module.cpp:461
        self._znear, self._zfar = 0.01, 100.0       # Module.cpp:459-460
        self._view_matrix = meshgen.look_at_view_matrix((0.0, 0.0, 2.0))
        self._coarse_to_fine_target = 0
        self._adjoint_mode = _lib.DTV_ADJOINT_EXACT
        self._checkpoint_interval = 64
        self._tile_order = 0
        self._rendered_mesh = None

    def __del__(self):
        try:
            if getattr(self, "_view", None) is not None:
                torch.cuda.synchronize()
                self._lib.dtv_view_destroy(self._view)
                self._view = None
        except Exception:
            pass

    def get_renderer_type(self):
        return self._renderer_type

    def set_tet_mesh(self, tet_mesh):
        self._mesh = tet_mesh

    def get_tet_mesh(self):
        return self._mesh

    def get_attenuation(self):
        return self._attenuation

    def set_attenuation(self, attenuation_coefficient):
        self._attenuation = float(attenuation_coefficient)

    def set_coarse_to_fine_target_num_tets(self, target_num_tets):
        self._coarse_to_fine_target = int(target_num_tets)

    def set_use_early_ray_termination(self, use_early_ray_termination):
        self._use_ert = bool(use_early_ray_termination)

    def set_early_ray_out_thresh(self, threshold):
        self._ert_thresh = float(threshold)

    def set_early_ray_out_alpha(self, alpha):
        self._ert_thresh = 1.0 - float(alpha)

    def set_clear_color(self, color):
        self._clear = (color.x, color.y, color.z, color.w)

    def set_viewport_size(self, image_width, image_height, recreate_swapchain=True):
        if image_width <= 0 or image_height <= 0:
            raise RuntimeError("set_viewport_size: width and height must be positive.")
        self._width, self._height = int(image_width), int(image_height)

    def reuse_intermediate_buffers_from(self, renderer_other):
        """Buffers are per view and grown on demand; nothing to share."""

    def set_camera_fovy(self, fovy):
        self._fovy = float(fovy)

    def set_view_matrix(self, view_matrix_array):
        a = np.asarray(view_matrix_array, np.float32).reshape(-1)
        if a.size != 16:
            raise RuntimeError("set_view_matrix: 16 values (column-major) expected.")
        self._view_matrix = a.copy()

    def set_adjoint_mode(self, mode):
        """'exact' (default): derivative of the forward.  'reference': AdjointIntersectRasterization.glsl as
        written, including its skipped first sub-segment (INTERSECTION renderers only)."""
        m = {"exact": _lib.DTV_ADJOINT_EXACT, "reference": _lib.DTV_ADJOINT_REFERENCE}.get(str(mode).lower())
        if m is None:
            raise RuntimeError(f"set_adjoint_mode: unknown mode '{mode}'.")
        self._adjoint_mode = m

    def set_checkpoint_interval(self, k):
        self._checkpoint_interval = int(k)

    def set_tile_size(self, tile_size):
        """Binning tile edge in pixels: 8 (default), 16 or 32 (rendered as 8x8 sub-tiles; images do not depend
        on it).  No reference counterpart."""
        if int(tile_size) not in (8, 16, 32):
            raise RuntimeError("set_tile_size: 8, 16 or 32 expected.")
        self._tile_size = int(tile_size)

    def set_segments_per_pass(self, segments):
        """Ray segments the forward shades per round: 64, 128 (default) or 256.  No reference counterpart."""
        if int(segments) not in (64, 128, 256):
            raise RuntimeError("set_segments_per_pass: 64, 128 or 256 expected.")
        self._segments_per_pass = int(segments)

    def set_tile_launch(self, launch):
        """How the tile kernels are launched: 'grid' (default: one workgroup per 8x8 tile) or 'queue' (a
        persistent grid of the workgroups the GPU holds at once, taking the tiles in launch order from a device
        work queue; the adjoint follows its render).  Images are identical either way.  No reference counterpart."""
        o = {"grid": _lib.DTV_TILE_LAUNCH_GRID, "queue": _lib.DTV_TILE_LAUNCH_QUEUE}.get(str(launch).lower())
        if o is None:
            raise RuntimeError(f"set_tile_launch: unknown launch '{launch}' ('grid' or 'queue').")
        self._tile_launch = o

    def set_autotune(self, enabled=True, with_adjoint=True, repeats=3):
        """With autotuning on, the first render of a (mesh size class, viewport) times every tile size and
        segments-per-pass setting (``dtv_autotune``: forward, and adjoint with with_adjoint) and the fastest is
        used from then on; choices are cached per process in ``autotune_cache()``.  Tile renderer only."""
        self._autotune = (bool(enabled), bool(with_adjoint), int(repeats))

    def get_tuning(self):
        """The (tile_size, segments_per_pass) the last render used."""
        return (self._tile_size, self._segments_per_pass)

    def _tune_key(self, items):
        def cls(m):
            return (int(m.get_num_cells()).bit_length(), int(m.get_color_storage()))
        return (int(self._renderer_type), tuple((cls(m), w, h) for m, w, h in items))

    def _maybe_tune(self, arr, n, items):
        on, with_adj, repeats = getattr(self, "_autotune", (False, True, 3))
        if not on or self._renderer_type == RendererType.PROJECTION:
            return
        key = self._tune_key(items)
        if key not in _TUNE_CACHE:
            res = _lib.DtvTuneResult()
            prm = self._params()
            _lib.check(self._lib.dtv_autotune(self._view, arr, n, ctypes.byref(prm), int(with_adj), repeats,
                                              ctypes.byref(res), _stream_ptr()), "dtv_autotune")
            _TUNE_CACHE[key] = {"tile_size": res.tile_size, "segments_per_pass": res.segments_per_pass,
                                "ms": res.ms, "ms_default": res.ms_default,
                                "candidates": [(res.cand_tile_size[i], res.cand_segments_per_pass[i], res.cand_ms[i])
                                               for i in range(res.num_candidates)]}
        choice = _TUNE_CACHE[key]
        self._tile_size, self._segments_per_pass = choice["tile_size"], choice["segments_per_pass"]

    def set_tile_order(self, order):
        """Launch order of the screen tiles: 'cost' (default: heaviest first) or 'row-major'.  Images and
        gradients do not depend on it (up to the order of the gradients' float atomics)."""
        o = {"cost": 0, "row-major": 1}.get(str(order).lower())
        if o is None:
            raise RuntimeError(f"set_tile_order: unknown order '{order}'.")
        self._tile_order = o

    def set_record_capacity(self, capacity):
        """Planned rendering (``dtv_view_set_record_capacity``): > 0 sizes the (tile, face) record buffers for
        ``capacity`` records and removes the per-render host synchronisation, so that render / adjoint / Adam
        can be captured in a CUDA/HIP graph (after one ordinary call).  0 = off."""
        _lib.check(self._lib.dtv_view_set_record_capacity(self._view, int(capacity)), "set_record_capacity")

    def records_overflowed(self):
        """(overflowed, records) of the last render; synchronises the device."""
        o, k = ctypes.c_int32(), ctypes.c_int64()
        _lib.check(self._lib.dtv_view_records_overflowed(self._view, ctypes.byref(o), ctypes.byref(k)),
                   "records_overflowed")
        return bool(o.value), int(k.value)

    def _camera(self):
        cam = _lib.DtvCamera()
        for i in range(16):
            cam.view[i] = float(self._view_matrix[i])
        cam.fovy, cam.znear, cam.zfar = self._fovy, self._znear, self._zfar
        cam.width, cam.height = self._width, self._height
        return cam

    def _params(self):
        p = _lib.DtvRenderParams()
        p.attenuation = self._attenuation
        for i in range(4):
            p.clear_color[i] = self._clear[i]
        # ERT is compiled in only for meshes with gradients (TetMeshVolumeRenderer.cpp:417-419).
        p.use_early_ray_termination = int(self._use_ert and self._mesh is not None and self._mesh.get_use_gradients())
        p.early_ray_out_alpha = 1.0 - self._ert_thresh
        p.checkpoint_interval = self._checkpoint_interval
        p.tile_order = self._tile_order
        # PPLL composites between consecutive faces by their linear depths (PPLL/Forward.glsl:53,188)
        p.segment_length = (_lib.DTV_SEGMENT_LINEAR_DEPTH if self._renderer_type == RendererType.PPLL
                            else _lib.DTV_SEGMENT_EUCLIDEAN)
        p.renderer = (_lib.DTV_RENDERER_PROJECTION if self._renderer_type == RendererType.PROJECTION
                      else _lib.DTV_RENDERER_TILE)
        proj = self._renderer_type == RendererType.PROJECTION
        p.tile_size = 8 if proj else getattr(self, "_tile_size", 8)
        p.segments_per_pass = getattr(self, "_segments_per_pass", 128)
        p.tile_launch = _lib.DTV_TILE_LAUNCH_GRID if proj else getattr(self, "_tile_launch", _lib.DTV_TILE_LAUNCH_GRID)
        return p

    def camera_uniforms(self):
        out = (ctypes.c_float * 68)()
        cam = self._camera()
        _lib.check(self._lib.dtv_camera_uniforms(ctypes.byref(cam), out), "dtv_camera_uniforms")
        return np.frombuffer(out, np.float32).copy()

    def render(self):
        """Renders the mesh; returns premultiplied RGBA [H, W, 4] float32 (Module.cpp:912-921)."""
        if self._mesh is None or self._mesh.get_is_empty():
            raise RuntimeError("Missing valid tet mesh in Renderer.render.")
        img = torch.empty((self._height, self._width, 4), dtype=torch.float32, device=_device())
        desc = self._mesh._desc()
        cam = self._camera()
        if getattr(self, "_autotune", (False,))[0]:
            item = (_lib.DtvRenderItem * 1)()
            item[0].mesh, item[0].camera, item[0].out_image = desc, cam, img.data_ptr()
            self._maybe_tune(item, 1, [(self._mesh, self._width, self._height)])
        prm = self._params()
        _lib.check(self._lib.dtv_render(self._view, ctypes.byref(desc), ctypes.byref(cam), ctypes.byref(prm),
                                        ctypes.c_void_p(img.data_ptr()), _stream_ptr()), "Renderer.render")
        self._rendered_mesh = self._mesh
        return img

    def render_adjoint(self, image_adjoint, use_abs_grad=False):
        """Accumulates d(loss)/d(positions, colours) into the mesh tensors' .grad (Module.cpp:922-959)."""
        if not torch.is_tensor(image_adjoint) or image_adjoint.device.type != "cuda":
            raise RuntimeError(
                "Error in render_adjoint: The tensors must be on the device used at initialization time.")
        if image_adjoint.dim() != 3:
            raise RuntimeError("Error in render_adjoint: imageAdjointTensor.sizes().size() != 3.")
        if image_adjoint.shape[2] != 4:
            raise RuntimeError("Error in render_adjoint: The number of image channels is not equal to 4.")
        if image_adjoint.dtype != torch.float32:
            raise RuntimeError("Error in render_adjoint: The only data type currently supported is 32-bit float.")
        if tuple(image_adjoint.shape[:2]) != (self._height, self._width):
            raise RuntimeError("Error in render_adjoint: image size differs from the viewport.")
        if self._rendered_mesh is None or self._rendered_mesh is not self._mesh:
            raise RuntimeError("Error in render_adjoint: render() must be called first.")
        image_adjoint = image_adjoint.contiguous()
        mesh = self._mesh
        pos, col = mesh.get_vertex_positions(), mesh._color_tensor()
        if pos.grad is None:
            pos.grad = torch.zeros_like(pos)
        if col.grad is None:
            col.grad = torch.zeros_like(col)
        desc = mesh._desc()
        _lib.check(self._lib.dtv_render_adjoint(self._view, ctypes.byref(desc),
                                                ctypes.c_void_p(image_adjoint.data_ptr()), int(bool(use_abs_grad)),
                                                self._adjoint_mode, ctypes.c_void_p(pos.grad.data_ptr()),
                                                ctypes.c_void_p(col.grad.data_ptr()), _stream_ptr()),
                   "Renderer.render_adjoint")

    # -- inspection (counterparts of the PPLL export, Module.cpp:960-981) ----------------------------------
    def get_fragment_counts(self):
        """int32 [H, W]: fragments composited per pixel by the last render."""
        out = torch.empty((self._height, self._width), dtype=torch.int32, device=_device())
        _lib.check(self._lib.dtv_view_fragment_counts(self._view, ctypes.c_void_p(out.data_ptr()), _stream_ptr()),
                   "get_fragment_counts")
        return out

    def get_fragment_ids(self, max_per_pixel):
        """int32 [H, W, max_per_pixel]: fragment ids (tet*4 + face) front to back, -1 padded."""
        out = torch.empty((self._height, self._width, int(max_per_pixel)), dtype=torch.int32, device=_device())
        desc = self._rendered_mesh._desc()
        _lib.check(self._lib.dtv_view_export_fragments(self._view, ctypes.byref(desc), int(max_per_pixel),
                                                       ctypes.c_void_p(out.data_ptr()), _stream_ptr()),
                   "get_fragment_ids")
        return out

    # -- PPLL linked-list export (Module.cpp:960-981, TetMeshRendererPPLL.cpp:809-860) -------------------------
    def set_export_linked_list_data(self):
        if self._renderer_type != RendererType.PPLL:
            raise RuntimeError("Error in set_export_linked_list_data: The renderer must be of type RendererType.PPLL.")
        self._export_ll = True

    def _linked_list_checked(self, fn, what):
        if self._renderer_type != RendererType.PPLL:
            raise RuntimeError(f"Error in {fn}: The renderer must be of type RendererType.PPLL.")
        if not getattr(self, "_export_ll", False):
            raise RuntimeError(f"Error in {what}: exportLinkedListData is set to false.")
        if self._rendered_mesh is None:
            raise RuntimeError(f"Error in {fn}: render() must be called first.")
        return self._linked_list()

    def get_fragment_buffer(self):
        """float32 [N, 3] of 12-byte nodes {faceBits, depth, next} (LinkedListHeader.glsl:35-42; the first and
        third words are uint32 bit patterns): one node per composited fragment, faceBits = (fragment id tet*4 +
        face) << 2 | boundary << 1 | 1 (the entered face, front-facing), depth = the face's window depth at the
        pixel centre, next = the pixel's next fragment front to back (0xFFFFFFFF at the end)."""
        return self._linked_list_checked("get_fragment_buffer", "getFragmentBufferTensor")[0]

    def get_start_offset_buffer(self):
        """int32 [H, W]: index of each pixel's first (nearest) node, -1 where none."""
        return self._linked_list_checked("get_start_offset_buffer", "getStartOffsetBufferTensor")[1]

    def _linked_list(self):
        mesh = self._rendered_mesh
        counts = self.get_fragment_counts()
        m = max(int(counts.max()), 1)
        ids = self.get_fragment_ids(m)
        valid = ids >= 0
        fid = ids[valid].to(torch.int64)  # pixel-major, front to back within a pixel
        n = int(fid.numel())
        tet, face = fid >> 2, fid & 3
        ft = torch.as_tensor(meshgen.TET_FACE_TABLE[:, :3], device=fid.device)
        cells = mesh.get_cell_indices().to(torch.int64)
        vid = cells[tet.unsqueeze(1), ft[face]]  # [n, 3]
        u = torch.as_tensor(self.camera_uniforms(), device=fid.device)
        vp = u[:16].view(4, 4).t()
        p = mesh.get_vertex_positions().detach()[vid]  # [n, 3, 3]
        clip = p @ vp[:3, :3].t() + vp[:3, 3]
        w = p @ vp[3, :3] + vp[3, 3]
        ndc = clip / w.unsqueeze(-1)
        pix = valid.nonzero()[:, :2].to(torch.float32)  # (y, x) of each node
        px = 2.0 * ((pix[:, 1] + 0.5) / self._width) - 1.0
        py = 2.0 * ((pix[:, 0] + 0.5) / self._height) - 1.0
        X, Y, Z = ndc[..., 0], ndc[..., 1], ndc[..., 2]
        den = (X[:, 1] - X[:, 0]) * (Y[:, 2] - Y[:, 0]) - (X[:, 2] - X[:, 0]) * (Y[:, 1] - Y[:, 0])
        l0 = ((X[:, 1] - px) * (Y[:, 2] - py) - (X[:, 2] - px) * (Y[:, 1] - py)) / den
        l1 = ((X[:, 2] - px) * (Y[:, 0] - py) - (X[:, 0] - px) * (Y[:, 2] - py)) / den
        depth = l0 * Z[:, 0] + l1 * Z[:, 1] + (1.0 - l0 - l1) * Z[:, 2]
        boundary = torch.as_tensor(mesh._face_boundary_bits(), device=fid.device)[tet, face].to(torch.int64)
        bits = ((fid << 2) | (boundary << 1) | 1).to(torch.int32)
        start = torch.full((self._height, self._width), -1, dtype=torch.int32, device=fid.device)
        first = torch.cumsum(counts.reshape(-1).to(torch.int64), 0) - counts.reshape(-1).to(torch.int64)
        has = counts.reshape(-1) > 0
        start.view(-1)[has] = first[has].to(torch.int32)
        nxt = torch.arange(1, n + 1, dtype=torch.int64, device=fid.device)
        last = (first + counts.reshape(-1).to(torch.int64) - 1)[has]
        nxt[last] = 0xFFFFFFFF
        nxt = nxt.to(torch.int64).bitwise_and(0xFFFFFFFF)
        nxt32 = torch.where(nxt > 0x7FFFFFFF, nxt - (1 << 32), nxt).to(torch.int32)
        buf = torch.stack([bits.view(torch.float32), depth.to(torch.float32), nxt32.view(torch.float32)], dim=1)
        return buf, start

    def set_timing(self, enabled=True):
        """Enables per-stage hipEvent timers on this renderer's view."""
        _lib.check(self._lib.dtv_view_set_timing(self._view, int(bool(enabled))), "set_timing")

    def get_timings(self):
        """Milliseconds of the most recent run of each pipeline stage (synchronises on its events)."""
        out = (ctypes.c_float * len(_lib.STAGES))()
        _lib.check(self._lib.dtv_view_get_timings(self._view, out), "get_timings")
        return {name: float(out[i]) for i, name in enumerate(_lib.STAGES) if out[i] >= 0.0}

    def get_stats(self):
        st = _lib.DtvViewStats()
        _lib.check(self._lib.dtv_view_get_stats(self._view, ctypes.byref(st)), "get_stats")
        return {"num_triangles": st.num_triangles, "num_pairs": st.num_pairs, "num_tiles": st.num_tiles,
                "sort_passes": st.sort_passes, "width": st.width, "height": st.height, "num_items": st.num_items,
                "num_pixels": st.num_pixels, "num_gradient_atomics": st.num_gradient_atomics}


class BatchRenderer(Renderer):
    """Renders many (mesh, camera) items as one batch: one binning grid over all their tets, one record sort and
    one tile-renderer launch over all their tiles (``dtv_render_batch``).  Each image equals what ``Renderer``
    gives for its item alone.  No reference counterpart (the reference renders one view per call); the render
    settings (attenuation, clear colour, early ray termination, fovy, clip planes) are ``Renderer``'s and apply
    to every item.

    ``set_items([(tet_mesh, view_matrix, width, height), ...])`` or ``set_views(tet_mesh, view_matrices, width,
    height)``; ``render()`` returns [N, H, W, 4] when all items share a viewport, else a list of [H_i, W_i, 4];
    ``render_adjoint(image_adjoints)`` accumulates into every item mesh's ``.grad``.  ``renderer_type``
    (INTERSECTION by default) applies to every item of the batch.
    """

    def __init__(self, renderer_type=RendererType.INTERSECTION):
        super().__init__(renderer_type=renderer_type)
        self._items = []
        self._rendered_items = None
        self._uniform_size = True

    def set_items(self, items):
        parsed = []
        for mesh, view_matrix, width, height in items:
            a = np.asarray(view_matrix, np.float32).reshape(-1)
            if a.size != 16:
                raise RuntimeError("set_items: 16 values (column-major) expected per view matrix.")
            if int(width) <= 0 or int(height) <= 0:
                raise RuntimeError("set_items: width and height must be positive.")
            parsed.append((mesh, a.copy(), int(width), int(height)))
        if not parsed:
            raise RuntimeError("set_items: at least one item is required.")
        self._items = parsed
        self._mesh = parsed[0][0]

    def set_views(self, tet_mesh, view_matrices, width, height):
        self.set_items([(tet_mesh, v, width, height) for v in view_matrices])

    def get_num_items(self):
        return len(self._items)

    def _params(self):
        p = super()._params()
        p.use_early_ray_termination = int(self._use_ert and all(m.get_use_gradients() for m, _, _, _ in self._items))
        return p

    def render(self):
        if not self._items:
            raise RuntimeError("Missing items in BatchRenderer.render.")
        for mesh, _, _, _ in self._items:
            if mesh is None or mesh.get_is_empty():
                raise RuntimeError("Missing valid tet mesh in BatchRenderer.render.")
        sizes = [(h, w) for _, _, w, h in self._items]
        total = sum(h * w for h, w in sizes)
        flat = torch.empty(total * 4, dtype=torch.float32, device=_device())
        arr = (_lib.DtvRenderItem * len(self._items))()
        images, off = [], 0
        for i, (mesh, vm, w, h) in enumerate(self._items):
            arr[i].mesh = mesh._desc()
            cam = arr[i].camera
            for k in range(16):
                cam.view[k] = float(vm[k])
            cam.fovy, cam.znear, cam.zfar = self._fovy, self._znear, self._zfar
            cam.width, cam.height = w, h
            img = flat[off * 4:(off + h * w) * 4].view(h, w, 4)
            arr[i].out_image = img.data_ptr()
            images.append(img)
            off += h * w
        if getattr(self, "_autotune", (False,))[0]:
            self._maybe_tune(arr, len(self._items), [(m, w, h) for m, _, w, h in self._items])
        prm = self._params()
        _lib.check(self._lib.dtv_render_batch(self._view, arr, len(self._items), ctypes.byref(prm), _stream_ptr()),
                   "BatchRenderer.render")
        self._rendered_items = list(self._items)
        self._rendered_mesh = self._mesh
        self._uniform_size = len(set(sizes)) == 1
        if self._uniform_size:
            return flat.view(len(self._items), sizes[0][0], sizes[0][1], 4)
        return images

    def render_adjoint(self, image_adjoints, use_abs_grad=False):
        items = self._rendered_items
        if items is None:
            raise RuntimeError("Error in render_adjoint: render() must be called first.")
        if torch.is_tensor(image_adjoints):
            if image_adjoints.dim() != 4 or image_adjoints.shape[0] != len(items):
                raise RuntimeError("Error in render_adjoint: [N, H, W, 4] image adjoints expected.")
            image_adjoints = list(image_adjoints.unbind(0))
        if len(image_adjoints) != len(items):
            raise RuntimeError("Error in render_adjoint: one image adjoint per item expected.")
        arr = (_lib.DtvAdjointItem * len(items))()
        keep = []  # adjoint images stay referenced until the launch is enqueued
        for i, ((mesh, _, w, h), d) in enumerate(zip(items, image_adjoints)):
            if not torch.is_tensor(d) or d.device.type != "cuda" or d.dtype != torch.float32:
                raise RuntimeError("Error in render_adjoint: float32 device tensors expected.")
            if tuple(d.shape) != (h, w, 4):
                raise RuntimeError("Error in render_adjoint: image size differs from the viewport.")
            d = d.contiguous()
            if d.data_ptr() % 16:
                d = d.clone()
            keep.append(d)
            pos, col = mesh.get_vertex_positions(), mesh._color_tensor()
            if pos.grad is None:
                pos.grad = torch.zeros_like(pos)
            if col.grad is None:
                col.grad = torch.zeros_like(col)
            arr[i].mesh = mesh._desc()
            arr[i].dimage = d.data_ptr()
            arr[i].grad_positions = pos.grad.data_ptr()
            arr[i].grad_colors = col.grad.data_ptr()
        _lib.check(self._lib.dtv_render_adjoint_batch(self._view, arr, len(items), int(bool(use_abs_grad)),
                                                      self._adjoint_mode, _stream_ptr()),
                   "BatchRenderer.render_adjoint")

    def get_fragment_counts(self):
        """int32 [N, H, W] (uniform viewports) or a list of [H_i, W_i]: fragments composited per pixel."""
        items = self._rendered_items
        if items is None:
            raise RuntimeError("get_fragment_counts: render() must be called first.")
        total = sum(h * w for _, _, w, h in items)
        out = torch.empty(total, dtype=torch.int32, device=_device())
        _lib.check(self._lib.dtv_view_fragment_counts(self._view, ctypes.c_void_p(out.data_ptr()), _stream_ptr()),
                   "get_fragment_counts")
        if self._uniform_size:
            return out.view(len(items), items[0][3], items[0][2])
        res, off = [], 0
        for _, _, w, h in items:
            res.append(out[off:off + h * w].view(h, w))
            off += h * w
        return res

    def get_fragment_ids(self, max_per_pixel):
        raise RuntimeError("get_fragment_ids: single-view renders only (use Renderer).")


class TetRegularizer:
    """Tet-quality regulariser (TetRegularizer, src/Module/Module.cpp:663-693, bound at :983-995).

    ``compute_grad()`` accumulates (+=) the gradient of sum_tets (reg_lambda / softplus_beta)
    log(1 + exp(-softplus_beta q)) into the mesh positions' ``.grad``, q being the Parthasarathy et al. quality
    6 sqrt(2) V / l_rms^3 (Data/Shaders/Optimizer/TetRegularizer.glsl:108-166).  As in the reference, the
    S^{5/2} denominator of dq/dp (S = sum of squared edge lengths) is clamped below at 1e-5.
    """

    def __init__(self, tet_mesh, reg_lambda, softplus_beta):
        _ctx()
        self._lib = _lib.load()
        self.tet_mesh = tet_mesh
        self.reg_lambda = float(reg_lambda)
        self.softplus_beta = float(softplus_beta)

    def compute_grad(self):
        mesh = self.tet_mesh
        pos = mesh.get_vertex_positions()
        if pos.grad is None:
            pos.grad = torch.zeros_like(pos)
        desc = mesh._desc()
        _lib.check(self._lib.dtv_tet_regularizer_grad(ctypes.byref(desc), self.reg_lambda, self.softplus_beta,
                                                      ctypes.c_void_p(pos.grad.data_ptr()), _stream_ptr()),
                   "TetRegularizer.compute_grad")


from . import optim  # noqa: E402,F401
from .pyutils import DifferentiableRenderer, DifferentiableRenderingFunction  # noqa: E402,F401
