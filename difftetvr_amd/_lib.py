"""ctypes binding of the C ABI declared in include/difftetvr_amd.h.

The shared library is built in-tree by ``difftetvr_amd/build.py`` (``__graft_entry__.build()``).  There is no
fallback: if the library is missing or no GPU is visible, every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdifftetvr_amd.so")

DTV_OK = 0
DTV_ADJOINT_EXACT = 0
DTV_ADJOINT_REFERENCE = 1
DTV_SEGMENT_EUCLIDEAN = 0
DTV_SEGMENT_LINEAR_DEPTH = 1
DTV_RENDERER_TILE = 0
DTV_RENDERER_PROJECTION = 1
DTV_TILE_LAUNCH_GRID = 0
DTV_TILE_LAUNCH_QUEUE = 1
STAGES = ["count_pairs", "scan", "emit", "sort", "ranges", "forward", "adjoint"]

c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_float = ctypes.c_float
c_void_p = ctypes.c_void_p


class DtvMesh(ctypes.Structure):
    _fields_ = [("cells", c_void_p), ("positions", c_void_p), ("colors", c_void_p),
                ("num_cells", c_int64), ("num_vertices", c_int64), ("cell_colors", c_void_p)]


class DtvCamera(ctypes.Structure):
    _fields_ = [("view", c_float * 16), ("fovy", c_float), ("znear", c_float), ("zfar", c_float),
                ("width", c_int32), ("height", c_int32)]


class DtvRenderParams(ctypes.Structure):
    _fields_ = [("attenuation", c_float), ("clear_color", c_float * 4), ("use_early_ray_termination", c_int32),
                ("early_ray_out_alpha", c_float), ("tile_size", c_int32), ("checkpoint_interval", c_int32),
                ("tile_order", c_int32), ("segment_length", c_int32), ("renderer", c_int32),
                ("segments_per_pass", c_int32), ("tile_launch", c_int32)]


class DtvViewStats(ctypes.Structure):
    _fields_ = [("num_triangles", c_int64), ("num_pairs", c_int64), ("num_tiles", c_int64),
                ("sort_passes", c_int32), ("width", c_int32), ("height", c_int32), ("num_items", c_int32),
                ("num_pixels", c_int64), ("num_gradient_atomics", c_int64)]


class DtvRenderItem(ctypes.Structure):
    _fields_ = [("mesh", DtvMesh), ("camera", DtvCamera), ("out_image", c_void_p)]


class DtvAdjointItem(ctypes.Structure):
    _fields_ = [("mesh", DtvMesh), ("dimage", c_void_p), ("grad_positions", c_void_p), ("grad_colors", c_void_p)]


class DtvServerRequest(ctypes.Structure):
    _fields_ = [("mesh_id", c_int32), ("camera", DtvCamera), ("out_image", c_void_p), ("target_image", c_void_p),
                ("grad_positions", c_void_p), ("grad_colors", c_void_p)]


class DtvServerStats(ctypes.Structure):
    _fields_ = [("submitted", c_int64), ("completed", c_int64), ("batches", c_int64), ("max_batch_seen", c_int64)]


class DtvAdamGroup(ctypes.Structure):
    _fields_ = [("param", c_void_p), ("grad", c_void_p), ("exp_avg", c_void_p), ("exp_avg_sq", c_void_p),
                ("numel", c_int64), ("lr", ctypes.c_double), ("beta1", ctypes.c_double), ("beta2", ctypes.c_double),
                ("eps", ctypes.c_double),
                ("step", c_int64), ("clamp_nonneg", c_int32), ("zero_grad", c_int32)]


class DtvTuneResult(ctypes.Structure):
    _fields_ = [("tile_size", c_int32), ("segments_per_pass", c_int32), ("ms", c_float), ("ms_default", c_float),
                ("num_candidates", c_int32), ("cand_tile_size", c_int32 * 9), ("cand_segments_per_pass", c_int32 * 9),
                ("cand_ms", c_float * 9)]


# (name, restype, argtypes) of every exported entry point; tests check the library exports exactly these.
SIGNATURES = [
    ("dtv_get_version", ctypes.c_int, []),
    ("dtv_last_error", ctypes.c_char_p, []),
    ("dtv_ctx_create", ctypes.c_int, [c_int32, ctypes.POINTER(c_void_p)]),
    ("dtv_ctx_destroy", ctypes.c_int, [c_void_p]),
    ("dtv_ctx_get_device", ctypes.c_int, [c_void_p, ctypes.POINTER(c_int32)]),
    ("dtv_view_create", ctypes.c_int, [c_void_p, ctypes.POINTER(c_void_p)]),
    ("dtv_view_destroy", ctypes.c_int, [c_void_p]),
    ("dtv_camera_uniforms", ctypes.c_int, [ctypes.POINTER(DtvCamera), ctypes.POINTER(c_float)]),
    ("dtv_render", ctypes.c_int, [c_void_p, ctypes.POINTER(DtvMesh), ctypes.POINTER(DtvCamera),
                                  ctypes.POINTER(DtvRenderParams), c_void_p, c_void_p]),
    ("dtv_render_adjoint", ctypes.c_int, [c_void_p, ctypes.POINTER(DtvMesh), c_void_p, c_int32, c_int32, c_void_p,
                                          c_void_p, c_void_p]),
    ("dtv_render_noalloc", ctypes.c_int, [c_void_p, ctypes.POINTER(DtvMesh), ctypes.POINTER(DtvCamera),
                                          ctypes.POINTER(DtvRenderParams), c_void_p, c_void_p, c_void_p]),
    ("dtv_render_batch", ctypes.c_int, [c_void_p, ctypes.POINTER(DtvRenderItem), c_int32,
                                        ctypes.POINTER(DtvRenderParams), c_void_p]),
    ("dtv_render_adjoint_batch", ctypes.c_int, [c_void_p, ctypes.POINTER(DtvAdjointItem), c_int32, c_int32, c_int32,
                                                c_void_p]),
    ("dtv_autotune", ctypes.c_int, [c_void_p, ctypes.POINTER(DtvRenderItem), c_int32, ctypes.POINTER(DtvRenderParams),
                                    c_int32, c_int32, ctypes.POINTER(DtvTuneResult), c_void_p]),
    ("dtv_view_fragment_counts", ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    ("dtv_view_export_fragments", ctypes.c_int, [c_void_p, ctypes.POINTER(DtvMesh), c_int32, c_void_p, c_void_p]),
    ("dtv_view_get_stats", ctypes.c_int, [c_void_p, ctypes.POINTER(DtvViewStats)]),
    ("dtv_view_set_record_capacity", ctypes.c_int, [c_void_p, c_int64]),
    ("dtv_view_records_overflowed", ctypes.c_int, [c_void_p, ctypes.POINTER(c_int32), ctypes.POINTER(c_int64)]),
    ("dtv_view_set_timing", ctypes.c_int, [c_void_p, c_int32]),
    ("dtv_view_get_timings", ctypes.c_int, [c_void_p, ctypes.POINTER(c_float)]),
    ("dtv_adam_step", ctypes.c_int, [ctypes.POINTER(DtvAdamGroup), c_int32, c_void_p]),
    ("dtv_adam_step_device", ctypes.c_int, [ctypes.POINTER(DtvAdamGroup), c_int32, c_void_p, c_void_p]),
    ("dtv_tet_regularizer_grad", ctypes.c_int, [ctypes.POINTER(DtvMesh), c_float, c_float, c_void_p, c_void_p]),
    ("dtv_split_by_largest_gradient_magnitudes", ctypes.c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_int64,
                                                                c_void_p, c_int32, c_float, ctypes.POINTER(c_void_p)]),
    ("dtv_subdivide_vertices", ctypes.c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_int64, c_void_p, c_int64,
                                              ctypes.POINTER(c_void_p)]),
    ("dtv_split_result_sizes", ctypes.c_int, [c_void_p, ctypes.POINTER(c_int64), ctypes.POINTER(c_int64)]),
    ("dtv_split_result_copy", ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("dtv_split_result_destroy", ctypes.c_int, [c_void_p]),
    ("dtv_mse_loss_grad", ctypes.c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    ("dtv_mse_loss_grad_batch", ctypes.c_int, [c_void_p, c_void_p, c_int64, c_int32, c_void_p, c_void_p, c_void_p]),
    ("dtv_server_last_error", ctypes.c_char_p, []),
    ("dtv_server_create", ctypes.c_int, [c_void_p, c_int32, c_int32, c_int64, ctypes.POINTER(DtvRenderParams),
                                         ctypes.POINTER(c_void_p)]),
    ("dtv_server_destroy", ctypes.c_int, [c_void_p]),
    ("dtv_server_add_mesh", ctypes.c_int, [c_void_p, ctypes.POINTER(DtvMesh), ctypes.POINTER(c_int32)]),
    ("dtv_server_submit", ctypes.c_int, [c_void_p, ctypes.POINTER(DtvServerRequest), ctypes.POINTER(ctypes.c_uint64)]),
    ("dtv_server_wait", ctypes.c_int, [c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]),
    ("dtv_server_get_stats", ctypes.c_int, [c_void_p, ctypes.POINTER(DtvServerStats)]),
]

_lib = None


class DtvError(RuntimeError):
    pass


def load():
    """Loads the in-tree HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DtvError(f"difftetvr_amd HIP library not built: {LIB_PATH} is missing "
                       "(run __graft_entry__.build() or python difftetvr_amd/build.py)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status, what="", server=False):
    if status != DTV_OK:
        msg = load().dtv_server_last_error() if server else load().dtv_last_error()
        raise DtvError(f"{what}: {msg.decode() if msg else 'error'} (status {status})")
