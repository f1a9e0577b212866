"""Render server: resident meshes, continuous batching of render requests (``dtv_server_*`` in the C ABI).

Worker threads inside the library each own a view and a HIP stream; a free worker takes the pending requests in
arrival order (up to ``max_batch``, within ``memory_budget`` bytes of estimated transient device memory for all
batches in flight) and renders them as one batch, so requests arriving while others are on the GPU are admitted
as workers and memory free up.  A request with a target image is also differentiated (MSE loss, adjoint into the
mesh's ``.grad``).  Each result equals the request rendered alone.  The reference has no serving mode.
"""
import ctypes
import math

import numpy as np
import torch

from . import _ctx, _device, _lib


class Ticket:
    """A submitted request: ``image`` is valid after ``Server.wait``."""

    def __init__(self, ticket, image, keep):
        self.ticket = ticket
        self.image = image
        self._keep = keep


class Server:
    def __init__(self, num_workers=2, max_batch=16, memory_budget=0, attenuation=1.0, clear_color=(0, 0, 0, 0),
                 use_early_ray_termination=False, early_ray_out_alpha=1.0 - 1e-4, fovy=math.pi / 4.0,
                 znear=0.01, zfar=100.0):
        ctx = _ctx()
        self._lib = _lib.load()
        p = _lib.DtvRenderParams()
        p.attenuation = float(attenuation)
        for i in range(4):
            p.clear_color[i] = float(clear_color[i])
        p.use_early_ray_termination = int(bool(use_early_ray_termination))
        p.early_ray_out_alpha = float(early_ray_out_alpha)
        p.tile_size = 8
        p.checkpoint_interval = 64
        h = ctypes.c_void_p()
        _lib.check(self._lib.dtv_server_create(ctx, int(num_workers), int(max_batch), int(memory_budget),
                                               ctypes.byref(p), ctypes.byref(h)), "dtv_server_create", server=True)
        self._h = h
        self._meshes = []
        self._fovy, self._znear, self._zfar = float(fovy), float(znear), float(zfar)

    def close(self):
        if getattr(self, "_h", None) is not None:
            self._lib.dtv_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_mesh(self, tet_mesh):
        """Registers a TetMesh (kept resident; it must stay alive and unchanged while requests use it)."""
        desc = tet_mesh._desc()
        mid = ctypes.c_int32()
        _lib.check(self._lib.dtv_server_add_mesh(self._h, ctypes.byref(desc), ctypes.byref(mid)),
                   "dtv_server_add_mesh", server=True)
        self._meshes.append(tet_mesh)
        return mid.value

    def submit(self, mesh_id, view_matrix, width, height, target=None, out=None):
        """Queues a render of mesh ``mesh_id`` (column-major view matrix, viewport width x height).  With a
        ``target`` image the request also runs the MSE loss and the adjoint into the mesh's ``.grad``.  When
        ``out`` is not given it is allocated here and the current stream is synchronised first, so that no
        pending work on that stream still uses the memory the server is about to write."""
        if out is None:
            out = torch.empty((int(height), int(width), 4), dtype=torch.float32, device=_device())
            torch.cuda.current_stream().synchronize()
        req = _lib.DtvServerRequest()
        req.mesh_id = int(mesh_id)
        vm = np.asarray(view_matrix, np.float32).reshape(-1)
        for i in range(16):
            req.camera.view[i] = float(vm[i])
        req.camera.fovy, req.camera.znear, req.camera.zfar = self._fovy, self._znear, self._zfar
        req.camera.width, req.camera.height = int(width), int(height)
        req.out_image = out.data_ptr()
        keep = [out]
        if target is not None:
            mesh = self._meshes[int(mesh_id)]
            pos, col = mesh.get_vertex_positions(), mesh._color_tensor()
            if pos.grad is None:
                pos.grad = torch.zeros_like(pos)
            if col.grad is None:
                col.grad = torch.zeros_like(col)
            target = target.contiguous()
            req.target_image = target.data_ptr()
            req.grad_positions = pos.grad.data_ptr()
            req.grad_colors = col.grad.data_ptr()
            keep.append(target)
        t = ctypes.c_uint64()
        _lib.check(self._lib.dtv_server_submit(self._h, ctypes.byref(req), ctypes.byref(t)), "dtv_server_submit",
                   server=True)
        return Ticket(t.value, out, keep)

    def wait(self, ticket):
        """Blocks until the request completed; returns (image, latency_ms, queue_ms)."""
        lat, q = ctypes.c_double(), ctypes.c_double()
        _lib.check(self._lib.dtv_server_wait(self._h, ctypes.c_uint64(ticket.ticket), ctypes.byref(lat),
                                             ctypes.byref(q)), "dtv_server_wait", server=True)
        return ticket.image, lat.value, q.value

    def stats(self):
        st = _lib.DtvServerStats()
        _lib.check(self._lib.dtv_server_get_stats(self._h, ctypes.byref(st)), "dtv_server_get_stats", server=True)
        return {"submitted": st.submitted, "completed": st.completed, "batches": st.batches,
                "max_batch_seen": st.max_batch_seen}
