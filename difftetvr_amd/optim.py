"""Fused Adam for the mesh parameters: one HIP launch updates every parameter group.

``FusedAdam`` is a ``torch.optim.Optimizer`` with torch.optim.Adam's update rule and defaults (the optimiser
pytests/train.py:318,351 uses) plus two per-group options: ``clamp_nonneg`` applies train.py:352-356's colour
clamp ``c -= min(c, 0)`` inside the same kernel, and ``zero_grad_after`` zeroes the consumed gradient (the
next step's ``optimizer.zero_grad(set_to_none=False)``, train.py:342).
"""
import ctypes

import torch

from . import _lib


class FusedAdam(torch.optim.Optimizer):
    """``device_step=True`` keeps the step count in device memory (``dtv_adam_step_device``), so that ``step()``
    can be captured in a CUDA/HIP graph and replayed: every replay advances the bias corrections."""

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, clamp_nonneg=False, zero_grad_after=False,
                 device_step=False):
        defaults = dict(lr=lr, betas=betas, eps=eps, clamp_nonneg=clamp_nonneg, zero_grad_after=zero_grad_after)
        super().__init__(params, defaults)
        self._device_step = bool(device_step)
        self._step_state = []  # one device counter per launch of up to 8 groups

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        groups = []
        for group in self.param_groups:
            for p in group["params"]:
                if p.grad is None:
                    continue
                if not (p.is_cuda and p.dtype == torch.float32 and p.is_contiguous() and p.grad.is_contiguous()):
                    raise RuntimeError("FusedAdam: parameters must be contiguous float32 GPU tensors.")
                st = self.state[p]
                if len(st) == 0:
                    st["step"] = 0
                    st["exp_avg"] = torch.zeros_like(p)
                    st["exp_avg_sq"] = torch.zeros_like(p)
                st["step"] += 1
                g = _lib.DtvAdamGroup()
                g.param, g.grad = p.data_ptr(), p.grad.data_ptr()
                g.exp_avg, g.exp_avg_sq = st["exp_avg"].data_ptr(), st["exp_avg_sq"].data_ptr()
                g.numel = p.numel()
                g.lr = group["lr"]
                g.beta1, g.beta2 = group["betas"]
                g.eps = group["eps"]
                g.step = st["step"]
                g.clamp_nonneg = int(bool(group["clamp_nonneg"]))
                g.zero_grad = int(bool(group["zero_grad_after"]))
                groups.append(g)
        lib = _lib.load()
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        for k, i in enumerate(range(0, len(groups), 8)):
            chunk = groups[i:i + 8]
            arr = (_lib.DtvAdamGroup * len(chunk))(*chunk)
            if self._device_step:
                if len(self._step_state) <= k:
                    self._step_state.append(torch.zeros(1, dtype=torch.int64, device=torch.device("cuda")))
                _lib.check(lib.dtv_adam_step_device(arr, len(chunk), ctypes.c_void_p(self._step_state[k].data_ptr()),
                                                    stream), "dtv_adam_step_device")
            else:
                _lib.check(lib.dtv_adam_step(arr, len(chunk), stream), "dtv_adam_step")
        return loss


def mse_loss_grad(image, target):
    """Returns (loss tensor [], dL/dimage) for torch.nn.MSELoss(image, target) using one HIP launch."""
    image = image.contiguous()
    target = target.contiguous()
    dimg = torch.empty_like(image)
    loss = torch.empty((), dtype=torch.float32, device=image.device)
    _lib.check(_lib.load().dtv_mse_loss_grad(ctypes.c_void_p(image.data_ptr()), ctypes.c_void_p(target.data_ptr()),
                                             image.numel(), ctypes.c_void_p(dimg.data_ptr()),
                                             ctypes.c_void_p(loss.data_ptr()),
                                             ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
               "dtv_mse_loss_grad")
    return loss, dimg


def mse_loss_grad_views(images, targets):
    """Sum over views of torch.nn.MSELoss(images[i], targets[i]) for [N, ...] stacks, and its image adjoint,
    in one HIP launch (the multi-view loss of a batched step)."""
    images = images.contiguous()
    targets = targets.contiguous()
    if images.shape != targets.shape or images.dim() < 2:
        raise RuntimeError("mse_loss_grad_views: [N, ...] images and targets of one shape expected.")
    dimg = torch.empty_like(images)
    loss = torch.empty((), dtype=torch.float32, device=images.device)
    n = images.shape[0]
    _lib.check(_lib.load().dtv_mse_loss_grad_batch(ctypes.c_void_p(images.data_ptr()),
                                                   ctypes.c_void_p(targets.data_ptr()), images.numel() // n, n,
                                                   ctypes.c_void_p(dimg.data_ptr()), ctypes.c_void_p(loss.data_ptr()),
                                                   ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
               "dtv_mse_loss_grad_batch")
    return loss, dimg
