"""Autograd glue, the counterpart of pytests/pyutils.py:31-79.

``DifferentiableRenderingFunction.forward`` renders; ``backward`` runs ``Renderer.render_adjoint``, which
accumulates the gradients straight into the ``.grad`` of the mesh's position/colour tensors, and therefore
returns no gradients itself (pyutils.py:41-56).
"""
import torch

from . import ColorStorage


class DifferentiableRenderingFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, renderer, tet_regularizer, use_abs_grad, vertex_positions, colors):
        image = renderer.render()
        ctx.renderer = renderer
        ctx.tet_regularizer = tet_regularizer
        ctx.use_abs_grad = use_abs_grad
        return image

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, image_adj):
        if ctx.tet_regularizer is not None:
            ctx.tet_regularizer.compute_grad()
        ctx.renderer.render_adjoint(image_adj, ctx.use_abs_grad)
        del ctx.renderer
        del ctx.tet_regularizer
        del ctx.use_abs_grad
        return None, None, None, None, None


class DifferentiableRenderer(torch.nn.Module):
    def __init__(self, renderer, use_tet_regularizer=False, tet_reg_lambda=0.0, tet_reg_softplus_beta=1.0):
        super().__init__()
        self.renderer = renderer
        self.use_abs_grad = False
        self.tet_regularizer = None
        if use_tet_regularizer:
            from . import TetRegularizer
            self.tet_regularizer = TetRegularizer(renderer.get_tet_mesh(), tet_reg_lambda, tet_reg_softplus_beta)

    def set_use_abs_grad(self, use_abs_grad):
        self.use_abs_grad = use_abs_grad

    def forward(self):
        tet_mesh = self.renderer.get_tet_mesh()
        if tet_mesh.get_color_storage() == ColorStorage.PER_VERTEX:
            colors = tet_mesh.get_vertex_colors()
        else:
            colors = tet_mesh.get_cell_colors()
        return DifferentiableRenderingFunction.apply(self.renderer, self.tet_regularizer, self.use_abs_grad,
                                                     tet_mesh.get_vertex_positions(), colors)
