"""PyTorch autograd integration — drop-in for rmesh_wgpu.autograd
(crates/rmesh-python/python/rmesh_wgpu/autograd.py:15-151).

``RMeshForward`` / ``RMeshModule`` keep the reference's signatures: forward
uploads the parameters and renders one view through the tiled pipeline
(``render_mode="tiled"``) or the interval-tiled one
(``render_mode="interval_tiled"``); backward hands dL/d(image) to the
matching hand-written adjoint and returns gradients for (vertices,
base_colors, densities, color_grads).  With parameters on the GPU the tiled
mode runs zero-copy: the renderer works on torch's current stream, parameters,
image, image gradient and parameter gradients move device to device (the
reference copies every one through host arrays, autograd.py:26-64).  The reference's interval branch passes
forward_interval_tiled's dict where an array is expected (autograd.py:40,49);
here the branch uses the dict's image, so it runs.
"""

from __future__ import annotations

import numpy as np
import torch
import torch.nn as nn

from .renderer import RMeshRenderer


def _np(t):
    return np.ascontiguousarray(t.detach().cpu().numpy().ravel().astype(np.float32))


class RMeshForward(torch.autograd.Function):
    """Custom autograd function delegating rendering to the HIP renderer."""

    @staticmethod
    def forward(ctx, renderer, cam_pos, vp, c2w_intrinsics, vertices, base_colors, densities, color_grads,
                render_mode="tiled"):
        ctx.render_mode = render_mode
        ctx.renderer = renderer
        ctx.device_path = vertices.is_cuda and render_mode != "interval_tiled"
        if ctx.device_path:
            renderer.set_stream(torch.cuda.current_stream(vertices.device).cuda_stream)
            params = [t.detach().reshape(-1).contiguous().float() for t in (vertices, base_colors, densities,
                                                                          color_grads)]
            renderer.update_params_device(*(t.data_ptr() for t in params))
            image = torch.empty((renderer.height, renderer.width, 4), dtype=torch.float32, device=vertices.device)
            renderer.forward_tiled_device(_np(cam_pos), _np(vp), _np(c2w_intrinsics), image.data_ptr())
            ctx.save_for_backward(vertices, base_colors, densities, color_grads)
            return image
        renderer.update_params(_np(vertices), _np(base_colors), _np(densities), _np(color_grads))
        if render_mode == "interval_tiled":
            image_np = renderer.forward_interval_tiled(_np(cam_pos), _np(vp), _np(c2w_intrinsics))["image"]
        else:
            image_np = renderer.forward_tiled(_np(cam_pos), _np(vp), _np(c2w_intrinsics))
        ctx.save_for_backward(vertices, base_colors, densities, color_grads)
        return torch.from_numpy(image_np.copy()).to(vertices.device)

    @staticmethod
    def backward(ctx, grad_output):
        if ctx.device_path:
            v, b, d, g = ctx.saved_tensors
            dl = grad_output.detach().contiguous().float()
            out = [torch.empty(t.numel(), dtype=torch.float32, device=t.device) for t in (v, d, b, g)]
            ctx.renderer.backward_device(dl.data_ptr(), *(t.data_ptr() for t in out))
            dv, dd, db, dg = out
            return (None, None, None, None, dv.view(v.shape), db.view(b.shape), dd.view(d.shape), dg.view(g.shape),
                    None)
        dl = np.ascontiguousarray(grad_output.detach().cpu().numpy().astype(np.float32))
        if ctx.render_mode == "interval_tiled":
            grads = ctx.renderer.backward_interval_tiled(dl)
        else:
            grads = ctx.renderer.backward(dl)
        device = ctx.saved_tensors[0].device
        t = lambda k: torch.from_numpy(grads[k].copy()).to(device)  # noqa: E731
        return (None, None, None, None, t("d_vertices"), t("d_base_colors"), t("d_densities"),
                t("d_color_grads"), None)


class RMeshModule(nn.Module):
    """nn.Module holding learnable parameters + the renderer (autograd.py:79-151)."""

    def __init__(self, vertices, indices, base_colors, densities, color_grads, circumdata, width, height, **kw):
        super().__init__()
        self.vertices = nn.Parameter(torch.tensor(vertices.ravel(), dtype=torch.float32))
        self.base_colors = nn.Parameter(torch.tensor(base_colors.ravel(), dtype=torch.float32))
        self.densities = nn.Parameter(torch.tensor(densities.ravel(), dtype=torch.float32))
        self.color_grads = nn.Parameter(torch.tensor(color_grads.ravel(), dtype=torch.float32))
        self._indices = indices.ravel().astype(np.uint32)
        self._circumdata = circumdata.ravel().astype(np.float32)
        self._width = width
        self._height = height
        self.renderer = RMeshRenderer(
            vertices.ravel().astype(np.float32), self._indices, base_colors.ravel().astype(np.float32),
            densities.ravel().astype(np.float32), color_grads.ravel().astype(np.float32), self._circumdata,
            width, height, **kw)

    def forward(self, cam_pos, vp, c2w_intrinsics, render_mode="tiled"):
        return RMeshForward.apply(self.renderer, cam_pos, vp, c2w_intrinsics, self.vertices, self.base_colors,
                                  self.densities, self.color_grads, render_mode)
