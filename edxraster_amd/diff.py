"""Differentiable rendering through the C ABI: a torch.autograd.Function over
edxr_render_batch (forward, float RGB out) and edxr_backward (hand-written
HIP adjoint with visibility held fixed).

``render_views`` renders one mesh from several views of equal size and
returns an (N, H, W, 3) float32 image (rows bottom-up, as GetBackBuffer).
Gradients flow to the per-vertex positions, normals and colours (CUDA
float32 tensors of length V) and to the light direction and the material
(ambient, kd, ks, shininess) (CPU float tensors of length 3 and 4).  The
forward's visibility (triangle IDs) is constant; barycentrics are
differentiated through the continuous projection of the winning triangle.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .renderer import Renderer, _check


def _ptr(t):
    return C.cast(C.c_void_p(t.data_ptr()), L.F)


PLANE_NAMES = ("px", "py", "pz", "nx", "ny", "nz", "cr", "cg", "cb")


def check_inputs(indices, planes):
    """Raises TypeError / ValueError unless the mesh tensors are what the
    kernels read through raw pointers: int32 indices (3 per triangle) and
    nine float32 planes of equal length, all on one CUDA device."""
    if indices.dtype != torch.int32:
        raise TypeError("indices must be int32 (got %s)" % indices.dtype)
    if indices.dim() != 1 or indices.numel() % 3:
        raise ValueError("indices must be a flat tensor of 3 entries per triangle")
    V = planes[0].numel()
    for name, t in zip(PLANE_NAMES, planes):
        if t.dtype != torch.float32:
            raise TypeError("%s must be float32 (got %s)" % (name, t.dtype))
        if t.dim() != 1 or t.numel() != V:
            raise ValueError("%s must be a flat tensor of %d values (got shape %s)" % (name, V, tuple(t.shape)))
    dev = planes[0].device
    for name, t in zip(("indices",) + PLANE_NAMES, (indices,) + tuple(planes)):
        if t.device.type != "cuda" or t.device != dev:
            raise ValueError("%s must be on the CUDA device of px (got %s)" % (name, t.device))


class _RenderViews(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rend, views, shader, indices, px, py, pz, nx, ny, nz, cr, cg, cb, light, material):
        check_inputs(indices, (px, py, pz, nx, ny, nz, cr, cg, cb))
        planes = [t.contiguous() for t in (px, py, pz, nx, ny, nz, cr, cg, cb)]
        idx = indices.contiguous()
        V, T = planes[0].numel(), idx.numel() // 3
        mesh = L.Mesh(*[_ptr(t) for t in planes], C.cast(C.c_void_p(idx.data_ptr()), L.U32), V, T, L.MEM_DEVICE)
        lt = [float(x) for x in light.detach().cpu().double()]
        mt = [float(x) for x in material.detach().cpu().double()]
        mat = L.Material()
        mat.ambient, mat.kd, mat.ks, mat.shininess = mt
        mat.albedo[:] = [1.0, 1.0, 1.0]
        mat.clamp_specular = 0
        jobs = [Renderer.make_job(mesh, v.width, v.height, v.mv, v.proj, v.raster, lt, shader, mat) for v in views]
        W, H = views[0].width, views[0].height
        torch.cuda.synchronize()
        L.lib().edxr_set_state(rend.ctx, L.STATE_FLOAT_COLOR, 1)
        rend.render_batch(jobs)
        out = torch.empty((len(views), H, W, 3), dtype=torch.float32, device=px.device)
        _check(rend.ctx, L.lib().edxr_copy_float_color(rend.ctx, C.c_void_p(out.data_ptr()), out.numel() * 4))
        ctx.rend = rend
        # edxr_backward differentiates the context's last render: remember
        # which one this image is.
        ctx.generation = L.lib().edxr_render_generation(rend.ctx)
        ctx.njobs = len(views)
        ctx.V = V
        ctx.light_device = light.device
        ctx.save_for_backward(px)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        (px,) = ctx.saved_tensors
        if L.lib().edxr_render_generation(ctx.rend.ctx) != ctx.generation:
            raise RuntimeError("the renderer has rendered again since this image was made; the backward pass "
                               "differentiates only the renderer's last render (use one Renderer per image "
                               "kept for backward)")
        dev = px.device
        g = grad_out.contiguous().float()
        gv = torch.zeros((9, ctx.V), dtype=torch.float32, device=dev)
        gl = torch.zeros(3, dtype=torch.float32, device=dev)
        gm = torch.zeros(4, dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        gr = L.Grads(*[_ptr(gv[k]) for k in range(9)], _ptr(gl), _ptr(gm))
        arr = (L.Grads * ctx.njobs)(*([gr] * ctx.njobs))
        _check(ctx.rend.ctx, L.lib().edxr_backward(ctx.rend.ctx, _ptr(g), arr, ctx.njobs))
        return (None, None, None, None, gv[0], gv[1], gv[2], gv[3], gv[4], gv[5], gv[6], gv[7], gv[8],
                gl.to(ctx.light_device), gm.to(ctx.light_device))


def render_views(rend, views, indices, px, py, pz, nx, ny, nz, cr, cg, cb, light, material,
                 shader=L.SHADER_VC_BLINN_PHONG):
    """(N, H, W, 3) float32 images of the mesh from `views` (configs.Workload
    cameras of one size); differentiable w.r.t. the tensors that require it."""
    return _RenderViews.apply(rend, views, shader, indices, px, py, pz, nx, ny, nz, cr, cg, cb, light, material)


def mesh_tensors(mesh, device="cuda"):
    """(indices int32, [px, py, pz, nx, ny, nz, cr, cg, cb] float32) on device."""
    idx = torch.from_numpy(mesh.indices.view(np.int32).copy()).to(device)
    return idx, [torch.from_numpy(a.copy()).to(device) for a in mesh.planes()]
