"""torch.autograd bridge: the GPU forward (gz, TMI) of a ForwardContext as a function of
the vertex positions, with the hand-written adjoint kernel as its backward.

    gz, tmi = polyhedral_forward(vertices, ctx)      # vertices: (n_v, 3) float64 on cuda
    (w_gz @ gz + w_tmi @ tmi).backward()             # vertices.grad = adjoint

and the Jacobian-vector product kernel as its forward-mode derivative
(torch.autograd.forward_ad: a dual vertex tensor gives the tangents of gz and TMI).

The library runs on its own HIP stream; tensors are handed over after the current torch
stream has finished with them and handed back after the library stream is done.
"""
from __future__ import annotations

import torch

from .lithofwd import ForwardContext


class PolyhedralForward(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vertices: torch.Tensor, fctx: ForwardContext):
        if not vertices.is_cuda or vertices.dtype != torch.float64:
            raise TypeError("vertices must be a float64 CUDA tensor of shape (n_v, 3)")
        if vertices.shape != (fctx.n_v, 3):
            raise ValueError(f"expected vertices of shape ({fctx.n_v}, 3)")
        v = vertices.detach().contiguous()
        gz = torch.zeros(fctx.n_s, dtype=torch.float64, device=vertices.device)
        tm = torch.zeros(fctx.n_s, dtype=torch.float64, device=vertices.device)
        torch.cuda.current_stream().synchronize()
        fctx.set_vertices_device(v.data_ptr())
        fctx.forward_device(gz.data_ptr() if "gz" in fctx.want else 0,
                            tm.data_ptr() if "tmi" in fctx.want else 0)
        fctx.sync()
        ctx.fctx = fctx
        ctx.save_for_backward(v)
        ctx.save_for_forward(v)
        return gz, tm

    @staticmethod
    def backward(ctx, g_gz, g_tmi):
        fctx = ctx.fctx
        (v,) = ctx.saved_tensors
        dev = v.device
        wg = (g_gz if g_gz is not None else torch.zeros(fctx.n_s, device=dev)).to(torch.float64).contiguous()
        wt = (g_tmi if g_tmi is not None else torch.zeros(fctx.n_s, device=dev)).to(torch.float64).contiguous()
        grad = torch.zeros(fctx.n_v * 3, dtype=torch.float64, device=dev)
        torch.cuda.current_stream().synchronize()
        # the adjoint is taken at the geometry of this forward
        fctx.set_vertices_device(v.data_ptr())
        fctx.adjoint_device(wg.data_ptr(), wt.data_ptr(), grad.data_ptr())
        fctx.sync()
        return grad.view(fctx.n_v, 3), None

    @staticmethod
    def jvp(ctx, v_t, _):
        fctx = ctx.fctx
        (v,) = ctx.saved_tensors
        dev = v.device
        vt = v_t.to(torch.float64).contiguous()
        gz_t = torch.zeros(fctx.n_s, dtype=torch.float64, device=dev)
        tm_t = torch.zeros(fctx.n_s, dtype=torch.float64, device=dev)
        torch.cuda.current_stream().synchronize()
        # the tangent is taken at the geometry of this forward
        fctx.set_vertices_device(v.data_ptr())
        fctx.jvp_device(vt.data_ptr(), gz_t.data_ptr(), tm_t.data_ptr())
        fctx.sync()
        return gz_t, tm_t


def polyhedral_forward(vertices: torch.Tensor, fctx: ForwardContext):
    return PolyhedralForward.apply(vertices, fctx)
