"""Differentiable closest hit: a torch.autograd.Function over the C ABI.

Forward: rtbvh_gpu_intersect / rtbvh_gpu_intersect_mbvh (lowest-(t, id)
rule).  Backward: rtbvh_gpu_adjoint, the analytic adjoint of
o + t d = (1-u-v) v0 + u v1 + v v2 with the hit triangle fixed by the
forward pass (no counterpart in the reference, whose triangle test,
src/builders/spatial_sah.rs:132-163, is forward only).
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib as L
from .api import RAY_FLOATS, TraceTree, _ptr, _require_device, _stream, hits_prim, rays_tensor


def adjoint(vertices: torch.Tensor, rays: torch.Tensor, hits: torch.Tensor,
            grad_tuv: torch.Tensor, grad_vertices: bool = True, grad_rays: bool = True):
    """dL/d(vertices) (N, 9), dL/d(origin) (R, 3), dL/d(direction) (R, 3) from
    dL/d(t, u, v) (R, 3) of the hits (R, 4) of `rays` (R, 8)."""
    _require_device(vertices, "vertices")
    _require_device(rays, "rays")
    _require_device(hits, "hits")
    _require_device(grad_tuv, "grad_tuv")
    verts = vertices.reshape(-1)
    n = verts.numel() // 9
    r = rays.shape[0]
    if rays.shape[1] != RAY_FLOATS or hits.shape != (r, 4) or grad_tuv.shape != (r, 3):
        raise ValueError("rays (R, 8), hits (R, 4) and grad_tuv (R, 3) must agree")
    gv = torch.zeros((n, 9), dtype=torch.float32, device=vertices.device) if grad_vertices else None
    go = torch.empty((r, 3), dtype=torch.float32, device=rays.device) if grad_rays else None
    gd = torch.empty((r, 3), dtype=torch.float32, device=rays.device) if grad_rays else None
    L.check(L.lib().rtbvh_gpu_adjoint(_ptr(verts), n, _ptr(rays), _ptr(hits), r, _ptr(grad_tuv),
                                      _ptr(gv), _ptr(go), _ptr(gd), _stream(rays)),
            "rtbvh_gpu_adjoint")
    return gv, go, gd


class TraceTUV(torch.autograd.Function):
    """(vertices, origins, directions) -> (t, u, v, prim) of the closest hit."""

    @staticmethod
    def forward(ctx, vertices, origins, directions, tree, t_min, t_max):
        verts = vertices.detach().contiguous()
        rays = rays_tensor(origins.detach().contiguous(), directions.detach().contiguous(),
                           t_min, t_max)
        if isinstance(tree, TraceTree):
            # a traversal tree traces its own snapshot of the triangles: the
            # gradients are only those of `vertices` if the snapshot is theirs
            if not tree.built_from(vertices):
                raise ValueError("the TraceTree was built from other vertices, or these were "
                                 "modified since: rebuild it (Bvh.trace_tree) from the vertices "
                                 "passed here")
            hits = tree.intersect(rays)
        else:
            hits = tree.intersect(verts.reshape(-1, 9), rays)
        prim = hits_prim(hits)
        ctx.save_for_backward(verts, rays, hits)
        ctx.mark_non_differentiable(prim)
        return hits[:, 0].clone(), hits[:, 1].clone(), hits[:, 2].clone(), prim

    @staticmethod
    def backward(ctx, gt, gu, gv, _gprim):
        verts, rays, hits = ctx.saved_tensors
        r = rays.shape[0]
        z = torch.zeros(r, dtype=torch.float32, device=rays.device)
        g = torch.stack([gt if gt is not None else z, gu if gu is not None else z,
                         gv if gv is not None else z], dim=1).contiguous()
        need_v = ctx.needs_input_grad[0]
        need_r = ctx.needs_input_grad[1] or ctx.needs_input_grad[2]
        d_v, d_o, d_d = adjoint(verts.reshape(-1, 9), rays, hits, g, need_v, need_r)
        d_v = d_v.view(verts.shape) if d_v is not None else None
        return (d_v, d_o if ctx.needs_input_grad[1] else None,
                d_d if ctx.needs_input_grad[2] else None, None, None, None)


def trace_tuv(tree, vertices: torch.Tensor, origins: torch.Tensor, directions: torch.Tensor,
              t_min: float = 1e-4, t_max: float = 1e34):
    """Closest hit (t, u, v, prim id) of rays against `tree` (a Bvh, Mbvh or
    TraceTree over `vertices`), differentiable in vertices, origins and
    directions.
    Misses return t = t_max, u = v = 0, prim = -1 and zero gradients."""
    return TraceTUV.apply(vertices, origins, directions, tree, t_min, t_max)


__all__ = ["adjoint", "trace_tuv", "TraceTUV", "ctypes"]
