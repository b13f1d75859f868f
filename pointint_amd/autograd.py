"""torch.autograd.Function over the volume matrix and its hand-written adjoint.

    V = intersection_volume_matrix(geoms, kgrid, verts_0, ..., verts_{n-1})

takes each object's vertex positions (TriangleMeshDoF: a (nv_i, 3) tensor,
FP32 or FP64) and returns the (n, n) FP64 volume matrix; the backward pass is
the library's VJP kernel (no dense dA/dvertex).  PyTorch is only the autograd
front end here: the form factors, the Gram and the adjoint run in
libpointint_hip.so.

With CUDA (ROCm) tensors on the library's device the vertices, V, dL/dV and
the vertex gradients stay in device memory (pii_volume_matrix_device /
pii_volume_matrix_vjp_device, ordered after torch's current stream); CPU
tensors go through the host entries.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from . import core


class _Meta:
    def __init__(self, geoms, kgrid, gram):
        self.geoms = list(geoms)
        self.kgrid = kgrid
        self.gram = gram


def _host_dofs(verts):
    return [v.detach().to("cpu", torch.float64).contiguous().numpy().reshape(-1) for v in verts]


def _device_dofs(verts):
    # FP64 copies on the same device (a device-side conversion for FP32 input)
    return [v.detach().to(torch.float64).contiguous().reshape(-1) for v in verts]


class VolumeMatrixFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, meta: _Meta, *verts):
        n = len(verts)
        kinds = ["trimesh"] * n
        ctx.meta = meta
        ctx.shapes = [(v.shape, v.dtype, v.device) for v in verts]
        ctx.on_device = all(v.is_cuda for v in verts)
        if ctx.on_device:
            xs = _device_dofs(verts)
            V = torch.empty((n, n), dtype=torch.float64, device=verts[0].device)
            core.compute_volume_matrix_device(V, meta.geoms, meta.kgrid, xs, kinds, ff="fast",
                                              gram=meta.gram)
            ctx.dofs = xs
            return V
        dofs = _host_dofs(verts)
        V = core.compute_volume_matrix(meta.geoms, meta.kgrid, dofs, kinds, ff="fast",
                                       gram=meta.gram)
        ctx.dofs = dofs
        return torch.from_numpy(V).to(verts[0].device)

    @staticmethod
    def backward(ctx, gV):
        kinds = ["trimesh"] * len(ctx.dofs)
        if ctx.on_device:
            G = gV.detach().to(device=ctx.dofs[0].device, dtype=torch.float64).contiguous()
            grads = [torch.empty_like(x) for x in ctx.dofs]
            core.volume_matrix_vjp_device(grads, ctx.meta.geoms, ctx.meta.kgrid, G, ctx.dofs,
                                          kinds)
            out = [g.reshape(shape).to(dtype=dt) for g, (shape, dt, _) in zip(grads, ctx.shapes)]
            return (None, *out)
        G = gV.detach().to("cpu", torch.float64).numpy()
        grads, _ = core.volume_matrix_vjp(ctx.meta.geoms, ctx.meta.kgrid, G, ctx.dofs, kinds,
                                          return_volume=False)
        out = [torch.from_numpy(g.reshape(shape)).to(device=dev, dtype=dt)
               for g, (shape, dt, dev) in zip(grads, ctx.shapes)]
        return (None, *out)


def intersection_volume_matrix(geoms: Sequence[core.Geometry], kgrid: core.KGrid, verts,
                               gram: str = "fp16x3") -> torch.Tensor:
    """Differentiable V(verts) for triangle meshes with vertex DoFs."""
    return VolumeMatrixFunction.apply(_Meta(geoms, kgrid, gram), *verts)


def pairwise_overlap_loss(V: torch.Tensor) -> torch.Tensor:
    """sum_{i<j} V_ij."""
    return torch.triu(V, diagonal=1).sum()
