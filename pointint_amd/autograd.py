"""torch.autograd.Function over the volume matrix and its hand-written adjoint.

    V = intersection_volume_matrix(geoms, kgrid, verts_0, ..., verts_{n-1})

takes each object's vertex positions (TriangleMeshDoF: a (nv_i, 3) tensor on
any device, FP32 or FP64) and returns the (n, n) FP64 volume matrix; the
backward pass is the library's VJP kernel (no dense dA/dvertex).  PyTorch is
only the autograd front end here: the form factors, the Gram and the adjoint
run in libpointint_hip.so.
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


class VolumeMatrixFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, meta: _Meta, *verts):
        dofs = _host_dofs(verts)
        V = core.compute_volume_matrix(meta.geoms, meta.kgrid, dofs, ["trimesh"] * len(dofs),
                                       ff="fast", gram=meta.gram)
        ctx.meta = meta
        ctx.dofs = dofs
        ctx.shapes = [(v.shape, v.dtype, v.device) for v in verts]
        return torch.from_numpy(V).to(verts[0].device)

    @staticmethod
    def backward(ctx, gV):
        G = gV.detach().to("cpu", torch.float64).numpy()
        grads, _ = core.volume_matrix_vjp(ctx.meta.geoms, ctx.meta.kgrid, G, ctx.dofs,
                                          ["trimesh"] * len(ctx.dofs), return_volume=False)
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
