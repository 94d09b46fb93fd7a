"""A thin ``torch.autograd.Function`` over the C ABI's device entry points.

The topology and creases of a cage are refined once (``Subdivider``); after
that the refined points are a linear function of the cage points, so the
forward pass is ``ccd_RefinePoints`` and the backward pass is the adjoint
``ccd_RefinePointsAdjoint``, both launched on torch's current stream with
device pointers.  No CPU path exists: without the HIP library or a GPU the
calls raise.
"""
from __future__ import annotations

import torch

from . import capi
from .cage import Cage


class Subdivider:
    """Refines a cage's topology and creases once and then maps cage points
    (V0 x 3, float32, on the GPU) to the points of the finest level."""

    def __init__(self, cage: Cage, depth: int, creases: bool = True):
        self.cage = cage.contiguous()
        self.depth = depth
        self.steps = 0 if creases else capi.NOCREASES
        self.ds = capi.DeviceSubd(self.cage, depth)
        # the topology pass runs on the stream that is current now; the
        # points and adjoint kernels may later run on another torch stream,
        # so wait for it here
        self.ds.refine(capi.HALFEDGES | capi.CREASES, _stream())
        self.ds.synchronize()
        n = self.ds.n
        self.V0 = n["V"][0]
        self.VD = n["V"][depth]
        self.Vcum = n["Vcum"][depth]
        self.finest_offset = n["Vcum"][depth - 1]

    def __call__(self, cage_points: torch.Tensor) -> torch.Tensor:
        return RefinePoints.apply(cage_points, self)

    def cage_points(self, device: str = "cuda") -> torch.Tensor:
        return torch.as_tensor(self.cage.vertex_points, device=device)

    def close(self):
        self.ds.close()


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


class RefinePoints(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cage_points: torch.Tensor, sub: Subdivider) -> torch.Tensor:
        if not cage_points.is_cuda or cage_points.dtype != torch.float32:
            raise TypeError("cage points must be a float32 CUDA tensor")
        pts = cage_points.contiguous()
        if pts.shape != (sub.V0, 3):
            raise ValueError(f"expected ({sub.V0}, 3) cage points, got {tuple(pts.shape)}")
        out = torch.empty((sub.Vcum, 3), device=pts.device, dtype=torch.float32)
        sub.ds.refine_points(pts.data_ptr(), out.data_ptr(), sub.steps, _stream())
        ctx.sub = sub
        return out[sub.finest_offset:]

    @staticmethod
    def backward(ctx, grad_finest: torch.Tensor):
        sub = ctx.sub
        g = grad_finest.contiguous().to(torch.float32)
        grad_cage = torch.empty((sub.V0, 3), device=g.device, dtype=torch.float32)
        sub.ds.refine_points_adjoint(g.data_ptr(), grad_cage.data_ptr(), sub.steps, _stream())
        return grad_cage, None
