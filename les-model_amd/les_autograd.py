"""torch.autograd binding of the differentiated LES step map.

``LesRollout.apply(u0, v0, w0, les, prm)`` advances the initial velocity
(flat float64 CUDA tensors of nx*ny*nz cells, index i*ny*nz + j*nz + k) by
``prm.nsteps`` fixed-dt steps on the GPU (les_rollout) and returns the final
velocity; its backward runs the hand-written adjoint kernels (les_adjoint),
recomputing the steps from checkpoints kept every ``prm.checkpoint_every``
steps.  The context ``les`` (lesgpu.LesGpu) must hold the mask and the
stencil built from it.  The step map and its restrictions are those of the
C ABI (include/les_gpu.h, "differentiation").

There is no counterpart in the reference; the forward is its step
(src/main.cpp:115-144) at a fixed dt with each pressure solve started from 0.
"""
from __future__ import annotations

import torch

from lesgpu import LesGpu, AdjParams


def _handover():
    # the library works on its own stream: torch's pending work on the
    # tensors it is handed must be complete first (the library synchronises
    # its own stream before returning)
    torch.cuda.current_stream().synchronize()


class LesRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, v0, w0, les: LesGpu, prm: AdjParams):
        for t in (u0, v0, w0):
            if not (t.is_cuda and t.dtype == torch.float64 and t.numel() == les.N):
                raise ValueError("expected float64 CUDA tensors of nx*ny*nz cells")
        u0, v0, w0 = (t.detach().contiguous() for t in (u0, v0, w0))
        _handover()
        les.set_field("u", u0)
        les.set_field("v", v0)
        les.set_field("w", w0)
        stats = les.rollout(prm)
        out = [torch.empty_like(u0) for _ in range(3)]
        for name, t in zip("uvw", out):
            les.get_field(name, t)
        les.synchronize()
        ctx.les, ctx.prm = les, prm
        ctx.save_for_backward(u0, v0, w0)
        # the context keeps one rollout's checkpoints: remember whose they are
        les._rollout_owner = ctx.owner = object()
        ctx.stats = stats
        return tuple(out)

    @staticmethod
    def backward(ctx, gu, gv, gw):
        les, prm = ctx.les, ctx.prm
        if getattr(les, "_rollout_owner", None) is not ctx.owner:
            # another rollout ran since: recompute this one's checkpoints
            u0, v0, w0 = ctx.saved_tensors
            les.set_field("u", u0)
            les.set_field("v", v0)
            les.set_field("w", w0)
            les.rollout(prm)
            les._rollout_owner = ctx.owner
        g = []
        for t, ref in zip((gu, gv, gw), ctx.saved_tensors):
            g.append(torch.zeros_like(ref) if t is None else t.detach().contiguous().clone())
        _handover()
        les.adjoint(prm, *g)
        les.synchronize()
        return g[0], g[1], g[2], None, None


def rollout(u0, v0, w0, les: LesGpu, prm: AdjParams):
    """Differentiable fixed-dt rollout (see LesRollout)."""
    return LesRollout.apply(u0, v0, w0, les, prm)
