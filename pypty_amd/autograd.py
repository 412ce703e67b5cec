"""torch.autograd.Function around the fused forward/adjoint pass.

``multislice_loss(obj, probe, corrections, ctx)`` returns the scalar loss
(FP64) of one full pass over ``ctx.sequence``; its backward hands PyTorch the
object, probe and sub-pixel-position gradients that the same launch
computed.  The forward and adjoint run together in ``pty_loss_and_direction``
(the adjoint needs the forward's stash), so the forward already produces the
gradients; backward only scales them.

Gradient convention: the kernels return PyPty's Wirtinger derivative
g = dL/dz* for complex inputs (pypty/utils.py:260-276).  PyTorch's complex
autograd expects 2 dL/dz* (so that z - lr * z.grad descends), hence the
factor 2 for object and probe.  Positions are real: dL/dd as is.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import lbfgs
from .objective import PassConfig, PtychoPass


@dataclass
class PassContext:
    """The fixed (non-differentiated) inputs of a pass."""

    cfg: PassConfig
    positions: torch.Tensor  # (P, 2) int32, device
    measured: torch.Tensor   # (P, N, N) float32, device
    sequence: torch.Tensor   # (n_seq,) int32, device


class _MultisliceLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, obj, probe, corrections, pctx: PassContext):
        need_obj, need_probe, need_pos = ctx.needs_input_grad[0], ctx.needs_input_grad[1], ctx.needs_input_grad[2]
        c = pctx.cfg
        shift = bool(torch.any(corrections != 0).item()) or need_pos
        cfg = PassConfig(**{**c.__dict__, "step_obj": need_obj, "step_probe": need_probe, "step_pos": need_pos,
                            "shift_active": shift})
        eng = PtychoPass(cfg, obj.device)
        obj_c = obj.detach().contiguous()
        probe_mnn = probe.detach().permute(2, 0, 1).contiguous()
        og = torch.empty_like(obj_c) if need_obj else None
        pg = torch.empty_like(probe_mnn) if need_probe else None
        posg = torch.zeros_like(corrections) if need_pos else None
        loss_sse = eng.run(obj_c, probe_mnn, pctx.positions, corrections.detach().contiguous(), pctx.measured,
                           pctx.sequence, obj_grad=og, probe_grad=pg, pos_grad=posg, sync=False)
        ctx.save_for_backward(*(t if t is not None else torch.empty(0) for t in (og, pg, posg)))
        ctx.flags = (need_obj, need_probe, need_pos)
        return loss_sse[0].clone()

    @staticmethod
    def backward(ctx, grad_out):
        og, pg, posg = ctx.saved_tensors
        need_obj, need_probe, need_pos = ctx.flags
        s = float(grad_out)
        g_obj = lbfgs.lincomb(og, 2.0 * s) if need_obj else None
        g_probe = lbfgs.lincomb(pg, 2.0 * s).permute(1, 2, 0) if need_probe else None
        g_pos = lbfgs.lincomb(posg, s) if need_pos else None
        return g_obj, g_probe, g_pos, None


def multislice_loss(obj: torch.Tensor, probe: torch.Tensor, corrections: torch.Tensor, pctx: PassContext):
    """Loss of the far-field multislice model (reference layouts: obj (Ny, Nx, S, K), probe (N, N, M),
    corrections (P, 2)); differentiable in all three."""
    return _MultisliceLoss.apply(obj, probe, corrections, pctx)
