"""Forward/backward operator pair of the window cost in the autograd.Function
style.  The forward is the HIP cost kernel (patch_eval_color_grad_weighted,
pmhstereo.cl:170-228, over a plane field); the backward is the hand-written
HIP adjoint (windows recomputed, nothing stored by the forward).

    cost = WindowCost.apply(states, I1, G1, I2, G2, P)      # (H, W)
    cost = window_cost(states, left, right, P)              # gradients through gradxy too
"""
from __future__ import annotations

import torch

from . import ops


class WindowCost(torch.autograd.Function):
    @staticmethod
    def forward(ctx, states, I1, G1, I2, G2, P: int):
        states = states.contiguous()
        ctx.save_for_backward(states, I1, G1, I2, G2)
        ctx.P = int(P)
        return ops.cost(I1, G1, I2, G2, states, ctx.P)

    @staticmethod
    def backward(ctx, gout):
        states, I1, G1, I2, G2 = ctx.saved_tensors
        images = any(ctx.needs_input_grad[1:5])
        dst, dI1, dG1, dI2, dG2 = ops.cost_backward(I1, G1, I2, G2, states, gout.contiguous(),
                                                    ctx.P, images=images)
        need = ctx.needs_input_grad
        return (dst if need[0] else None, dI1 if need[1] else None, dG1 if need[2] else None,
                dI2 if need[3] else None, dG2 if need[4] else None, None)


class GradXY(torch.autograd.Function):
    """pmhstereo_gradxy (pmhstereo.cl:785-831) with its adjoint.  An RGBA
    image enters through its luminance g = 0.2126 R + 0.7152 G + 0.0722 B
    (pmhstereo.cl:800,808), so its adjoint spreads dL/dg over R, G, B with
    those weights (alpha: none)."""

    @staticmethod
    def forward(ctx, img):
        ctx.rgba = img.dim() == 3
        return ops.gradxy(img.contiguous())

    @staticmethod
    def backward(ctx, dG):
        dg = ops.gradxy_backward(dG.contiguous())
        if not ctx.rgba:
            return dg
        lum = torch.tensor([0.2126, 0.7152, 0.0722, 0.0], dtype=dg.dtype, device=dg.device)
        return dg[..., None] * lum


def window_cost(states, left, right, P: int):
    """Cost field of `states` (left view) with gradients to the planes and to
    both intensity images (directly and through their gradient images)."""
    GL = GradXY.apply(left)
    GR = GradXY.apply(right)
    return WindowCost.apply(states, left, GL, right, GR, P)
