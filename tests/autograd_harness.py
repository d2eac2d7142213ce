"""Verification harness: phi+(w+, W-) of one propagator evaluation as a
torch.autograd.Function whose backward is the library's adjoint (lfts_vjp) and whose
forward-mode rule is its tangent (lfts_jvp). Used only by the tests (torch gradcheck
compares both against finite differences of the forward evaluation); the library
itself has no torch dependency."""
from __future__ import annotations

import numpy as np
import torch


class PhiPlus(torch.autograd.Function):
    """phi+ = f(w+, W-) for a fixed lfts_cpu_amd.Context (float64 CPU tensors of length M)."""

    @staticmethod
    def forward(ctx, wp: torch.Tensor, wm: torch.Tensor, lctx) -> torch.Tensor:
        M = lctx.M
        w = np.concatenate([wm.detach().cpu().numpy(), wp.detach().cpu().numpy()])
        lctx.set_fields(w)
        lctx.calc_concs()
        out = lctx.get_fields(4)[3 * M:].copy()
        ctx.lctx = lctx
        ctx.w = w
        return torch.from_numpy(out)

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        lctx = ctx.lctx
        M = lctx.M
        lctx.set_fields(ctx.w)
        ubar = np.zeros(2 * M)
        ubar[M:] = g.detach().cpu().numpy()
        wbar = lctx.vjp(ubar)
        return torch.from_numpy(wbar[M:].copy()), torch.from_numpy(wbar[:M].copy()), None

    @staticmethod
    def jvp(ctx, dwp: torch.Tensor, dwm: torch.Tensor, _):
        lctx = ctx.lctx
        M = lctx.M
        lctx.set_fields(ctx.w)
        a = None if dwm is None else dwm.detach().cpu().numpy()
        b = None if dwp is None else dwp.detach().cpu().numpy()
        dphi = lctx.jvp(a, b)
        return torch.from_numpy(dphi[M:].copy())
