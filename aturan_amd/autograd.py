"""torch.autograd bridge of the window gradients (atu_window_forward / atu_window_backward).

`WindMisfit.apply(t0, nue, core, state, opts)` runs a window of steps from `state` (the 46 dense
fields of a UranusCore) with the initial temperature replaced by `t0` and the eddy viscosity set to
the uniform value `nue`, and returns the misfit L of the final zonal-mean zonal wind to the target
jet. Its backward runs the hand-written discrete adjoint on the GPU (checkpoint every
`opts["checkpoint_every"]` steps, the segment recomputed with its stage states taped) and returns
dL/dt0 as a tensor of t0's shape and device and dL/dnue as a 0-d tensor.

The adjoint has no counterpart in the reference. PyTorch is the interface here, not the compute
path: both passes are the HIP kernels of libaturan_amd.so.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from .core import UranusCore


class WindMisfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, t0: torch.Tensor, nue: torch.Tensor, core: UranusCore, state: np.ndarray, opts: dict):
        if t0.shape != core.shape:
            raise ValueError(f"t0 shape {tuple(t0.shape)} != grid {core.shape}")
        F = state.copy()
        F[L.FIELD["t"]] = t0.detach().to("cpu", torch.float64).numpy()
        F[L.FIELD["nue"]] = float(nue.detach().cpu())
        core.set_state(F)
        loss = core.window_forward(**opts)
        ctx.core = core
        ctx.device = t0.device
        return torch.tensor(loss, dtype=torch.float64, device=t0.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        g_t0, g_nue = ctx.core.window_backward(float(grad_out.detach().cpu()))
        return (torch.from_numpy(g_t0).to(ctx.device), torch.tensor(g_nue, dtype=torch.float64, device=ctx.device),
                None, None, None)


def wind_misfit(t0: torch.Tensor, nue: torch.Tensor, core: UranusCore, state: np.ndarray, **opts) -> torch.Tensor:
    """Differentiable window loss; see WindMisfit."""
    return WindMisfit.apply(t0, nue, core, state, opts)
