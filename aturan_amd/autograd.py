"""torch.autograd bridge of the window gradients (atu_window_forward / atu_window_backward_device).

`wind_misfit(t0, nue, core, state=None, **opts)` runs a window of steps on the UranusCore `core`
from its current state (or from `state`, the 46 dense host fields, uploaded first) with the
initial temperature replaced by the device tensor `t0` and the eddy viscosity set to the uniform
value `nue`, and returns the misfit L of the final zonal-mean zonal wind to the target jet. Its
backward runs the hand-written discrete adjoint on the GPU (checkpoint every
`opts["checkpoint_every"]` steps, each segment recomputed with its stage states taped) and returns
dL/dt0 as a tensor of t0's shape and device and dL/dnue as a 0-d tensor. t0 and the gradient move
between torch and the library device to device (atu_set_field_device, atu_window_backward_device).

The backward differentiates the forward it belongs to: when another forward ran on the same core
in between, or a frozen input of the window was written since, it raises instead of returning the
gradient of a different window.

The adjoint has no counterpart in the reference. PyTorch is the interface here, not the compute
path: both passes are the HIP kernels of libaturan_amd.so, which must be loaded before torch so
that both run on the same HIP runtime (aturan_amd._lib.load raises otherwise).
"""
from __future__ import annotations

from . import _lib as L

L.load()   # before torch: the library and torch then share the system HIP runtime

import numpy as np  # noqa: E402
import torch  # noqa: E402

from .core import UranusCore  # noqa: E402


def _dense_f64(x: torch.Tensor, shape, device) -> torch.Tensor:
    x = x.detach().to(device=device, dtype=torch.float64)
    if tuple(x.shape) != tuple(shape):
        raise ValueError(f"t0 shape {tuple(x.shape)} != grid {tuple(shape)}")
    return x.contiguous()


class WindMisfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, t0: torch.Tensor, nue: torch.Tensor, core: UranusCore, state, opts: dict):
        if t0.device.type != "cuda":
            raise ValueError("t0 must be a tensor on the GPU the core runs on")
        if state is not None:
            core.set_state(np.asarray(state))
        t = _dense_f64(t0, core.shape, t0.device)
        nu = torch.full(core.shape, float(nue.detach().cpu()), dtype=torch.float64, device=t0.device)
        torch.cuda.synchronize(t0.device)            # the copies read t and nu on the library's stream
        core.set_field_device("t", t.data_ptr())
        core.set_field_device("nue", nu.data_ptr())
        loss = core.window_forward(**opts)
        ctx.core = core
        ctx.generation = core.window_generation()
        ctx.shape = core.shape
        ctx.device = t0.device
        return torch.tensor(loss, dtype=torch.float64, device=t0.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        core = ctx.core
        if core.window_generation() != ctx.generation:
            raise RuntimeError("wind_misfit backward: the core ran another window forward, or a frozen input of "
                               "the window was written, after this forward; its gradient is no longer available")
        g = torch.empty(ctx.shape, dtype=torch.float64, device=ctx.device)
        torch.cuda.synchronize(ctx.device)
        g_nue = core.window_backward_device(float(grad_out.detach().cpu()), g.data_ptr())
        return g, torch.tensor(g_nue, dtype=torch.float64, device=ctx.device), None, None, None


def wind_misfit(t0: torch.Tensor, nue: torch.Tensor, core: UranusCore, state=None, **opts) -> torch.Tensor:
    """Differentiable window loss; see WindMisfit."""
    return WindMisfit.apply(t0, nue, core, state, opts)
