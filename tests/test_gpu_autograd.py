"""The torch.autograd binding of the window gradients on device tensors: wind_misfit(t0, nue, core)
with t0 on the GPU gives, through loss.backward(), exactly the gradients of atu_window_backward, and
they agree with central finite differences along two seeded directions (1e-4). A backward whose
forward is no longer the core's held window raises. torch and the library share the system HIP
runtime (the library is loaded first)."""
import numpy as np
import pytest

import aturan_amd as A
from aturan_amd import _lib as L
from aturan_amd.autograd import wind_misfit

import torch  # noqa: E402  (after the library: one HIP runtime)

pytestmark = pytest.mark.gpu

SEED = 20261203
EPS = 1e-6


def _setup(oracle):
    g = A.default_grid(1)
    g.im, g.jm, g.km = 12, 19, 37
    p = A.default_params(g)
    F = A.synthetic_state(g, p, SEED)
    oracle.bc(g, p, F)
    return g, p, F


def test_autograd_device_gradients_match_library_and_fd(oracle):
    print("[case] autograd wind_misfit on cuda tensors, 12x19x37, 5 steps", flush=True)
    assert len(L.mapped_runtimes()) == 1, L.mapped_runtimes()
    g, p, F = _setup(oracle)
    opts = dict(nsteps=5, checkpoint_every=2, solve_iters=12)
    dev = torch.device("cuda:0")
    T = A.FIELD["t"]
    nu0 = float(F[A.FIELD["nue"]].flat[0])
    with A.UranusCore(g, p) as core:
        t0 = torch.tensor(F[T], dtype=torch.float64, device=dev, requires_grad=True)
        nue = torch.tensor(nu0, dtype=torch.float64, device=dev, requires_grad=True)
        loss = wind_misfit(t0, nue, core, F, **opts)
        loss.backward()
        assert t0.grad.device == dev and t0.grad.shape == t0.shape
        # the same window through the C ABI directly
        core.set_state(F)
        core.set_field("nue", np.full(F[T].shape, nu0))
        L_ref = core.window_forward(**opts)
        gT, gN = core.window_backward(1.0)
        assert float(loss.detach()) == L_ref
        assert np.array_equal(t0.grad.cpu().numpy(), gT)
        assert float(nue.grad) == gN

        def f(t_field):
            with torch.no_grad():
                return float(wind_misfit(torch.tensor(t_field, device=dev), torch.tensor(nu0, device=dev), core, F,
                                         **opts))

        rng = np.random.default_rng(3)
        gnorm = float(np.linalg.norm(gT))   # the standard deviation of dL/dT0 . d over d ~ N(0, I)
        for q in range(2):
            d = rng.standard_normal(F[T].shape)
            fd = (f(F[T] + EPS * d) - f(F[T] - EPS * d)) / (2 * EPS)
            ad = float(np.sum(gT * d))
            e = abs(fd - ad) / max(abs(ad), gnorm)
            print(f"[autograd] dir {q}: AD {ad:+.12e} FD {fd:+.12e} rel {e:.2e}", flush=True)
            assert e <= 1e-4


def test_autograd_backward_refuses_a_stale_forward(oracle):
    print("[case] autograd backward after another forward on the same core", flush=True)
    g, p, F = _setup(oracle)
    opts = dict(nsteps=2, checkpoint_every=2, solve_iters=8)
    dev = torch.device("cuda:0")
    T = A.FIELD["t"]
    nu0 = float(F[A.FIELD["nue"]].flat[0])
    with A.UranusCore(g, p) as core:
        t_a = torch.tensor(F[T], dtype=torch.float64, device=dev, requires_grad=True)
        t_b = torch.tensor(1.001 * F[T], dtype=torch.float64, device=dev, requires_grad=True)
        nue = torch.tensor(nu0, dtype=torch.float64, device=dev)
        loss_a = wind_misfit(t_a, nue, core, F, **opts)
        loss_b = wind_misfit(t_b, nue, core, F, **opts)
        with pytest.raises(RuntimeError):
            loss_a.backward()
        loss_b.backward()
        assert torch.isfinite(t_b.grad).all()
