"""The hand-written adjoint of the window of steps against central finite differences of the forward,
with the converging solve (differentiated as an exact solve) and with the reference's red-black
sweeps (differentiated sweep by sweep).

Loss: mean over the interior (r, theta) of (zonal-mean zonal wind w - 1.2 cos 3 lat)^2 after the
window. Gradients: with respect to the initial temperature field (checked along seeded random
directions d: dL/dT0 . d against (L(T0 + eps d) - L(T0 - eps d)) / (2 eps), eps = 1e-6) and the
uniform eddy viscosity nue (relative step 1e-6). The solves run a fixed iteration count, so the
forward is a smooth map; the adjoint differentiates the solve as exact, which the fixed count
reaches to the solver's rounding floor. Bar: |FD - AD| <= 1e-4 of the scale of the directional
derivative, max(|AD|, ||dL/dT0||_2): for d drawn from N(0, I), dL/dT0 . d has standard deviation
||dL/dT0||_2, and a direction drawn nearly orthogonal to the gradient (|AD| far below that) would
otherwise measure the central difference's rounding floor (~1e-16 L / eps), not the gradient.
"""
import time

import numpy as np
import pytest

import aturan_amd as A

pytestmark = pytest.mark.gpu

EPS = 1e-6
BAR = 1e-4


def _state(g, p, oracle, seed=20261203):
    F = A.synthetic_state(g, p, seed)
    oracle.bc(g, p, F)
    return F


def _fd_check(g, p, F, opts, ndir, seed, eps=EPS):
    rng = np.random.default_rng(seed)
    T = A.FIELD["t"]
    nue = A.FIELD["nue"]
    t0 = time.time()
    with A.UranusCore(g, p) as core:
        core.set_state(F)
        L0 = core.window_forward(**opts)
        t1 = time.time()
        gT, gN = core.window_backward(1.0)
        t2 = time.time()
        print(f"[adjoint] forward {t1 - t0:.2f} s (with upload), backward {t2 - t1:.2f} s", flush=True)

        def loss(t_field, nu_field):
            core.window_rewind()                 # the window's initial state, on the device
            core.set_field("t", t_field)
            core.set_field("nue", nu_field)
            return core.window_forward(**opts)

        worst = 0.0
        gnorm = float(np.linalg.norm(gT))
        for q in range(ndir):
            d = rng.standard_normal(F[T].shape)
            fd = (loss(F[T] + eps * d, F[nue]) - loss(F[T] - eps * d, F[nue])) / (2 * eps)
            ad = float(np.sum(gT * d))
            e = abs(fd - ad) / max(abs(ad), gnorm, 1e-300)
            worst = max(worst, e)
            print(f"[adjoint] dir {q}: AD {ad:+.12e} FD {fd:+.12e} rel {abs(fd - ad) / max(abs(ad), 1e-300):.2e} "
                  f"(of the scale {max(abs(ad), gnorm):.3e}: {e:.2e})", flush=True)
        nu = float(F[nue].flat[0])
        h = eps * nu
        fdn = (loss(F[T], F[nue] + h) - loss(F[T], F[nue] - h)) / (2 * h)
        en = abs(fdn - gN) / max(abs(gN), 1e-300)
        print(f"[adjoint] nue: AD {gN:+.12e} FD {fdn:+.12e} rel {en:.2e}; L0 {L0:.12e}; "
              f"{time.time() - t0:.1f} s in all", flush=True)
    return worst, en, gT, gN


def test_adjoint_small_shell_matches_fd(oracle):
    print("[case] adjoint vs FD, 12x19x37 shell, 6 steps, checkpoint every 4", flush=True)
    g = A.default_grid(1)
    g.im, g.jm, g.km = 12, 19, 37
    p = A.default_params(g)
    F = _state(g, p, oracle)
    opts = dict(nsteps=6, checkpoint_every=4, solve_iters=14)
    worst, en, gT, gN = _fd_check(g, p, F, opts, ndir=4, seed=7)
    assert np.isfinite(gT).all() and np.isfinite(gN)
    assert worst <= BAR and en <= BAR


def test_adjoint_backward_is_repeatable_and_linear(oracle):
    print("[case] adjoint determinism and linearity in the seed", flush=True)
    g = A.default_grid(1)
    g.im, g.jm, g.km = 12, 19, 37
    p = A.default_params(g)
    F = _state(g, p, oracle)
    with A.UranusCore(g, p) as core:
        core.set_state(F)
        core.window_forward(nsteps=3, checkpoint_every=2, solve_iters=10)
        g1, n1 = core.window_backward(1.0)
        g2, n2 = core.window_backward(1.0)
        g3, n3 = core.window_backward(-2.5)
    print(f"[adjoint] repeat max diff {np.max(np.abs(g1 - g2)):.3e} nue {n1!r} {n2!r}; linearity "
          f"{np.max(np.abs(g3 + 2.5 * g1)) / np.max(np.abs(g1)):.3e}", flush=True)
    assert np.array_equal(g1, g2) and n1 == n2          # gather-form kernels: no atomics
    # linear in the seed up to rounding: the scaled seed rounds differently along the way, and the
    # adjoint solves stop at the solver's rounding floor (~1e-11 relative residual)
    assert np.max(np.abs(g3 + 2.5 * g1)) <= 1e-9 * np.max(np.abs(g1))
    assert abs(n3 + 2.5 * n1) <= 1e-9 * abs(n1)


def test_mixed_window_adjoint_tracks_fp64(oracle):
    """The window on a mixed context (FP32 stencils and stage states, refined solves) against FP64."""
    print("[case] mixed vs FP64 window adjoint, 12x19x37 shell, 6 steps", flush=True)
    g = A.default_grid(1)
    g.im, g.jm, g.km = 12, 19, 37
    p = A.default_params(g)
    F = _state(g, p, oracle)
    res = {}
    for prec in ("fp64", "mixed"):
        with A.UranusCore(g, p, prec) as core:
            core.set_state(F)
            loss = core.window_forward(nsteps=6, checkpoint_every=4, rel_tol=1e-10)
            res[prec] = (loss,) + core.window_backward(1.0)
    (l64, g64, n64), (lmx, gmx, nmx) = res["fp64"], res["mixed"]
    eg = float(np.linalg.norm(gmx - g64) / np.linalg.norm(g64))
    en = abs(nmx - n64) / abs(n64)
    print(f"[adjoint] mixed vs fp64: loss {abs(lmx - l64) / l64:.2e}, grad T0 {eg:.2e}, grad nue {en:.2e}", flush=True)
    assert abs(lmx - l64) <= 1e-6 * l64 and eg <= 1e-4 and en <= 1e-4


def test_rb_sweep_adjoint_matches_fd(oracle):
    """The window with the reference's red-black sweeps as its solve (rb_sweeps = 20 per step, omega 1),
    differentiated sweep by sweep: walls folded back, the black then the red pass transposed."""
    print("[case] red-black sweep adjoint vs FD, 12x19x37 shell, 6 steps, 20 sweeps per solve", flush=True)
    g = A.default_grid(1)
    g.im, g.jm, g.km = 12, 19, 37
    p = A.default_params(g)
    F = _state(g, p, oracle)
    opts = dict(nsteps=6, checkpoint_every=4, rb_sweeps=20)
    worst, en, gT, gN = _fd_check(g, p, F, opts, ndir=4, seed=13)
    assert np.isfinite(gT).all() and np.isfinite(gN)
    assert worst <= BAR and en <= BAR


def test_rb_sweep_adjoint_full_shell_matches_fd(oracle):
    """The same at 41x181x361: 3 steps of 20 sweeps. Here dL/dT0 . d is ~4e-10 against L ~ 2e-2, so the
    central difference's rounding floor, ~1e-16 L / eps, is 5e-3 of it at eps = 1e-6: the step is
    eps = 1e-3 (floor ~5e-6; the third-order truncation, ~eps^2, below it)."""
    print("[case] red-black sweep adjoint vs FD, 41x181x361, 3 steps, 20 sweeps per solve", flush=True)
    g = A.default_grid(1)
    p = A.default_params(g)
    F = _state(g, p, oracle)
    opts = dict(nsteps=3, checkpoint_every=2, rb_sweeps=20)
    worst, en, gT, gN = _fd_check(g, p, F, opts, ndir=2, seed=17, eps=1e-3)
    assert worst <= BAR and en <= BAR


def test_mixed_rb_window_adjoint_tracks_fp64(oracle):
    """The red-black window on a mixed context against FP64 (same bars as the converging window)."""
    print("[case] mixed vs FP64 red-black window adjoint, 12x19x37 shell, 6 steps", flush=True)
    g = A.default_grid(1)
    g.im, g.jm, g.km = 12, 19, 37
    p = A.default_params(g)
    F = _state(g, p, oracle)
    res = {}
    for prec in ("fp64", "mixed"):
        with A.UranusCore(g, p, prec) as core:
            core.set_state(F)
            loss = core.window_forward(nsteps=6, checkpoint_every=4, rb_sweeps=20)
            res[prec] = (loss,) + core.window_backward(1.0)
    (l64, g64, n64), (lmx, gmx, nmx) = res["fp64"], res["mixed"]
    eg = float(np.linalg.norm(gmx - g64) / np.linalg.norm(g64))
    en = abs(nmx - n64) / abs(n64)
    print(f"[adjoint] red-black mixed vs fp64: loss {abs(lmx - l64) / l64:.2e}, grad T0 {eg:.2e}, grad nue {en:.2e}",
          flush=True)
    assert abs(lmx - l64) <= 1e-6 * l64 and eg <= 1e-4 and en <= 1e-4


@pytest.mark.slow
def test_adjoint_full_shell_tolerance_solves(oracle):
    """The configuration bench.py times: 41x181x361, 50 steps, checkpoint every 10, every solve to a
    relative residual of 1e-8 (the adjoint differentiates the converged solve)."""
    print("[case] adjoint vs FD, 41x181x361, 50 steps, checkpoint every 10, solves to 1e-8, 4 directions",
          flush=True)
    g = A.default_grid(1)
    p = A.default_params(g)
    F = _state(g, p, oracle)
    opts = dict(nsteps=50, checkpoint_every=10, rel_tol=1e-8)
    worst, en, gT, gN = _fd_check(g, p, F, opts, ndir=4, seed=12)
    assert worst <= BAR and en <= BAR


@pytest.mark.slow
def test_adjoint_full_shell_50_steps(oracle):
    print("[case] adjoint vs FD, 41x181x361, 50 steps, checkpoint every 10, 16 directions", flush=True)
    g = A.default_grid(1)
    p = A.default_params(g)
    F = _state(g, p, oracle)
    opts = dict(nsteps=50, checkpoint_every=10, solve_iters=10)
    worst, en, gT, gN = _fd_check(g, p, F, opts, ndir=16, seed=11)
    assert worst <= BAR and en <= BAR
