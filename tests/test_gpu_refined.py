"""The refined shell, 81x361x721 (dr 0.0125, 0.5 deg in theta and phi, 21.1 M cells): a grid the
reference cannot run (its im, jm, km are compile-time constants, planet/cUranusModel.h:867).

* FP64 GPU against the oracle over two converging steps (quadratic warm start; 1e-12 bar);
* the mixed-precision step against the FP64 GPU step after 10 steps (1e-5 bar, max relative);
* the window adjoint against central finite differences along two seeded directions of T0 and for
  the uniform eddy viscosity (1e-4 bar), with the device memory the context and its window hold
  reported against the device's total.
"""
import time

import numpy as np
import pytest

import aturan_amd as A

pytestmark = pytest.mark.gpu

SEED = 20261203
FIELDS = ["t", "u", "v", "w", "ch4", "tke", "dis", "nh4sh", "p_dyn", "p_hydro"]
GIB = 1024 ** 3


def rel_err(a, b):
    d = float(np.max(np.abs(a - b)))
    s = float(np.max(np.abs(b)))
    return d / s if s > 0 else d


@pytest.fixture(scope="module")
def refined(oracle):
    g = A.default_grid(2)
    p = A.default_params(g)
    F = A.synthetic_state(g, p, SEED)
    oracle.bc(g, p, F)
    return g, p, F


def test_refined_fp64_matches_oracle(refined, oracle):
    print("[case] refined 81x361x721: FP64 GPU vs oracle, 2 converging steps", flush=True)
    g, p, F0 = refined
    free0, total = A.core.device_memory()
    with A.UranusCore(g, p) as core:
        free1, _ = A.core.device_memory()
        core.set_state(F0)
        st = core.step(2, "converge", rel_tol=1e-8, max_sweeps=60, warm_start=2)
        G = {n: core.get_field(n) for n in FIELDS}
    print(f"[memory] FP64 context at 81x361x721: {(free0 - free1) / GIB:.2f} GiB of {total / GIB:.1f} GiB",
          flush=True)
    F = F0.copy()
    t0 = time.time()
    it, rel = oracle.step_converge(g, p, F, 2, 60, 1e-8, 2)
    errs = {n: rel_err(G[n], F[A.FIELD[n]]) for n in FIELDS}
    print(f"[parity] refined 2 steps: worst {max(errs.values()):.3e} (oracle {time.time() - t0:.1f} s); "
          f"last solve gpu {st.sweeps_done} it rel {st.rel_resid:.3e}, oracle {it} it rel {rel:.3e}", flush=True)
    assert max(errs.values()) <= 1e-12
    assert st.rel_resid <= 1e-8


def test_refined_mixed_tracks_fp64(refined):
    print("[case] refined 81x361x721: mixed vs FP64 GPU, 10 converging steps", flush=True)
    g, p, F0 = refined
    res = {}
    for prec in ("fp64", "mixed"):
        with A.UranusCore(g, p, prec) as core:
            core.set_state(F0)
            core.step(10, "converge", rel_tol=1e-8, max_sweeps=60, warm_start=2)
            res[prec] = {n: core.get_field(n) for n in FIELDS[:5] + ["p_dyn"]}
    errs = {n: rel_err(res["mixed"][n], res["fp64"][n]) for n in res["fp64"]}
    print("[parity] refined mixed vs fp64 after 10 steps: " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()),
          flush=True)
    assert max(errs.values()) <= 1e-5


def test_refined_adjoint_matches_fd(refined):
    print("[case] refined 81x361x721: window adjoint (4 steps, checkpoint every 2) vs FD", flush=True)
    g, p, F0 = refined
    eps, bar = 1e-6, 1e-4
    opts = dict(nsteps=4, checkpoint_every=2, solve_iters=12)
    rng = np.random.default_rng(5)
    T, NU = A.FIELD["t"], A.FIELD["nue"]
    free0, total = A.core.device_memory()
    with A.UranusCore(g, p) as core:
        core.set_state(F0)
        L0 = core.window_forward(**opts)
        gT, gN = core.window_backward(1.0)
        free1, _ = A.core.device_memory()
        print(f"[memory] FP64 context with a 4-step window (3 checkpoints, 2 tape slots) at 81x361x721: "
              f"{(free0 - free1) / GIB:.2f} GiB of {total / GIB:.1f} GiB", flush=True)

        def loss(t_field, nu_field):
            core.window_rewind()
            core.set_field("t", t_field)
            core.set_field("nue", nu_field)
            return core.window_forward(**opts)

        worst = 0.0
        gnorm = float(np.linalg.norm(gT))   # the standard deviation of dL/dT0 . d over d ~ N(0, I)
        for q in range(2):
            d = rng.standard_normal(F0[T].shape)
            fd = (loss(F0[T] + eps * d, F0[NU]) - loss(F0[T] - eps * d, F0[NU])) / (2 * eps)
            ad = float(np.sum(gT * d))
            e = abs(fd - ad) / max(abs(ad), gnorm, 1e-300)
            worst = max(worst, e)
            print(f"[adjoint] refined dir {q}: AD {ad:+.12e} FD {fd:+.12e} rel {e:.2e}", flush=True)
        h = eps * float(F0[NU].flat[0])
        fdn = (loss(F0[T], F0[NU] + h) - loss(F0[T], F0[NU] - h)) / (2 * h)
        en = abs(fdn - gN) / max(abs(gN), 1e-300)
        print(f"[adjoint] refined nue: AD {gN:+.12e} FD {fdn:+.12e} rel {en:.2e}; L0 {L0:.12e}", flush=True)
    assert np.isfinite(gT).all()
    assert worst <= bar and en <= bar
