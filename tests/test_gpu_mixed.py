"""Mixed-precision step (FP32 stencils, stage states and smoothing; FP64 y_n, accumulator, residual,
refinement and pole-row metric arithmetic) against the FP64 GPU step on the same seeded state.
Bar: every output field within 1e-5 relative (max |x - x64| / max |x64|) after 10 steps, and the
refined Poisson solve at the same 1e-8 FP64 relative residual."""
import numpy as np
import pytest

import aturan_amd as A

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def state(oracle):
    g = A.default_grid(1)
    p = A.default_params(g)
    F = A.synthetic_state(g, p, 20261203)
    oracle.bc(g, p, F)
    return g, p, F


def test_mixed_solve_reaches_fp64_tolerance(state, oracle):
    print("[case] mixed refinement solve to 1e-8", flush=True)
    g, p, F0 = state
    F = F0.copy()
    oracle.hydro(g, p, F)
    oracle.rk4(g, p, F)
    with A.UranusCore(g, p, "mixed") as core:
        core.set_state(F)
        st = core.PressureSolver("converge", rel_tol=1e-8, max_sweeps=60, sweeps=2)
        pd = core.get_field("p_dyn")
    it, rel, _ = oracle.pressure_converge(g, p, F, 60, 1e-13)
    ref = F[A.FIELD["p_dyn"]]
    e = float(np.max(np.abs(pd - ref)) / np.max(np.abs(ref)))
    print(f"[mixed] solve: {st.sweeps_done} FP32 iterations in {st.outer_iters} refinements, rel {st.rel_resid:.3e}, "
          f"p vs converged FP64 {e:.3e}", flush=True)
    assert st.rel_resid <= 1e-8
    assert e <= 1e-6


def test_mixed_step_tracks_fp64(state):
    print("[case] mixed vs fp64, 10 steps", flush=True)
    g, p, F = state
    opts = dict(mode="converge", rel_tol=1e-8, max_sweeps=60)
    with A.UranusCore(g, p, "fp64") as c64, A.UranusCore(g, p, "mixed") as cmx:
        c64.set_state(F)
        cmx.set_state(F)
        c64.step(10, **opts)
        cmx.step(10, sweeps=2, **opts)
        worst = {}
        for n in ("t", "u", "v", "w", "ch4", "p_dyn", "tke", "dis"):
            a, b = cmx.get_field(n), c64.get_field(n)
            worst[n] = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    print("[mixed] 10 steps: " + " ".join(f"{k}={v:.2e}" for k, v in worst.items()), flush=True)
    for n in ("t", "u", "v", "w", "ch4", "p_dyn"):
        assert worst[n] <= 1e-5, (n, worst[n])
