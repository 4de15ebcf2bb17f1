"""Parity over long runs, the species clamp after the BC, and the window's refusal to differentiate a
forward whose inputs changed.

FP64 bar: every prognostic field, p_dyn and p_hydro within 1e-12 relative (max |x - x_ref| /
max |x_ref|) of the oracle on the same seeded input; the converging step is bit-identical to the
oracle in practice (same operand order, same fixed-order sums), which the 100-step test checks too.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

import aturan_amd as A

pytestmark = pytest.mark.gpu

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
import make_golden as MG  # noqa: E402

SEED = 20261203
TOL = 1e-12
PROG = ["t", "u", "v", "w", "ch4", "ch4_cloud", "ch4_ice", "h2o", "h2o_cloud", "h2o_ice", "h2s", "h2s_cloud",
        "h2s_ice", "nh3", "nh3_cloud", "nh3_ice", "nh4sh", "tke", "dis"]
SPECIES = PROG[4:17]


def rel_err(a, b):
    d = float(np.max(np.abs(a - b)))
    s = float(np.max(np.abs(b)))
    return d / s if s > 0 else d


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_bc_species_clamp_matches_oracle(oracle, prec):
    """BC then clampNegativeSpecies (on by default, planet/BC_Uran.h:110-128) on an input whose species
    hold negative values and NaNs in the interior and on every boundary plane: no species value is
    left below zero or NaN, and every field equals the oracle's exactly."""
    print(f"[case] bc + species clamp, {prec} context", flush=True)
    g = A.default_grid(1)
    p = A.default_params(g)
    assert p.species_clamp == 1
    F = MG.case_input("bc_clamp", g, p)
    bad = sum(int(np.count_nonzero(~(F[A.FIELD[n]] >= 0.0))) for n in SPECIES)
    assert bad > 0.05 * len(SPECIES) * F[0].size
    with A.UranusCore(g, p, prec) as core:
        core.set_state(F)
        core.boundary_conditions()
        G = core.get_state()
    oracle.bc(g, p, F)
    for n in SPECIES:
        assert np.all(G[A.FIELD[n]] >= 0.0), n
    for n in A.FIELD_NAMES:
        if prec == "mixed" and n.startswith("aux"):
            continue                       # a mixed context keeps aux in FP32
        assert np.array_equal(G[A.FIELD[n]], F[A.FIELD[n]]), n
    print(f"[parity] bc + clamp {prec}: {bad} negative or NaN species values before, none after", flush=True)


def shell_mask(g):
    """The (i, j) rows the BC writes or reads (i < 3 or i > im-4, or j < 3 or j > jm-4)."""
    far = np.zeros((g.im, g.jm), bool)
    far[3:g.im - 3, 3:g.jm - 3] = True
    return ~far


def test_bc_species_clamp_then_step_matches_oracle(oracle):
    """The clamp's writes reach the stage buffers the next step's stencils read (a mixed context's FP32
    mirrors, an FP64 context's stage states): one converging step after the clamp equals the oracle.
    The step's own clamp is split (the last stage clamps the rows the BC does not touch, a pass after
    the BC the shell of rows it does): the same step without the clamp leaves negative species in both
    parts, and the split clamp gives the oracle's clamp-after-BC values bit for bit."""
    print("[case] bc + clamp, then one converging step", flush=True)
    g = A.default_grid(1)
    p = A.default_params(g)
    F = MG.case_input("bc_clamp", g, p)
    with A.UranusCore(g, p) as core:
        core.set_state(F)
        core.boundary_conditions()
        core.step(1, "converge", rel_tol=1e-8, max_sweeps=50)
        G = core.get_state()
    oracle.bc(g, p, F)
    p0 = A.default_params(g)
    p0.species_clamp = 0
    H = F.copy()
    oracle.step_converge(g, p0, H, 1, 50, 1e-8, 0)
    shell = shell_mask(g)
    neg_far = neg_shell = 0
    for n in SPECIES:
        bad = ~(H[A.FIELD[n]].reshape(g.im, g.jm, -1) >= 0.0)
        neg_far += int(bad[~shell].sum())
        neg_shell += int(bad[shell].sum())
    print(f"[parity] unclamped step: {neg_far} negative species values at the rows the BC does not touch, "
          f"{neg_shell} in the shell", flush=True)
    assert neg_far > 0 and neg_shell > 0
    oracle.step_converge(g, p, F, 1, 50, 1e-8, 0)
    errs = {n: rel_err(G[A.FIELD[n]], F[A.FIELD[n]]) for n in PROG + ["p_dyn", "p_hydro"]}
    print(f"[parity] clamp then step: worst {max(errs.values()):.3e}", flush=True)
    assert max(errs.values()) == 0.0


def test_mixed_step_species_clamp():
    """The split clamp on a mixed context: after two steps from the clamped input no species value is
    negative or NaN, and the FP32 mirror the next stencils read holds the clamped values (a third step
    stays within 1e-5 of the FP64 context's)."""
    print("[case] mixed context: bc + clamp, then steps with the split clamp", flush=True)
    g = A.default_grid(1)
    p = A.default_params(g)
    F = MG.case_input("bc_clamp", g, p)
    out = {}
    for prec in ("fp64", "mixed"):
        with A.UranusCore(g, p, prec) as core:
            core.set_state(F)
            core.boundary_conditions()
            core.step(2, "converge", rel_tol=1e-8, max_sweeps=50)
            G2 = core.get_state()
            core.step(1, "converge", rel_tol=1e-8, max_sweeps=50)
            out[prec] = (G2, core.get_state())
    for n in SPECIES:
        assert np.all(out["mixed"][0][A.FIELD[n]] >= 0.0), n
    errs = {n: rel_err(out["mixed"][1][A.FIELD[n]], out["fp64"][1][A.FIELD[n]]) for n in PROG}
    print(f"[parity] mixed vs fp64 after 3 steps: worst {max(errs.values()):.3e}", flush=True)
    assert max(errs.values()) <= 1e-5


def test_converge_100_steps_match_oracle(oracle):
    """The benchmark's step (hydro prep, RK4, BC, clamp, quadratic-warm-started solve to 1e-8), 100 steps
    in one call, against the oracle's 100 steps: the long-run form of the FP64 parity bar."""
    print("[case] 100 converging steps, warm start 2", flush=True)
    g = A.default_grid(1)
    p = A.default_params(g)
    F = A.synthetic_state(g, p, SEED)
    oracle.bc(g, p, F)
    with A.UranusCore(g, p) as core:
        core.set_state(F)
        st = core.step(100, "converge", rel_tol=1e-8, max_sweeps=60, warm_start=2)
        G = core.get_state()
    it, rel = oracle.step_converge(g, p, F, 100, 60, 1e-8, 2)
    errs = {n: rel_err(G[A.FIELD[n]], F[A.FIELD[n]]) for n in PROG + ["p_dyn", "p_hydro"]}
    worst = max(errs.values())
    print(f"[parity] 100 steps: worst {worst:.3e}; last solve gpu {st.sweeps_done} it rel {st.rel_resid:.3e}, "
          f"oracle {it} it rel {rel:.3e}", flush=True)
    assert worst <= TOL
    assert st.rel_resid <= 1e-8 and rel <= 1e-8


def test_profile_sample_brackets_every_nth_step(oracle):
    """atu_profile_sample: with every = 3, only steps 0, 3, 6 of a 7-step call carry the stage and step
    events (4 stage launches each); the state equals that of an unprofiled run."""
    print("[case] profiler sample: every 3rd step of 7", flush=True)
    g = A.default_grid(1)
    g.im, g.jm, g.km = 12, 19, 37
    p = A.default_params(g)
    F = A.synthetic_state(g, p, SEED)
    oracle.bc(g, p, F)
    opts = dict(rel_tol=1e-8, max_sweeps=60, warm_start=2)
    with A.UranusCore(g, p) as core:
        core.set_state(F)
        core.profile_begin("rk_stage,step", every=3)
        core.step(7, "converge", **opts)
        ms, n_rk = core.profile_query("rk_stage")
        _, n_step = core.profile_query("step")
        core.profile_end()
        G = core.get_state()
        core.set_state(F)
        core.profile_begin("rk_stage", every=1)
        core.step(2, "converge", **opts)
        _, n_all = core.profile_query("rk_stage")
        core.profile_end()
    with A.UranusCore(g, p) as core:
        core.set_state(F)
        core.step(7, "converge", **opts)
        H = core.get_state()
    print(f"[profile] sampled stage launches {n_rk}, steps {n_step}, unsampled 2 steps {n_all}, {ms:.3f} ms",
          flush=True)
    assert (n_rk, n_step, n_all) == (12, 3, 8) and ms > 0.0
    assert np.array_equal(G, H)


def test_window_backward_refuses_changed_inputs(oracle):
    """A frozen input written between the forward and the backward (here nue) makes the backward
    refuse; rewinding still works, and a new forward re-arms the backward with a new generation."""
    print("[case] window invalidation on a frozen-field write", flush=True)
    g = A.default_grid(1)
    g.im, g.jm, g.km = 12, 19, 37
    p = A.default_params(g)
    F = A.synthetic_state(g, p, SEED)
    oracle.bc(g, p, F)
    with A.UranusCore(g, p) as core:
        core.set_state(F)
        assert core.window_generation() == 0
        core.window_forward(nsteps=3, checkpoint_every=2, solve_iters=6)
        g1 = core.window_generation()
        assert g1 > 0
        core.set_field("t", F[A.FIELD["t"]])        # a prognostic field: restored from the checkpoints
        assert core.window_generation() == g1
        core.window_backward(1.0)
        core.set_field("nue", 1.1 * F[A.FIELD["nue"]])
        assert core.window_generation() == 0
        with pytest.raises(A.AtuError):
            core.window_backward(1.0)
        core.window_rewind()
        core.window_forward(nsteps=3, checkpoint_every=2, solve_iters=6)
        g2 = core.window_generation()
        assert g2 > g1
        core.window_backward(1.0)
