"""Pins the CPU oracle to golden vectors written by the compiled reference (tests/golden/ref_*.npz,
from tests/golden/make_golden.py): the domain BC, one RungeKuttaUran (hydro prep + 4 stages), the
red-black PressureSolver at 1 and 20 sweeps (alone, with its PRESS_LINE radial line pass, with its
PRESS_BAROTROPIC column-mean correction, and with both), two full steps, and the default PressureSolverUran after
one RungeKuttaUran and in two full steps, the Shapiro wiggle filter, the turbulence closure
(Turbulence::run and ::init under six sets of its knobs) and the loop's saturation-adjustment block
(four gases and the NH4SH equilibrium, on two inputs), on an input in which every term of
the path is non-zero. Per field: the sampled cells within 1e-12 of the field's largest sample, and
the exact sum and sum of squares within 1e-12 relative (the closure's own arrays prod, tken, disn and
vel_star likewise, and exactly)."""
import math
import sys
from pathlib import Path

import numpy as np
import pytest

import aturan_amd as A

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
import make_golden as MG  # noqa: E402

TOL = 1e-12
# p_hydro integrates the buoyancy anomaly (b - ref) / ref, which magnifies a rounding difference of the
# level mean ref by nd_rel / |anomaly| (~1e4 on these states). The level sums here run in a fixed tree
# order (closer to the exact sum than the reference's serial order, test_level_means_track_exact_sum),
# so p_hydro differs from the reference's by that magnified rounding: ~1e-11 of its largest value, and
# its sum, which cancels to ~1/300 of its magnitude, by ~1e-9 relative. Its bar is therefore 1e-10 of
# its largest sample, and the sums 1e-10 of sqrt(N * sum of squares); every other field keeps 1e-12.
PH_TOL = 1e-10
PH = A.FIELD["p_hydro"]


def run_oracle(oracle, name, g, p):
    F = MG.case_input(name, g, p)
    mode, steps, sweeps, _ = MG.CASES[name]
    level = np.zeros(g.im)
    if mode == "satadj":
        oracle.satadj(g, p, F)
        return F, level, None
    if mode in ("turb_run", "turb_init"):
        T3, vs = MG.turb_input(g)
        oracle.turb(g, p, F, T3[0], T3[1], T3[2], vs, init=(mode == "turb_init"), **MG.TURB_KNOBS[name])
        return F, level, (T3, vs)
    if mode in ("bc", "bc_clamp"):
        pb = A.default_params(g)
        pb.species_clamp = int(mode == "bc_clamp")      # "bc" is the BC alone
        oracle.bc(g, pb, F)
    elif mode == "rk4":
        level = oracle.hydro(g, p, F)
        oracle.rk4(g, p, F)
    elif mode == "pressure":
        oracle.pressure(g, p, F, sweeps, 1.0, MG.PRESS_FLAGS.get(name, 0))
    elif mode == "rk4_uran":
        level = oracle.hydro(g, p, F)
        rhs3 = oracle.rk4_rhs(g, p, F)
        oracle.pressure_uran(g, p, F, rhs3, 1)
    elif mode == "damp_t":
        oracle.damp_wiggles(g, F, A.FIELD["t"], False, True, True, 1.0, 1)
        oracle.damp_wiggles(g, F, A.FIELD["t"], True, False, False, 0.5, 1)
    elif mode == "damp_uvw":
        for f in ("u", "v", "w"):
            oracle.damp_wiggles(g, F, A.FIELD[f], True, True, True, 1.0, sweeps)
    elif mode == "step_uran":
        for _ in range(steps):
            level = oracle.hydro(g, p, F)
            rhs3 = oracle.rk4_rhs(g, p, F)
            oracle.bc(g, p, F)
            oracle.pressure_uran(g, p, F, rhs3, 1)
    else:
        for _ in range(steps):
            level = oracle.hydro(g, p, F)
            oracle.rk4(g, p, F)
            oracle.bc(g, p, F)
            oracle.pressure(g, p, F, sweeps, 1.0)
    return F, level, None


@pytest.mark.slow
@pytest.mark.parametrize("name", list(MG.CASES))
def test_oracle_matches_compiled_reference(oracle, name):
    ref = np.load(GOLDEN / f"ref_{name}.npz")
    g = A.default_grid(1)
    p = A.default_params(g)
    F, level, turb = run_oracle(oracle, name, g, p)
    got = MG.summarize(F, level)
    if turb is not None:
        t = MG.summarize_turb(*turb)
        for key in ("turb_samples", "turb_vs_samples", "turb_sums", "turb_sumsq"):
            assert np.array_equal(t[key], ref[key]), (name, key)
    worst = {}
    n = F[0].size
    for f, fname in enumerate(A.FIELD_NAMES):
        a, b = got["samples"][f], ref["samples"][f]
        scale = float(np.max(np.abs(b))) or 1.0
        worst[fname] = float(np.max(np.abs(a - b))) / scale
        for key in ("sums", "sumsq"):
            x, y = got[key][f], ref[key][f]
            bar = TOL * max(abs(y), 1e-300)
            if f == PH:
                bar = PH_TOL * (math.sqrt(n * ref["sumsq"][f]) if key == "sums" else abs(y))
            assert abs(x - y) <= bar or x == y, (name, fname, key, x, y)
    bad = {k: v for k, v in worst.items() if v > (PH_TOL if k == "p_hydro" else TOL)}
    print(f"[golden] {name}: worst sampled difference {max(worst.values()):.3e}")
    assert not bad, bad
    if MG.CASES[name][0] in ("rk4", "step", "rk4_uran", "step_uran"):
        assert np.max(np.abs(got["level"] - ref["level"])) <= TOL * np.max(np.abs(ref["level"]))


def test_fixture_constants_match_library_params():
    """The reference's mixture constants after T_INIT_ANCHOR equal atu_params_default's."""
    ref = np.load(GOLDEN / "ref_rk4.npz")
    line = str(ref["consts"][0])
    vals = dict(zip(line.split()[1::2], map(float, line.split()[2::2])))
    p = A.default_params(A.default_grid(1))
    for key in ("r_mix", "R_mix", "cp_mix", "mue_mix", "r_h2"):
        assert getattr(p, key) == vals[key], key


def test_satadj_defaults_match_reference_constants():
    """atu_satadj_opts_default (and the oracle's table) carry the gas constants the compiled reference's
    loop passes, as its harness printed them."""
    from oracle import binding as O
    ref = np.load(GOLDEN / "ref_satadj.npz")
    lines = [str(x) for x in ref["consts"] if str(x).startswith("ref_satadj:")]
    printed = {ln.split()[1]: [float(v) for v in ln.split()[2:]] for ln in lines}
    o = A.core.satadj_opts()
    names = ("t_0", "t_00", "ep", "lv", "ls", "C", "L0", "R", "del_alf", "del_bet", "C_i", "L0_i", "del_alf_i",
             "del_bet_i")
    for gas, n in A._lib.GASES.items():
        assert [getattr(o.gas[n], f) for f in names] == printed[gas], gas
        assert list(O.GASES[gas]) == printed[gas], gas
    assert [o.m_nh3, o.m_h2s, o.m_nh4sh] == printed["molar"] == list(O.MOLAR)


def test_enriched_input_is_deterministic():
    g = A.default_grid(1)
    g.im, g.jm, g.km = 8, 9, 13
    p = A.default_params(g)
    a = MG.enriched_state(g, p)
    b = MG.enriched_state(g, p)
    assert np.array_equal(a, b)
    assert math.isclose(float(np.mean(a[A.FIELD["nue"]] < 0)), 0.1, abs_tol=0.1)


def test_exact_sum_matches_fsum():
    """The fixtures' vectorised exact sums equal math.fsum: wide exponent spread, subnormals,
    cancellation, zeros and an empty array."""
    rng = np.random.default_rng(7)
    cases = [rng.standard_normal(200_000) * 10.0 ** rng.integers(-300, 300, 200_000),
             rng.standard_normal(100_000), np.array([1e308, -1e308, 1e-320, 3.0]),
             np.array([1.0, 1e100, 1.0, -1e100]), np.array([0.1] * 10), np.zeros(5), np.array([])]
    for a in cases:
        assert MG.exact_sum(a) == math.fsum(a.tolist())
        if a.size and np.max(np.abs(a)) < 1e150:   # squares stay finite
            assert MG.exact_sum(np.square(a)) == math.fsum(np.square(a).tolist())


def test_level_means_track_exact_sum(oracle):
    """The level means' fixed tree order against the correctly rounded sums (math.fsum of the same
    terms): within a few units of the last place, where the reference's serial order drifts further
    (its fixture's level means are the serial ones)."""
    g = A.default_grid(1)
    p = A.default_params(g)
    F = MG.case_input("rk4", g, p)
    lev = oracle.hydro(g, p, F.copy())
    t, ps = F[A.FIELD["t"]], F[A.FIELD["p_stat"]]
    wgt = np.sin(np.cumsum(np.r_[0.0, np.full(g.jm - 1, g.dthe)]))[None, :, None] * np.ones_like(t)
    b = p.g * ps / (p.r_mix * p.R_mix * t * p.t_ref)
    ok = (t > 0) & np.isfinite(b)
    exact = np.array([MG.exact_sum((wgt[i] * b[i])[ok[i]]) / MG.exact_sum(wgt[i][ok[i]]) for i in range(g.im)])
    serial = np.load(GOLDEN / "ref_rk4.npz")["level"]
    e_tree = np.max(np.abs(lev - exact) / np.abs(exact))
    e_serial = np.max(np.abs(serial - exact) / np.abs(exact))
    print(f"[level] max relative error vs the exact mean: tree {e_tree:.2e}, reference serial {e_serial:.2e}")
    assert e_tree <= 1e-14 and e_tree < e_serial
