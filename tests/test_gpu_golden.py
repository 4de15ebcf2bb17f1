"""The HIP path against the compiled reference's golden vectors (tests/golden/ref_*.npz) on the
enriched input in which every term of the step is non-zero: domain BC, one RungeKuttaUran (hydro
prep + 4 stages), PressureSolver at 1 and 20 sweeps (alone and with its PRESS_LINE and
PRESS_BAROTROPIC options), two full steps, and the default
PressureSolverUran after one RungeKuttaUran and in two full steps, the Shapiro filter, the
turbulence closure and the loop's saturation-adjustment block. FP64 bar: 1e-12 of each field's largest sampled magnitude, and 1e-12 relative on
the exact field sums; the closure's own arrays (prod, tken, disn, vel_star) the same."""
import sys
from pathlib import Path

import numpy as np
import pytest

import aturan_amd as A

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
import make_golden as MG  # noqa: E402

TOL = 1e-12
# p_hydro magnifies the rounding of the level means, which sum in a fixed tree order here and in the
# reference's serial order there (tests/test_golden_reference.py gives the derivation): 1e-10 of its
# largest sample, its sums 1e-10 of sqrt(N * sum of squares).
PH_TOL = 1e-10
PH = A.FIELD["p_hydro"]


@pytest.mark.parametrize("name", list(MG.CASES))
def test_gpu_matches_compiled_reference(oracle, name):
    print(f"[case] golden {name}", flush=True)
    ref = np.load(GOLDEN / f"ref_{name}.npz")
    g = A.default_grid(1)
    p = A.default_params(g)
    F = MG.case_input(name, g, p)
    mode, steps, sweeps, _ = MG.CASES[name]
    level = np.zeros(g.im)
    turb = None
    pc = A.default_params(g)
    if mode == "bc":
        pc.species_clamp = 0                 # the BC alone; "bc_clamp" runs the default clamp after it
    with A.UranusCore(g, pc) as core:
        core.set_state(F)
        if mode == "satadj":
            core.SaturationAdjustment()
        elif mode in ("turb_run", "turb_init"):
            T3, vs = MG.turb_input(g)
            for n, a in zip(("prod", "tken", "disn"), T3):
                core.set_turb_array(n, a)
            core.set_turb_array("vel_star", vs)
            if mode == "turb_init":
                core.Turbulence_init(**MG.TURB_KNOBS[name])
            else:
                core.Turbulence_run(**MG.TURB_KNOBS[name])
            turb = (np.stack([core.turb_array(n) for n in ("prod", "tken", "disn")]), core.turb_array("vel_star"))
        elif mode in ("bc", "bc_clamp"):
            core.boundary_conditions()
        elif mode == "rk4":
            level = core.computeHydrostaticPressure()
            core.RungeKuttaUran()
        elif mode == "pressure":
            fl = MG.PRESS_FLAGS.get(name, 0)
            core.PressureSolver("reference", sweeps=sweeps, omega=1.0, line=bool(fl & 0x100),
                                barotropic=bool(fl & 0x200))
        elif mode == "rk4_uran":
            core.keep_stage_rhs(True)
            level = core.computeHydrostaticPressure()
            core.RungeKuttaUran()
            core.PressureSolver("uran", sweeps=1)
        elif mode == "damp_t":
            core.damp_wiggles("t", False, True, True, 1.0, 1)
            core.damp_wiggles("t", True, False, False, 0.5, 1)
        elif mode == "damp_uvw":
            for f in ("u", "v", "w"):
                core.damp_wiggles(f, True, True, True, 1.0, sweeps)
        elif mode == "step_uran":
            core.step(steps, "uran", sweeps=1)
        else:
            # p_hydro after the loop is the last step's prep, in the reference and here alike
            core.step(steps, "reference", sweeps=sweeps, omega=1.0)
        G = core.get_state()
    got = MG.summarize(G, level)
    worst = 0.0
    n = G[0].size
    for f, fname in enumerate(A.FIELD_NAMES):
        a, b = got["samples"][f], ref["samples"][f]
        scale = float(np.max(np.abs(b))) or 1.0
        e = float(np.max(np.abs(a - b))) / scale
        if f != PH:
            worst = max(worst, e)
        assert e <= (PH_TOL if f == PH else TOL), (name, fname, e)
        for key in ("sums", "sumsq"):
            x, y = got[key][f], ref[key][f]
            bar = TOL * abs(y)
            if f == PH:
                bar = PH_TOL * ((n * ref["sumsq"][f]) ** 0.5 if key == "sums" else abs(y))
            assert x == y or abs(x - y) <= bar, (name, fname, key, x, y)
    if mode in ("rk4", "rk4_uran"):
        # the level sums run in a fixed parallel order (the oracle's too), not the reference's serial
        # (j, k) order: equal to the reference to rounding
        assert np.max(np.abs(got["level"] - ref["level"])) <= TOL * np.max(np.abs(ref["level"]))
    if turb is not None:
        t = MG.summarize_turb(*turb)
        for q in range(3):
            b = ref["turb_samples"][q]
            e = float(np.max(np.abs(t["turb_samples"][q] - b))) / (float(np.max(np.abs(b))) or 1.0)
            worst = max(worst, e)
            assert e <= TOL, (name, "turb", q, e)
        assert np.array_equal(t["turb_vs_samples"], ref["turb_vs_samples"]), name
        for key in ("turb_sums", "turb_sumsq"):
            x, y = t[key], ref[key]
            assert np.all((x == y) | (np.abs(x - y) <= TOL * np.abs(y))), (name, key, x, y)
    print(f"[golden-gpu] {name}: worst sampled difference {worst:.3e}", flush=True)
