"""The red-black PressureSolver's PRESS_LINE and PRESS_BAROTROPIC options (planet/PressureSolver.h:
284-303, 497-708) on the device against the oracle, whole fields, on the golden input and on the
refined shell.

Bar: bit-identical to the oracle, which equals the compiled reference at one thread (every sum in
the reference's order, the interior mean PRESS_LINE holds included). The reference itself on 8
threads differs from its one-thread result by 1.8e-12 of max |p| after 20 sweeps with both options
(its OpenMP reduction order follows the thread count, and the barotropic solve magnifies a one-ulp
change of the mean): the reason the device sums that mean in the one-thread order.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

import aturan_amd as A

pytestmark = pytest.mark.gpu

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
import make_golden as MG  # noqa: E402

LINE, BARO = 0x100, 0x200
P = A.FIELD["p_dyn"]


def _run_both(oracle, g, p, F, sweeps, flags):
    with A.UranusCore(g, p) as core:
        core.set_state(F)
        core.PressureSolver("reference", sweeps=sweeps, line=bool(flags & LINE), barotropic=bool(flags & BARO))
        G = core.get_field("p_dyn")
    H = F.copy()
    oracle.pressure(g, p, H, sweeps, 1.0, flags)
    return G, H[P]


@pytest.mark.parametrize("flags", [LINE, BARO, LINE | BARO], ids=["line", "baro", "line+baro"])
@pytest.mark.parametrize("sweeps", [1, 20])
def test_press_options_match_oracle(oracle, flags, sweeps):
    print(f"[case] PressureSolver options {flags:#x}, {sweeps} sweeps, 41x181x361", flush=True)
    g = A.default_grid(1)
    p = A.default_params(g)
    F = MG.case_input("pressure_20", g, p)
    G, H = _run_both(oracle, g, p, F, sweeps, flags)
    e = float(np.max(np.abs(G - H))) / float(np.max(np.abs(H)))
    print(f"[parity] options {flags:#x} x{sweeps}: max |gpu - oracle| / max |p| = {e:.3e}", flush=True)
    assert np.array_equal(G, H)


def test_press_options_refined_shell(oracle):
    """81x361x721 (N = 720 meridians, 359 rows, 79-deep columns): both options, 3 sweeps."""
    print("[case] PressureSolver line+baro, 3 sweeps, 81x361x721", flush=True)
    g = A.default_grid(2)
    p = A.default_params(g)
    F = MG.enriched_state(g, p)             # non-zero aux: a non-zero projection source
    oracle.bc(g, p, F)
    G, H = _run_both(oracle, g, p, F, 3, LINE | BARO)
    e = float(np.max(np.abs(G - H))) / float(np.max(np.abs(H)))
    print(f"[parity] refined line+baro x3: {e:.3e}", flush=True)
    assert np.array_equal(G, H)


def test_press_options_refused_outside_reference_mode(oracle):
    g = A.default_grid(1)
    g.im, g.jm, g.km = 12, 19, 37
    p = A.default_params(g)
    F = A.synthetic_state(g, p, 1)
    with A.UranusCore(g, p) as core:
        core.set_state(F)
        with pytest.raises(A.AtuError):
            core.PressureSolver("converge", line=True)
        with pytest.raises(A.AtuError):
            core.PressureSolver("uran", barotropic=True)
        st = core.PressureSolver("reference", sweeps=2, line=True, barotropic=True)
        assert st.sweeps_done == 2
