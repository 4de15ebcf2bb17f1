"""atu_pressure_residual_static, the entry point that allocates nothing, against atu_pressure_residual
of a context on the same arrays: two calls in a row, the second on a smaller grid, each bit-identical
to the context's residual (same weights, same kernels, same fixed-order sums), and the refusal of a
grid whose weights and partials exceed the static array."""
import sys
from pathlib import Path

import numpy as np
import pytest

import aturan_amd as A

pytestmark = pytest.mark.gpu

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
import make_golden as MG  # noqa: E402


def _compare(oracle, g, sweeps):
    p = A.default_params(g)
    F = MG.enriched_state(g, p)
    oracle.bc(g, p, F)
    src = oracle.pressure(g, p, F.copy(), 1, 1.0)      # the clamped source the context forms from aux
    with A.UranusCore(g, p) as core:
        core.set_state(F)
        core.PressureSolver("reference", sweeps=sweeps)
        rel_ctx, sn_ctx = core.pressure_residual()
        core.set_field("aux_u", src)                    # the source, in a device array of the caller's
        pd, ld = core.field_device_ptr("p_dyn")
        sp, ld2 = core.field_device_ptr("aux_u")
        assert ld == ld2
        rel_st, sn_st = A.core.pressure_residual_static(g, p, pd, sp, ld)
    print(f"[no-alloc] {g.im}x{g.jm}x{g.km}: context rel {rel_ctx!r} |src| {sn_ctx!r}; "
          f"static rel {rel_st!r} |src| {sn_st!r}", flush=True)
    assert rel_ctx > 0.0 and sn_ctx > 0.0
    assert rel_st == rel_ctx and sn_st == sn_ctx


def test_static_residual_matches_context_twice(oracle):
    print("[case] atu_pressure_residual_static: 41x181x361 then 12x19x37", flush=True)
    _compare(oracle, A.default_grid(1), 3)
    g = A.default_grid(1)
    g.im, g.jm, g.km = 12, 19, 37
    _compare(oracle, g, 2)


def test_static_residual_refuses_an_oversized_grid():
    print("[case] atu_pressure_residual_static: a grid beyond the static array", flush=True)
    g = A.default_grid(2)
    g.im = 400                                           # 400x361x721: ~14 MB of weights and partials
    p = A.default_params(A.default_grid(2))
    with pytest.raises(A.AtuError, match="bytes of scratch"):
        A.core.pressure_residual_static(g, p, 1 << 40, 1 << 40, g.km)
