"""The step's species clamp is split: its last stage clamps the rows the BC neither writes nor reads
(3 <= i <= im-4 and 3 <= j <= jm-4, every interior k), and a pass after the BC clamps the remaining
shell of rows. That gives the values of clampNegativeSpecies after the BC (planet/BC_Uran.h:110-128,
planet/cUranusModel.cpp:1243-1248) because nothing between the last stage and the BC reads a species,
the BC reads only shell rows (its NEUMANN forms i = 1, 2, im-3, im-2 and j = 1, 2, jm-3, jm-2, the
floor form i = 1, im-2) and copies k = 0 to the seam, and the clamp is idempotent.

Checked here on the oracle alone (no GPU): on inputs whose species hold negative values and NaNs in
every row, BC-then-clamp equals far-rows-clamp, BC without the clamp, shell clamp, bit for bit, for
the default knobs and for the other floor and lid forms. A far region one row larger on every side
fails the comparison (the BC then reads clamped values).
"""
import sys
from pathlib import Path

import numpy as np
import pytest

import aturan_amd as A

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
import make_golden as MG  # noqa: E402

SPECIES = list(range(A.FIELD["ch4"], A.FIELD["nh4sh"] + 1))


def clamp(x):
    return np.where(x >= 0.0, x, 0.0)


def split_clamp_step(oracle, g, p, F, lo=3):
    """far rows clamped (interior k only, as the last stage writes them), BC without the clamp, then
    the shell rows clamped over their whole storage"""
    G = F.copy()
    nk = g.km - 1
    far = (slice(lo, g.im - lo), slice(lo, g.jm - lo))
    for f in SPECIES:
        v = G[f].reshape(g.im, g.jm, -1)
        v[far[0], far[1], :nk] = clamp(v[far[0], far[1], :nk])
    p0 = A.default_params(g)
    for name in ("lid_vw", "rigid_lid", "radius_positive"):
        setattr(p0, name, getattr(p, name))
    p0.species_clamp = 0
    oracle.bc(g, p0, G)
    shell = np.ones((g.im, g.jm), bool)
    shell[far] = False
    for f in SPECIES:
        v = G[f].reshape(g.im, g.jm, -1)
        v[shell] = clamp(v[shell])
    return G


def small_case(seed=7):
    g = A.default_grid(1)
    g.im, g.jm, g.km = 12, 19, 37
    p = A.default_params(g)
    F = MG.enriched_state(g, p)
    MG.species_negatives(F, seed)
    return g, p, F


@pytest.mark.parametrize("radius_positive,lid_vw", [(1, 2), (2, 2), (0, 3), (2, 0)])
def test_split_clamp_equals_clamp_after_bc(oracle, radius_positive, lid_vw):
    g, p, F = small_case()
    p.radius_positive, p.lid_vw = radius_positive, lid_vw
    assert p.species_clamp == 1
    bad = sum(int(np.count_nonzero(~(F[f] >= 0.0))) for f in SPECIES)
    assert bad > 0.05 * len(SPECIES) * F[0].size
    ref = F.copy()
    oracle.bc(g, p, ref)
    got = split_clamp_step(oracle, g, p, F)
    for f in range(F.shape[0]):
        assert np.array_equal(got[f], ref[f], equal_nan=True), f
    for f in SPECIES:
        assert np.all(got[f] >= 0.0)


@pytest.mark.parametrize("lo", [2])
def test_split_clamp_needs_the_shell(oracle, lo):
    """Clamping one more row in i and j before the BC (rows the NEUMANN forms read) changes the
    result: the shell of three rows is what the equivalence needs."""
    g, p, F = small_case()
    ref = F.copy()
    oracle.bc(g, p, ref)
    got = split_clamp_step(oracle, g, p, F, lo=lo)
    assert any(not np.array_equal(got[f], ref[f], equal_nan=True) for f in SPECIES)
