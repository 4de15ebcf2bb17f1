"""The C-ABI library loads, exports every symbol include/atu.h declares, and its host-side logic
(grid and parameter defaults, the cp table, the seeded generator) behaves. No GPU work here."""
import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

import aturan_amd as A
from aturan_amd import _lib as L

ROOT = Path(__file__).resolve().parent.parent


def header_symbols():
    text = (ROOT / "include" / "atu.h").read_text()
    return sorted(set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(atu_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(L.LIB_PATH))
    syms = header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # and the ctypes table covers the header exactly
    assert sorted(L.exported_symbols()) == syms


def test_grid_defaults():
    g = A.default_grid(1)
    assert (g.im, g.jm, g.km) == (41, 181, 361)
    assert g.dr == 0.025 and g.r0 == 1.0
    assert g.dthe == 1.0 / (180.0 / math.pi)
    g2 = A.default_grid(2)
    assert (g2.im, g2.jm, g2.km) == (81, 361, 721)


def test_params_mixture_and_anchor():
    g = A.default_grid(1)
    p = A.default_params(g)
    # mass-weighted mixture of planet/ChemistryUran.h:437-490 from the densities of cUranusModel.h
    r = dict(h2=0.864, he=0.136, ch4=0.19, nh3=0.1, h2s=0.3, h2o=0.08, nh4sh=0.007)
    cp = dict(h2=14.32e3, he=5.19e3, nh3=2.19e3, nh4sh=2.00e3, h2s=2.24e3, h2o=1.93e3, ch4=2.232e3)
    Rg = dict(h2=4124.2, he=2077.1, nh3=488.21, nh4sh=261.0, h2s=243.96, h2o=461.52, ch4=518.28)
    rs = sum(r.values())
    assert p.cp_mix == pytest.approx(sum(r[s] * cp[s] for s in r) / rs, rel=1e-14)
    assert p.R_mix == pytest.approx(sum(r[s] * Rg[s] for s in r) / rs, rel=1e-14)
    # the anchor makes r_mix R_mix T_bottom equal the 17.3 bar bottom pressure (cUranusModel.cpp:504-511)
    cpt = A.cp_table(2)

    def cp_lapse(T):
        x = (min(max(T, 5.0), 1500.0) - 5.0) / 0.5
        i = min(len(cpt) - 2, int(x))
        ch = cpt[i] + (x - i) * (cpt[i + 1] - cpt[i])
        return (r["h2"] * ch + r["he"] * cp["he"] + r["ch4"] * cp["ch4"]) / (r["h2"] + r["he"] + r["ch4"])

    T, n = 76.0, 2000
    h = math.log(17.3) / n
    for _ in range(n):
        k1 = T * 3614.91 / cp_lapse(T)
        Tm = T + 0.5 * h * k1
        T += h * Tm * 3614.91 / cp_lapse(Tm)
    assert p.r_mix * p.R_mix * T == pytest.approx(17.3e5, rel=1e-12)
    # the cp(T) adiabat of the H2-He-CH4 gas (equilibrium H2) through 76 K at 1 bar is cooler at
    # 17.3 bar than the constant-cp one (~232 K, planet/cUranusModel.h:132-137)
    assert 180.0 < T < 232.0
    assert p.coriolis_scale == 2250.0
    assert p.turb_active == 1 and p.lid_vw == 2 and p.rigid_lid == 1
    # clampNegativeSpecies after the BC is on by default (planet/BC_Uran.h:110-111)
    assert p.species_clamp == 1 and p.radius_positive == 1


def test_cp_table_known_values(oracle):
    for mode in (1, 2):
        t = A.cp_table(mode)
        assert t.shape == (2991,)
        # planet/CpH2.h:22-23: 298.15 K gives 28.69 J/(mol K) for the rigid rotor
        x = (298.15 - 5.0) / 0.5
        i = int(x)
        cp = t[i] + (x - i) * (t[i + 1] - t[i])
        assert cp * 2.01588e-3 == pytest.approx(28.69, abs=0.01)
        # at 5 K only translation is left: cp = 5/2 R / M
        assert t[0] == pytest.approx(2.5 * 8.314462618 / 2.01588e-3, rel=1e-6)
        # the oracle's independent restatement agrees bit for bit
        assert np.array_equal(t, oracle.cp_table(mode))
    # normal and equilibrium H2 differ at low T (ortho-para conversion), agree at high T
    t1, t2 = A.cp_table(1), A.cp_table(2)
    assert abs(t1[150] - t2[150]) / t1[150] > 0.05       # 80 K
    assert abs(t1[-1] - t2[-1]) / t1[-1] < 1e-3          # 1500 K


def test_cp_lapse_matches_oracle(oracle):
    g = A.default_grid(1)
    p = A.default_params(g)
    lib = A.load()
    for T in (4.0, 5.0, 49.15, 60.65, 76.4, 123.45, 250.0, 1499.9, 1500.0, 2000.0):
        assert lib.atu_cp_lapse(C.byref(p), T) == oracle.cp_lapse(p, T)


def test_synthetic_state_is_seeded_and_shaped():
    g = A.default_grid(1)
    g.im, g.jm, g.km = 11, 21, 37
    p = A.default_params(g)
    a = A.synthetic_state(g, p, 5)
    b = A.synthetic_state(g, p, 5)
    c = A.synthetic_state(g, p, 6)
    assert a.shape == (46, 11, 21, 37)
    assert np.array_equal(a, b)
    assert not np.array_equal(a[A.FIELD["w"]], c[A.FIELD["w"]])
    assert np.array_equal(a[A.FIELD["ch4"]], c[A.FIELD["ch4"]])   # no noise on CH4
    # the species other than CH4, every source and p_dyn start at zero
    for n in ("h2o", "nh4sh", "S_precip_ch4", "massflux_nh3", "p_dyn", "aux_u"):
        assert not a[A.FIELD[n]].any()
    assert np.all(a[A.FIELD["tke"]] == 1e-4)


def test_synthetic_state_profiles():
    g = A.default_grid(1)
    p = A.default_params(g)
    F = A.synthetic_state(g, p, 20261203)
    T = F[A.FIELD["t"]] * p.t_ref
    # top level: 49.15 K at the poles, 60.65 K at the equator, noise sigma 0.05 K
    assert abs(T[-1, 90].mean() - 60.65) < 0.02
    assert abs(T[-1, 0].mean() - 49.15) < 0.02
    assert abs(np.std(T[-1, 90]) - 0.05) < 0.01
    # 0.7 K/km warming downward over 225 km
    assert abs((T[0, 90].mean() - T[-1, 90].mean()) - 0.7 * 225.0) < 0.05
    # zonal wind cos(3 latitude) with 10 % noise
    w = F[A.FIELD["w"]]
    assert abs(w[20, 90].mean() - 1.0) < 0.02
    assert abs(w[20, 30].mean() - math.cos(3 * math.radians(60.0))) < 0.02
    # hydrostatic p_stat: 17.3 bar at the bottom, decreasing upward
    ps = F[A.FIELD["p_stat"]]
    assert ps[0, 90, 0] == 17.3 and np.all(np.diff(ps[:, 90, 0]) < 0)


def test_create_rejects_bad_grid():
    g = A.default_grid(1)
    g.km = 360          # km - 1 odd: the red-black colours would clash across the phi seam
    p = A.default_params(A.default_grid(1))
    ctx = C.c_void_p()
    rc = A.load().atu_create(C.byref(g), C.byref(p), 0, C.byref(ctx))
    assert rc == -1 and not ctx.value
    assert b"invalid grid" in A.load().atu_last_error(None)
    with pytest.raises(A.AtuError):
        A.UranusCore(g, p)


def _mapped_after(code: str):
    import subprocess
    import sys
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300,
                       cwd=str(ROOT))
    return r.returncode, r.stdout, r.stderr


def test_one_hip_runtime_per_process():
    """Loaded before torch, the library maps the system HIP runtime and torch binds to the same copy;
    torch imported first (its wheel bundles its own libamdhip64 under the same soname) makes load()
    refuse instead of running the library on that copy."""
    rc, out, err = _mapped_after("import aturan_amd as A; from aturan_amd import _lib as L; A.load(); import torch; "
                                 "print(L.mapped_runtimes())")
    assert rc == 0, err
    paths = eval(out.strip().splitlines()[-1])
    assert len(paths) == 1 and "torch" not in paths[0], paths
    rc, out, err = _mapped_after("import torch; import aturan_amd as A; A.load()")
    assert rc != 0 and "Import aturan_amd before" in err
