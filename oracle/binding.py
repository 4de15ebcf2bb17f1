"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (oracle.c).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker of the
HIP library. The product (aturan_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_DP = C.POINTER(C.c_double)
_lib = None


def build() -> Path:
    r = subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return LIB


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build()
    lib = C.CDLL(str(LIB))
    V = C.c_void_p
    sig = {
        "orc_cp_table": (C.c_int, [C.c_int, _DP]),
        "orc_cp_lapse": (C.c_double, [V, C.c_double]),
        "orc_rhs": (C.c_int, [V, V, _DP, _DP]),
        "orc_rk4": (C.c_int, [V, V, _DP]),
        "orc_rk4_rhs": (C.c_int, [V, V, _DP, _DP]),
        "orc_pressure_uran": (C.c_int, [V, V, _DP, _DP, C.c_int]),
        "orc_hydro": (C.c_int, [V, V, _DP, _DP]),
        "orc_bc": (C.c_int, [V, V, _DP]),
        "orc_pressure": (C.c_int, [V, V, _DP, C.c_int, C.c_double, _DP]),
        "orc_pressure_flags": (C.c_int, [V, V, _DP, C.c_int, C.c_double, C.c_int, _DP]),
        "orc_residual": (C.c_int, [V, V, _DP, _DP, _DP, _DP]),
        "orc_pressure_converge": (C.c_int, [V, V, _DP, C.c_int, C.c_double, C.POINTER(C.c_int),
                                            C.POINTER(C.c_double), _DP]),
        "orc_step": (C.c_int, [V, V, _DP, C.c_int, C.c_int, C.c_double]),
        "orc_damp_wiggles": (C.c_int, [V, _DP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int]),
        "orc_step_converge": (C.c_int, [V, V, _DP, C.c_int, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int),
                                        C.POINTER(C.c_double)]),
        "orc_turb": (C.c_int, [V, V, V, C.c_int, _DP, _DP, _DP, _DP, _DP]),
        "orc_satadj": (C.c_int, [V, V, V, C.POINTER(C.c_int), C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                 _DP]),
    }
    for n, (res, args) in sig.items():
        f = getattr(lib, n)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _p(x):
    return C.cast(C.pointer(x), C.c_void_p)


def _d(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_DP)


def _ok(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed: {rc}")


def cp_table(mode: int) -> np.ndarray:
    out = np.zeros(2991)
    n = load().orc_cp_table(mode, _d(out))
    assert n == 2991
    return out


def cp_lapse(params, T: float) -> float:
    return load().orc_cp_lapse(_p(params), T)


def rhs(grid, params, fields: np.ndarray) -> np.ndarray:
    """One RHS evaluation: (22, im, jm, km) = 19 tendencies + aux_u, aux_v, aux_w (interior only)."""
    out = np.zeros((22,) + fields.shape[1:])
    _ok(load().orc_rhs(_p(grid), _p(params), _d(fields), _d(out)), "orc_rhs")
    return out


def rk4(grid, params, fields: np.ndarray) -> None:
    _ok(load().orc_rk4(_p(grid), _p(params), _d(fields)), "orc_rk4")


def rk4_rhs(grid, params, fields: np.ndarray) -> np.ndarray:
    """RK4 as rk4(); returns rhs_u, rhs_v, rhs_w of the last stage (3 dense fields)."""
    out = np.zeros((3,) + fields.shape[1:])
    _ok(load().orc_rk4_rhs(_p(grid), _p(params), _d(fields), _d(out)), "orc_rk4_rhs")
    return out


def pressure_uran(grid, params, fields: np.ndarray, rhs3: np.ndarray, calls: int = 1) -> None:
    """PressureSolverUran::run `calls` times (the reference's default solver)."""
    _ok(load().orc_pressure_uran(_p(grid), _p(params), _d(fields), _d(np.ascontiguousarray(rhs3)), calls),
        "orc_pressure_uran")


def hydro(grid, params, fields: np.ndarray) -> np.ndarray:
    lev = np.zeros(grid.im)
    _ok(load().orc_hydro(_p(grid), _p(params), _d(fields), _d(lev)), "orc_hydro")
    return lev


def bc(grid, params, fields: np.ndarray) -> None:
    _ok(load().orc_bc(_p(grid), _p(params), _d(fields)), "orc_bc")


def pressure(grid, params, fields: np.ndarray, sweeps: int, omega: float, flags: int = 0) -> np.ndarray:
    """The red-black PressureSolver; flags 0x100 (PRESS_LINE) and 0x200 (PRESS_BAROTROPIC) add the
    reference's radial line pass and barotropic correction to every sweep."""
    src = np.zeros(fields.shape[1:])
    _ok(load().orc_pressure_flags(_p(grid), _p(params), _d(fields), sweeps, omega, flags, _d(src)),
        "orc_pressure_flags")
    return src


def pressure_converge(grid, params, fields: np.ndarray, max_iters: int, rel_tol: float):
    """Line + barotropic solve to rel_tol; returns (iterations, relative residual, clamped source)."""
    src = np.zeros(fields.shape[1:])
    it, rel = C.c_int(), C.c_double()
    _ok(load().orc_pressure_converge(_p(grid), _p(params), _d(fields), max_iters, rel_tol, C.byref(it), C.byref(rel),
                                     _d(src)), "orc_pressure_converge")
    return it.value, rel.value, src


def residual(grid, params, fields: np.ndarray, src: np.ndarray) -> tuple[float, float]:
    r, s = C.c_double(), C.c_double()
    _ok(load().orc_residual(_p(grid), _p(params), _d(fields), _d(src), C.byref(r), C.byref(s)), "orc_residual")
    return r.value, s.value


def step(grid, params, fields: np.ndarray, nsteps: int, sweeps: int, omega: float) -> None:
    _ok(load().orc_step(_p(grid), _p(params), _d(fields), nsteps, sweeps, omega), "orc_step")


def step_converge(grid, params, fields: np.ndarray, nsteps: int, max_iters: int, rel_tol: float,
                  warm: bool) -> tuple[int, float]:
    """nsteps outer steps with the converging solve (warm 1: start each solve after the first from
    2 p_n - p_{n-1}; 2: after the second from 3 p_n - 3 p_{n-1} + p_{n-2}); returns the last solve's
    iteration count and relative residual."""
    it, rel = C.c_int(), C.c_double()
    _ok(load().orc_step_converge(_p(grid), _p(params), _d(fields), nsteps, max_iters, rel_tol, int(warm),
                                 C.byref(it), C.byref(rel)), "orc_step_converge")
    return it.value, rel.value


def damp_wiggles(grid, fields: np.ndarray, field: int, along_i: bool, along_j: bool, along_k: bool,
                 strength: float = 1.0, passes: int = 1) -> None:
    """AtomUtils::damp_wiggles on one field (no fluid floor, periodic phi seam)."""
    _ok(load().orc_damp_wiggles(_p(grid), _d(fields), field, int(along_i), int(along_j), int(along_k), strength,
                                passes), "orc_damp_wiggles")


class TurbOpts(C.Structure):
    """atu_turb_opts (include/atu.h) with every default written out."""
    _fields_ = [("model", C.c_int32), ("free_flow", C.c_int32), ("abl_taper", C.c_int32), ("curvature", C.c_int32),
                ("k_bg", C.c_double), ("l_bg", C.c_double), ("nue_max", C.c_double), ("vel_star_max", C.c_double),
                ("abl_height", C.c_double)]


_TURB_MODELS = {"k_omega_SST": 0, "k_omega": 1, "k_epsilon": 2}


def turb(grid, params, fields: np.ndarray, prod: np.ndarray, tken: np.ndarray, disn: np.ndarray,
         vel_star: np.ndarray, init: bool = False, model: str = "k_omega_SST", free_flow: bool = True,
         abl_taper: bool = False, curvature: bool = True, k_bg: float = 1.0, l_bg: float = 0.0,
         nue_max: float = 1.0e5, vel_star_max: float = 0.0, abl_height: float = 20000.0) -> None:
    """Turbulence<cUranusModel>::init (init) or ::run on the dense arrays, in place. The defaults are
    the reference's (planet/Turbulence.h:125-156, 349-352; l_bg 0: L_atm / (im-1) in metres)."""
    o = TurbOpts()
    o.model = _TURB_MODELS[model]
    o.free_flow, o.abl_taper, o.curvature = int(free_flow), int(abl_taper), int(curvature)
    o.k_bg = k_bg
    o.l_bg = l_bg if l_bg > 0.0 else params.L_atm * 1.0e3 / float(grid.im - 1)
    o.nue_max = nue_max
    o.vel_star_max = vel_star_max if vel_star_max > 0.0 else float(np.sqrt(0.002)) * 200.0
    o.abl_height = abl_height
    _ok(load().orc_turb(_p(grid), _p(params), _p(o), int(init), _d(fields), _d(prod), _d(tken), _d(disn),
                        _d(vel_star)), "orc_turb")


class Gas(C.Structure):
    """atu_gas (include/atu.h)."""
    _fields_ = [(n, C.c_double) for n in ("t_0", "t_00", "ep", "lv", "ls", "C", "L0", "R", "del_alf", "del_bet",
                                          "C_i", "L0_i", "del_alf_i", "del_bet_i")]


# The constants the reference loop passes (planet/cUranusModel.h:880-906, 1045-1192 and the ice sets
# of planet/SaturationAdjustmentUran.cpp:222-231: H2S takes its liquid set), in atu_gas order.
GASES = {
    "H2O": (273.15, 210.15, 8.1253, 2.5009e6, 2.8339e6, 25.096, 3148.2, 461.52, 0.0, -8.7e-3,
            28.418, 3567.3, 0.0, -8.7e-3),
    "H2S": (187.65, 140.0, 14.5192, 5.5e5, 7.45e5, 17.064, 747.0, 243.96, 0.0, -2.9e-3,
            17.064, 747.0, 0.0, -2.9e-3),
    "NH3": (195.5, 140.0, 7.6752, 1.3720e6, 1.8320e6, 27.863, 2016.0, 488.21, -0.888, 0.0,
            34.948, 2692.5, -0.888, 0.0),
    "CH4": (90.69, 67.36, 7.6752, 5.11e5, 5.11e5, 1.627, 553.1, 518.28, 1.002, -4.1e-3,
            1.627, 553.1, 1.002, -4.1e-3),
}
MOLAR = (17.03052, 34.08088, 51.1114)    # m_nh3, m_h2s, m_nh4sh [kg/kmol]


def satadj(grid, params, fields: np.ndarray, gas_mask: int = 15, nh4sh_eq: bool = True, conserve: int = 1,
           newton: int = 1, evap: int = 1, ice_entry: int = 0, bergeron: int = 2, freeze: int = 1,
           dry_exit: int = 1) -> None:
    """The reference loop's saturation-adjustment block on the dense fields, in place (defaults: the
    reference's)."""
    gases = (Gas * 4)(*[Gas(*GASES[n]) for n in ("H2O", "H2S", "NH3", "CH4")])
    kn = (C.c_int * 7)(conserve, newton, evap, ice_entry, bergeron, freeze, dry_exit)
    _ok(load().orc_satadj(_p(grid), _p(params), C.cast(gases, C.c_void_p), kn, gas_mask, int(nh4sh_eq), *MOLAR,
                          _d(fields)), "orc_satadj")
