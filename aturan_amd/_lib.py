"""ctypes declarations of the C ABI in include/atu.h and the loader of libaturan_amd.so.

The library is loaded from this package directory only; a missing library raises at import of
the binding (there is no CPU fallback on the product path).

One HIP runtime per process: the library is built against the system ROCm's libamdhip64, and a
framework that bundles its own copy (torch's wheel does, under the same soname) must not have
loaded it first, or the library would run on that copy. load() therefore maps the system runtime
by its plain names first, so that a framework imported afterwards binds to the same copy, and
raises when another copy is already in the process or a second one appears.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libaturan_amd.so"

# ---- field ids (atu_field) ----
FIELD_NAMES = [
    "t", "u", "v", "w",
    "ch4", "ch4_cloud", "ch4_ice", "h2o", "h2o_cloud", "h2o_ice",
    "h2s", "h2s_cloud", "h2s_ice", "nh3", "nh3_cloud", "nh3_ice",
    "nh4sh", "tke", "dis",
    "nue", "rho_mix", "Q_precip",
    "S_precip_ch4", "S_precip_ch4_cloud", "S_precip_ch4_ice",
    "S_precip_h2o", "S_precip_h2o_cloud", "S_precip_h2o_ice",
    "S_precip_h2s", "S_precip_h2s_cloud", "S_precip_h2s_ice",
    "S_precip_nh3", "S_precip_nh3_cloud", "S_precip_nh3_ice",
    "massflux_h2s", "massflux_nh3", "massflux_nh4sh",
    "fluxlim_nh4sh", "tke_source", "dis_source",
    "p_stat", "p_dyn", "p_hydro", "aux_u", "aux_v", "aux_w",
]
FIELD = {n: i for i, n in enumerate(FIELD_NAMES)}
NPROG = 19
NFROZEN = 21
NFIELDS = 46

PREC_FP64 = 0
PREC_MIXED = 1
SOLVE_REFERENCE = 0
SOLVE_CONVERGE = 1
SOLVE_URAN = 2
SOLVE_MODE_MASK = 0xFF
SOLVE_PRESS_LINE = 0x100
SOLVE_PRESS_BAROTROPIC = 0x200

STATUS = {0: "ok", -1: "invalid argument", -2: "HIP error", -3: "out of memory", -4: "non-finite value",
          -5: "no convergence", -6: "unsupported"}


class Grid(C.Structure):
    _fields_ = [("im", C.c_int32), ("jm", C.c_int32), ("km", C.c_int32), ("pad_", C.c_int32),
                ("r0", C.c_double), ("dr", C.c_double), ("the0", C.c_double), ("dthe", C.c_double),
                ("phi0", C.c_double), ("dphi", C.c_double)]


_PARAM_DOUBLES = [
    "dt", "re", "pr", "g", "omega", "u_0", "t_ref", "L_atm",
    "Coriolis", "centrifugal", "buoyancy", "chemical_reaction",
    "sc_ch4", "sc_h2o", "sc_h2s", "sc_nh3", "sc_nh4sh",
    "alpha_sponge", "rho_cond_nh4sh", "r_p_nh4sh",
    "r_mix", "R_mix", "cp_mix", "mue_mix",
    "r_h2", "r_he", "r_ch4", "cp_h2", "cp_he", "cp_ch4",
    "R_metric_km", "sinthe_min", "t_min_K", "t_max_K", "tke_max_m2s2", "dis_min_nd",
    "turb_coupling", "Pr_t", "sig_k", "sig_w", "precip_coupling", "p_dyn_cap", "coriolis_scale",
]
_PARAM_INTS = ["cp_mode", "cp_gas", "turb_active", "lid_vw", "rigid_lid", "radius_positive",
               "species_clamp", "pad_"]


class Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in _PARAM_DOUBLES] + [(n, C.c_int32) for n in _PARAM_INTS]


class SolverOpts(C.Structure):
    _fields_ = [("mode", C.c_int32), ("sweeps", C.c_int32), ("max_sweeps", C.c_int32), ("warm_start", C.c_int32),
                ("omega", C.c_double), ("rel_tol", C.c_double)]


class SolveStats(C.Structure):
    _fields_ = [("sweeps_done", C.c_int32), ("outer_iters", C.c_int32), ("rel_resid", C.c_double),
                ("src_norm", C.c_double)]


class WindowOpts(C.Structure):
    _fields_ = [("nsteps", C.c_int32), ("checkpoint_every", C.c_int32), ("solve_iters", C.c_int32),
                ("rb_sweeps", C.c_int32), ("rel_tol", C.c_double), ("target_amp", C.c_double)]


class TurbOpts(C.Structure):
    _fields_ = [("model", C.c_int32), ("free_flow", C.c_int32), ("abl_taper", C.c_int32), ("curvature", C.c_int32),
                ("k_bg", C.c_double), ("l_bg", C.c_double), ("nue_max", C.c_double), ("vel_star_max", C.c_double),
                ("abl_height", C.c_double)]


class Gas(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("t_0", "t_00", "ep", "lv", "ls", "C", "L0", "R", "del_alf", "del_bet",
                                          "C_i", "L0_i", "del_alf_i", "del_bet_i")]


class SatAdjOpts(C.Structure):
    _fields_ = [("gas", Gas * 4), ("m_nh3", C.c_double), ("m_h2s", C.c_double), ("m_nh4sh", C.c_double),
                ("gas_mask", C.c_int32), ("nh4sh_eq", C.c_int32), ("conserve", C.c_int32), ("newton", C.c_int32),
                ("evap", C.c_int32), ("ice_entry", C.c_int32), ("bergeron", C.c_int32), ("freeze", C.c_int32),
                ("dry_exit", C.c_int32), ("pad_", C.c_int32)]


GASES = {"H2O": 0, "H2S": 1, "NH3": 2, "CH4": 3}
TURB_MODELS = {"k_omega_SST": 0, "k_omega": 1, "k_epsilon": 2}
TURB_ARRAYS = {"prod": 0, "tken": 1, "disn": 2, "vel_star": 3}

_DP = C.POINTER(C.c_double)
_SIGS = {
    "atu_grid_default": (C.c_int, [C.c_int, C.POINTER(Grid)]),
    "atu_params_default": (C.c_int, [C.POINTER(Grid), C.POINTER(Params)]),
    "atu_cp_table": (C.c_int, [C.c_int, _DP, C.POINTER(C.c_int)]),
    "atu_cp_lapse": (C.c_double, [C.POINTER(Params), C.c_double]),
    "atu_synthetic_state": (C.c_int, [C.POINTER(Grid), C.POINTER(Params), C.c_uint64, _DP]),
    "atu_create": (C.c_int, [C.POINTER(Grid), C.POINTER(Params), C.c_int, C.POINTER(C.c_void_p)]),
    "atu_destroy": (None, [C.c_void_p]),
    "atu_last_error": (C.c_char_p, [C.c_void_p]),
    "atu_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "atu_synchronize": (C.c_int, [C.c_void_p]),
    "atu_set_field": (C.c_int, [C.c_void_p, C.c_int, _DP]),
    "atu_get_field": (C.c_int, [C.c_void_p, C.c_int, _DP]),
    "atu_set_field_device": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "atu_get_field_device": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "atu_hydro_prep": (C.c_int, [C.c_void_p, _DP]),
    "atu_rk4": (C.c_int, [C.c_void_p]),
    "atu_apply_bc": (C.c_int, [C.c_void_p]),
    "atu_pressure_solve": (C.c_int, [C.c_void_p, C.POINTER(SolverOpts), C.POINTER(SolveStats)]),
    "atu_step": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(SolverOpts), C.POINTER(SolveStats)]),
    "atu_keep_stage_rhs": (C.c_int, [C.c_void_p, C.c_int]),
    "atu_pressure_residual": (C.c_int, [C.c_void_p, _DP, _DP]),
    "atu_pressure_residual_static": (C.c_int, [C.POINTER(Grid), C.POINTER(Params), C.c_void_p, C.c_void_p, C.c_int,
                                               C.c_void_p, _DP, _DP]),
    "atu_field_device_ptr": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int)]),
    "atu_window_forward": (C.c_int, [C.c_void_p, C.POINTER(WindowOpts), _DP]),
    "atu_window_backward": (C.c_int, [C.c_void_p, C.c_double, _DP, _DP]),
    "atu_window_rewind": (C.c_int, [C.c_void_p]),
    "atu_window_backward_device": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_int, _DP]),
    "atu_window_generation": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "atu_build_info": (C.c_char_p, []),
    "atu_runtime_versions": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "atu_copy_probe": (C.c_int, [C.c_size_t, C.c_int, _DP]),
    "atu_device_memory": (C.c_int, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "atu_damp_wiggles": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int]),
    "atu_satadj_opts_default": (C.c_int, [C.POINTER(SatAdjOpts)]),
    "atu_saturation_adjust": (C.c_int, [C.c_void_p, C.POINTER(SatAdjOpts)]),
    "atu_turb_opts_default": (C.c_int, [C.POINTER(TurbOpts)]),
    "atu_turbulence_init": (C.c_int, [C.c_void_p, C.POINTER(TurbOpts)]),
    "atu_turbulence_run": (C.c_int, [C.c_void_p, C.POINTER(TurbOpts)]),
    "atu_turb_set": (C.c_int, [C.c_void_p, C.c_int, _DP]),
    "atu_turb_get": (C.c_int, [C.c_void_p, C.c_int, _DP]),
    "atu_profile_begin": (C.c_int, [C.c_void_p]),
    "atu_profile_begin_classes": (C.c_int, [C.c_void_p, C.c_char_p]),
    "atu_profile_query": (C.c_int, [C.c_void_p, C.c_char_p, _DP, C.POINTER(C.c_int)]),
    "atu_profile_end": (C.c_int, [C.c_void_p]),
    "atu_profile_sample": (C.c_int, [C.c_void_p, C.c_int]),
}

_lib = None
RUNTIME_PATH: str | None = None   # the libamdhip64 this process runs the library on


def mapped_runtimes() -> list[str]:
    """Real paths of every libamdhip64 mapped into this process."""
    try:
        with open("/proc/self/maps") as f:
            paths = {ln.split()[-1] for ln in f if "libamdhip64" in ln and "/" in ln}
    except OSError:
        return []
    return sorted({os.path.realpath(p) for p in paths})


def load() -> C.CDLL:
    """Load libaturan_amd.so from this package; raise if it has not been built, or if the process
    would run it on a HIP runtime other than the one it was built for."""
    global _lib, RUNTIME_PATH
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: build it with `python aturan_amd/build.py` "
                          "(the HIP library is required; there is no fallback)")
    before = mapped_runtimes()
    if not before:
        # the system runtime under the names a framework's libraries ask for, so that they bind to it
        for name in ("libhsa-runtime64.so", "libamdhip64.so"):
            C.CDLL(name, mode=C.RTLD_GLOBAL)
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    after = mapped_runtimes()
    built, running = C.c_int(0), C.c_int(0)
    rc = lib.atu_runtime_versions(C.byref(built), C.byref(running))
    if len(after) > 1 or (rc == 0 and built.value // 100000 != running.value // 100000):
        raise ImportError(
            f"libaturan_amd needs the HIP runtime it was built for (HIP {built.value // 10000000}."
            f"{built.value // 100000 % 100}); this process maps {after} (running HIP {running.value // 10000000}."
            f"{running.value // 100000 % 100}). Import aturan_amd before any framework that bundles its own ROCm "
            f"runtime (such as torch).")
    RUNTIME_PATH = after[0] if after else None
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return list(_SIGS)
