"""Python host mirror of the reference's hot-path interface over the C ABI (include/atu.h).

The reference exposes this path as parameterless member functions of its model object
(planet/cUranusModel.h:1226-1227: RungeKuttaUran, RHSUran; BC_Uran::bcRadius/bcTheta/bcPhi,
planet/BC_Uran.h:29-41; PressureSolver<cUranusModel>::run, planet/PressureSolver.h:145). `UranusCore`
keeps those names and their order of effects; the arrays it hands out are NumPy copies of the
device fields in the reference's [i][j][k] order (u radial, v meridional +southward, w zonal).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L


class AtuError(RuntimeError):
    """A C-ABI call returned a negative status."""


def _check(rc: int, ctx: C.c_void_p | None, what: str) -> None:
    if rc != 0:
        lib = L.load()
        msg = lib.atu_last_error(ctx if ctx else None)
        raise AtuError(f"{what}: {L.STATUS.get(rc, rc)} ({msg.decode() if msg else ''})")


def default_grid(refine: int = 1) -> L.Grid:
    g = L.Grid()
    _check(L.load().atu_grid_default(refine, C.byref(g)), None, "atu_grid_default")
    return g


def default_params(grid: L.Grid) -> L.Params:
    p = L.Params()
    _check(L.load().atu_params_default(C.byref(grid), C.byref(p)), None, "atu_params_default")
    return p


def synthetic_state(grid: L.Grid, params: L.Params, seed: int = 20261203) -> np.ndarray:
    """All 46 fields of the seeded synthetic state, shape (46, im, jm, km), FP64 (BC not applied)."""
    out = np.zeros((L.NFIELDS, grid.im, grid.jm, grid.km), dtype=np.float64)
    _check(L.load().atu_synthetic_state(C.byref(grid), C.byref(params), C.c_uint64(seed),
                                        out.ctypes.data_as(L._DP)), None, "atu_synthetic_state")
    return out


def runtime_versions() -> tuple[int, int]:
    """(HIP_VERSION the library was built with, hipRuntimeGetVersion of the runtime it runs on)."""
    b, r = C.c_int(), C.c_int()
    _check(L.load().atu_runtime_versions(C.byref(b), C.byref(r)), None, "atu_runtime_versions")
    return b.value, r.value


def device_memory() -> tuple[int, int]:
    """(free, total) bytes of the current device (atu_device_memory)."""
    f, t = C.c_size_t(), C.c_size_t()
    _check(L.load().atu_device_memory(C.byref(f), C.byref(t)), None, "atu_device_memory")
    return f.value, t.value


def copy_probe(nbytes: int, launches: int = 20) -> float:
    """Mean milliseconds of a device-to-device copy of nbytes (atu_copy_probe)."""
    ms = C.c_double()
    _check(L.load().atu_copy_probe(C.c_size_t(nbytes), launches, C.byref(ms)), None, "atu_copy_probe")
    return ms.value


def pressure_residual_static(grid: L.Grid, params: L.Params, p_dyn_dptr: int, src_dptr: int, ld_k: int,
                             stream: int | None = None) -> tuple[float, float]:
    """atu_pressure_residual_static: (relative residual, source norm) of device arrays p_dyn and src
    ([im][jm][ld_k] FP64) without a context, allocating nothing."""
    lib = L.load()
    r, s = C.c_double(), C.c_double()
    _check(lib.atu_pressure_residual_static(C.byref(grid), C.byref(params), C.c_void_p(p_dyn_dptr),
                                            C.c_void_p(src_dptr), int(ld_k), C.c_void_p(stream), C.byref(r),
                                            C.byref(s)), None, "atu_pressure_residual_static")
    return r.value, s.value


def cp_table(mode: int = 2) -> np.ndarray:
    lib = L.load()
    n = C.c_int(0)
    _check(lib.atu_cp_table(mode, None, C.byref(n)), None, "atu_cp_table")
    out = np.zeros(n.value)
    _check(lib.atu_cp_table(mode, out.ctypes.data_as(L._DP), C.byref(n)), None, "atu_cp_table")
    return out


def satadj_opts(gases=("H2O", "H2S", "NH3", "CH4"), nh4sh_eq: bool = True, **knobs) -> L.SatAdjOpts:
    """The saturation-adjustment options: the reference's gas constants and knobs, the listed gases."""
    o = L.SatAdjOpts()
    _check(L.load().atu_satadj_opts_default(C.byref(o)), None, "atu_satadj_opts_default")
    o.gas_mask = sum(1 << L.GASES[g] for g in gases)
    o.nh4sh_eq = int(nh4sh_eq)
    for k, v in knobs.items():
        if k not in ("conserve", "newton", "evap", "ice_entry", "bergeron", "freeze", "dry_exit"):
            raise ValueError(f"unknown saturation-adjustment knob {k}")
        setattr(o, k, int(v))
    return o


def turb_opts(model: str = "k_omega_SST", free_flow: bool = True, abl_taper: bool = False, curvature: bool = True,
              k_bg: float = 0.0, l_bg: float = 0.0, nue_max: float = 0.0, vel_star_max: float = 0.0,
              abl_height: float = 0.0) -> L.TurbOpts:
    """The closure's knobs (the reference's ATURAN_TURB_MODEL, _TURB_FREE, _ABL_TAPER, _TURB_CURV,
    _TURB_K_BG, _TURB_L_BG, _NUE_MAX, _VEL_STAR_MAX); 0 selects the reference's default."""
    o = L.TurbOpts()
    _check(L.load().atu_turb_opts_default(C.byref(o)), None, "atu_turb_opts_default")
    o.model = L.TURB_MODELS[model]
    o.free_flow, o.abl_taper, o.curvature = int(free_flow), int(abl_taper), int(curvature)
    for n, v in (("k_bg", k_bg), ("l_bg", l_bg), ("nue_max", nue_max), ("vel_star_max", vel_star_max),
                 ("abl_height", abl_height)):
        if v > 0.0:
            setattr(o, n, float(v))
    return o


@dataclass
class SolveResult:
    sweeps_done: int
    outer_iters: int
    rel_resid: float
    src_norm: float


def solver_opts(mode: str = "reference", sweeps: int = 1, omega: float = 1.0, rel_tol: float = 1e-8,
                max_sweeps: int = 200, warm_start: int = 0, line: bool = False,
                barotropic: bool = False) -> L.SolverOpts:
    """line / barotropic: the reference solver's PRESS_LINE / PRESS_BAROTROPIC options (mode
    "reference" only; the library refuses them with another mode)."""
    o = L.SolverOpts()
    o.mode = {"reference": L.SOLVE_REFERENCE, "converge": L.SOLVE_CONVERGE, "uran": L.SOLVE_URAN}[mode]
    o.mode |= (L.SOLVE_PRESS_LINE if line else 0) | (L.SOLVE_PRESS_BAROTROPIC if barotropic else 0)
    o.sweeps = sweeps
    o.max_sweeps = max_sweeps
    o.warm_start = int(warm_start)
    o.omega = omega
    o.rel_tol = rel_tol
    return o


class UranusCore:
    """One device context: the model-owned fields and the step pieces of the reference."""

    def __init__(self, grid: L.Grid | None = None, params: L.Params | None = None, precision: str = "fp64"):
        self._lib = L.load()
        self.grid = grid if grid is not None else default_grid(1)
        self.params = params if params is not None else default_params(self.grid)
        prec = {"fp64": L.PREC_FP64, "mixed": L.PREC_MIXED}[precision]
        ctx = C.c_void_p()
        _check(self._lib.atu_create(C.byref(self.grid), C.byref(self.params), prec, C.byref(ctx)), None,
               "atu_create")
        self._ctx = ctx
        self.shape = (self.grid.im, self.grid.jm, self.grid.km)

    # ---- lifetime ----
    def close(self) -> None:
        if getattr(self, "_ctx", None):
            self._lib.atu_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- fields ----
    def set_field(self, name: str | int, arr: np.ndarray) -> None:
        f = L.FIELD[name] if isinstance(name, str) else name
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"field shape {a.shape} != grid {self.shape}")
        _check(self._lib.atu_set_field(self._ctx, f, a.ctypes.data_as(L._DP)), self._ctx, "atu_set_field")

    def get_field(self, name: str | int) -> np.ndarray:
        f = L.FIELD[name] if isinstance(name, str) else name
        out = np.empty(self.shape, dtype=np.float64)
        _check(self._lib.atu_get_field(self._ctx, f, out.ctypes.data_as(L._DP)), self._ctx, "atu_get_field")
        return out

    def set_field_device(self, name: str | int, dptr: int, ld_k: int | None = None) -> None:
        """Copy a dense FP64 device array [im][jm][ld_k] (an address in this device's memory, e.g. a
        tensor's data_ptr()) into a field (atu_set_field_device)."""
        f = L.FIELD[name] if isinstance(name, str) else name
        _check(self._lib.atu_set_field_device(self._ctx, f, C.c_void_p(dptr), ld_k or self.grid.km), self._ctx,
               "atu_set_field_device")

    def get_field_device(self, name: str | int, dptr: int, ld_k: int | None = None) -> None:
        """Copy a field into a dense FP64 device array [im][jm][ld_k] (atu_get_field_device)."""
        f = L.FIELD[name] if isinstance(name, str) else name
        _check(self._lib.atu_get_field_device(self._ctx, f, C.c_void_p(dptr), ld_k or self.grid.km), self._ctx,
               "atu_get_field_device")

    def set_state(self, fields: np.ndarray) -> None:
        for f in range(L.NFIELDS):
            self.set_field(f, fields[f])

    def get_state(self) -> np.ndarray:
        return np.stack([self.get_field(f) for f in range(L.NFIELDS)])

    # ---- the step pieces, named as in the reference ----
    def computeHydrostaticPressure(self) -> np.ndarray:
        """computeBuoyancyRefLevel + computeHydrostaticPressure; returns buoy_ref_level."""
        lev = np.zeros(self.grid.im)
        _check(self._lib.atu_hydro_prep(self._ctx, lev.ctypes.data_as(L._DP)), self._ctx, "atu_hydro_prep")
        return lev

    def RungeKuttaUran(self) -> None:
        _check(self._lib.atu_rk4(self._ctx), self._ctx, "atu_rk4")

    def keep_stage_rhs(self, on: bool = True) -> None:
        """Keep rhs_u, rhs_v, rhs_w of the last RK stage (read by the "uran" solve)."""
        _check(self._lib.atu_keep_stage_rhs(self._ctx, int(on)), self._ctx, "atu_keep_stage_rhs")

    def boundary_conditions(self) -> None:
        """BC_Uran(m).bcRadius(); bcTheta(); bcPhi() on the prognostic fields."""
        _check(self._lib.atu_apply_bc(self._ctx), self._ctx, "atu_apply_bc")

    def PressureSolver(self, mode: str = "reference", sweeps: int = 1, omega: float = 1.0,
                       rel_tol: float = 1e-8, max_sweeps: int = 200, line: bool = False,
                       barotropic: bool = False) -> SolveResult:
        """PressureSolver::run: the projection source, then the solve of `mode`; with mode "reference",
        `line` and `barotropic` switch on the reference's PRESS_LINE and PRESS_BAROTROPIC."""
        o = solver_opts(mode, sweeps, omega, rel_tol, max_sweeps, line=line, barotropic=barotropic)
        st = L.SolveStats()
        _check(self._lib.atu_pressure_solve(self._ctx, C.byref(o), C.byref(st)), self._ctx, "atu_pressure_solve")
        return SolveResult(st.sweeps_done, st.outer_iters, st.rel_resid, st.src_norm)

    def step(self, nsteps: int = 1, mode: str = "reference", sweeps: int = 1, omega: float = 1.0,
             rel_tol: float = 1e-8, max_sweeps: int = 200, warm_start: int = 0, line: bool = False,
             barotropic: bool = False) -> SolveResult:
        """The benchmark's outer step: hydro prep, RK4, BC, Poisson solve; nsteps times (converge with
        warm_start 1: each solve after the first starts from 2 p_n - p_{n-1}; 2: after the second from
        3 p_n - 3 p_{n-1} + p_{n-2})."""
        o = solver_opts(mode, sweeps, omega, rel_tol, max_sweeps, warm_start, line, barotropic)
        st = L.SolveStats()
        _check(self._lib.atu_step(self._ctx, nsteps, C.byref(o), C.byref(st)), self._ctx, "atu_step")
        return SolveResult(st.sweeps_done, st.outer_iters, st.rel_resid, st.src_norm)

    def damp_wiggles(self, field: str | int, along_i: bool, along_j: bool, along_k: bool, strength: float = 1.0,
                     passes: int = 1) -> None:
        """AtomUtils::damp_wiggles on one prognostic field (atu_damp_wiggles)."""
        f = L.FIELD[field] if isinstance(field, str) else int(field)
        _check(self._lib.atu_damp_wiggles(self._ctx, f, int(along_i), int(along_j), int(along_k), strength, passes),
               self._ctx, "atu_damp_wiggles")

    # ---- the saturation-adjustment block of the reference loop ----
    def SaturationAdjustment(self, gases=("H2O", "H2S", "NH3", "CH4"), nh4sh_eq: bool = True, **knobs) -> None:
        """The loop's four SaturationAdjustmentUran calls (in the loop's order, for the listed gases)
        and ChemistryUran::EquilibriumNH4SH("loop") (planet/cUranusModel.cpp:997-1033). Knobs: conserve,
        newton, evap, ice_entry, bergeron, freeze, dry_exit (the reference's ATURAN_SATADJ_*)."""
        o = satadj_opts(gases, nh4sh_eq, **knobs)
        _check(self._lib.atu_saturation_adjust(self._ctx, C.byref(o)), self._ctx, "atu_saturation_adjust")

    # ---- the turbulence closure, Turbulence<cUranusModel> (planet/Turbulence.h) ----
    def Turbulence_init(self, **knobs) -> None:
        """Turbulence::init (planet/Turbulence.h:282-303): seed tke and dis, then fill nue, prod and the
        sources. Knobs as turb_opts()."""
        o = turb_opts(**knobs)
        _check(self._lib.atu_turbulence_init(self._ctx, C.byref(o)), self._ctx, "atu_turbulence_init")

    def Turbulence_run(self, **knobs) -> None:
        """Turbulence::run (planet/Turbulence.h:306-332), the loop's call after restoreVar
        (planet/cUranusModel.cpp:1270). Knobs as turb_opts()."""
        o = turb_opts(**knobs)
        _check(self._lib.atu_turbulence_run(self._ctx, C.byref(o)), self._ctx, "atu_turbulence_run")

    def turb_array(self, name: str) -> np.ndarray:
        """The closure's own arrays: prod, tken, disn ([im][jm][km]) or vel_star ([jm][km])."""
        g = self.grid
        shape = (g.jm, g.km) if name == "vel_star" else (g.im, g.jm, g.km)
        out = np.empty(shape, dtype=np.float64)
        _check(self._lib.atu_turb_get(self._ctx, L.TURB_ARRAYS[name], out.ctypes.data_as(L._DP)), self._ctx,
               "atu_turb_get")
        return out

    def set_turb_array(self, name: str, values: np.ndarray) -> None:
        g = self.grid
        shape = (g.jm, g.km) if name == "vel_star" else (g.im, g.jm, g.km)
        a = np.ascontiguousarray(values, dtype=np.float64)
        if a.shape != shape:
            raise ValueError(f"{name}: expected shape {shape}, got {a.shape}")
        _check(self._lib.atu_turb_set(self._ctx, L.TURB_ARRAYS[name], a.ctypes.data_as(L._DP)), self._ctx,
               "atu_turb_set")

    def field_device_ptr(self, name: str | int) -> tuple[int, int]:
        """(device address, ld_k) of one field (atu_field_device_ptr)."""
        f = L.FIELD[name] if isinstance(name, str) else int(name)
        ptr, ld = C.c_void_p(), C.c_int()
        _check(self._lib.atu_field_device_ptr(self._ctx, f, C.byref(ptr), C.byref(ld), None), self._ctx,
               "atu_field_device_ptr")
        return int(ptr.value), ld.value

    def pressure_residual(self) -> tuple[float, float]:
        r, s = C.c_double(), C.c_double()
        _check(self._lib.atu_pressure_residual(self._ctx, C.byref(r), C.byref(s)), self._ctx,
               "atu_pressure_residual")
        return r.value, s.value

    def synchronize(self) -> None:
        _check(self._lib.atu_synchronize(self._ctx), self._ctx, "atu_synchronize")

    # ---- gradients over a window of steps ----
    def window_forward(self, nsteps: int, checkpoint_every: int = 10, solve_iters: int = 0,
                       rel_tol: float = 1e-10, target_amp: float = 1.2, rb_sweeps: int = 0) -> float:
        """Run `nsteps` steps from the current state, checkpointing every `checkpoint_every` steps, and
        return the zonal-wind misfit L of the final state (atu_window_forward). rb_sweeps > 0: each
        solve is that many red-black sweeps of the reference, differentiated sweep by sweep."""
        o = L.WindowOpts()
        o.nsteps, o.checkpoint_every, o.solve_iters, o.rb_sweeps = nsteps, checkpoint_every, solve_iters, rb_sweeps
        o.rel_tol, o.target_amp = rel_tol, target_amp
        loss = C.c_double()
        _check(self._lib.atu_window_forward(self._ctx, C.byref(o), C.byref(loss)), self._ctx, "atu_window_forward")
        return loss.value

    def window_backward(self, dloss: float = 1.0) -> tuple[np.ndarray, float]:
        """dloss * dL/dT0 (dense [im][jm][km]) and dloss * dL/dnue of the last window_forward."""
        g = np.empty((self.grid.im, self.grid.jm, self.grid.km), dtype=np.float64)
        gn = C.c_double()
        _check(self._lib.atu_window_backward(self._ctx, float(dloss), g.ctypes.data_as(L._DP), C.byref(gn)),
               self._ctx, "atu_window_backward")
        return g, gn.value

    def window_backward_device(self, dloss: float, grad_t0_dptr: int, ld_k: int | None = None) -> float:
        """As window_backward, dL/dT0 written to a dense FP64 device array [im][jm][ld_k]; returns
        dloss * dL/dnue."""
        gn = C.c_double()
        _check(self._lib.atu_window_backward_device(self._ctx, float(dloss), C.c_void_p(grad_t0_dptr),
                                                    ld_k or self.grid.km, C.byref(gn)),
               self._ctx, "atu_window_backward_device")
        return gn.value

    def window_generation(self) -> int:
        """Identifies the forward the next window_backward differentiates (0: none, or its inputs changed)."""
        g = C.c_uint64()
        _check(self._lib.atu_window_generation(self._ctx, C.byref(g)), self._ctx, "atu_window_generation")
        return g.value

    def window_rewind(self) -> None:
        """State (prognostic fields, p_dyn) back to the start of the last window_forward."""
        _check(self._lib.atu_window_rewind(self._ctx), self._ctx, "atu_window_rewind")

    # ---- per-kernel timing ----
    def profile_begin(self, classes: str | None = None, every: int = 1) -> None:
        """Bracket kernel launches with HIP events: every class, or the comma-separated `classes`; in
        step(), every `every`-th step only."""
        _check(self._lib.atu_profile_sample(self._ctx, every), self._ctx, "atu_profile_sample")
        if classes is None:
            _check(self._lib.atu_profile_begin(self._ctx), self._ctx, "atu_profile_begin")
        else:
            _check(self._lib.atu_profile_begin_classes(self._ctx, classes.encode()), self._ctx,
                   "atu_profile_begin_classes")

    def profile_query(self, cls: str) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_int()
        _check(self._lib.atu_profile_query(self._ctx, cls.encode(), C.byref(ms), C.byref(n)), self._ctx,
               "atu_profile_query")
        return ms.value, n.value

    def profile_end(self) -> None:
        _check(self._lib.atu_profile_end(self._ctx), self._ctx, "atu_profile_end")
