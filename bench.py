"""Benchmark of the dynamical-core step on one MI355X.

A step is one outer time step of the 41x181x361 shell with every input resident in HBM: the
hydrostatic prep, the four fused RK4 stages, the domain BC and the red-black Poisson solve of p_dyn
to a relative residual of 1e-8. `value` is the plain FP64 GPU step (the benchmark's headline
configuration); the same line carries the mixed-precision step (FP32 stencils, stage states and
smoothing; FP64 accumulation, residual and refinement) with its per-field error against the FP64
run after 1, 10 and 100 steps, the roofline of the dominant kernel measured with HIP events on the
library's own stream, the HBM bandwidth a plain copy kernel reaches on the same device, the CPU
oracle timed on a bounded sample of the same step with the GPU's state after the same steps compared
against it, and the reference's CPU configuration (one RK4 step, then 20 red-black SOR sweeps at
omega = 1.5) timed on both.

The process loads no framework besides the library: the library runs on the HIP runtime it was
built for, which the line records.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBS = 8000.0          # MI355X HBM3E, vendor figure
N_INT = 39 * 179 * 360         # integrated interior cells
# Algorithmic HBM bytes of the four fused RK stages per interior cell: array passes (stage 0: 80,
# stages 1-2: 118 each, stage 3: 102) x bytes per value of each array.
RK_PASSES_FP64 = (80, 118, 118, 102)
# mixed: stencil states, frozen inputs and aux in FP32; p_dyn, p_hydro and acc in FP64; stages 1-2
# form the next input from y_n's FP32 mirror, stage 3 the update from FP64 y_n
_MX_READ = 40 * 4 + 2 * 8                                        # 19 states + 21 frozen (4 B), p_dyn, p_hydro
RK_BYTES_MIXED = (
    _MX_READ + 19 * 8 + 19 * 4,                                  # stage 0: writes acc, next input
    _MX_READ + 19 * 4 + 19 * 8 + 19 * 8 + 19 * 4,                # stages 1-2: + y_n mirror, acc read
    _MX_READ + 19 * 4 + 19 * 8 + 19 * 8 + 19 * 4,
    _MX_READ + 19 * 8 + 19 * 8 + 19 * 8 + 19 * 4 + 3 * 4,        # stage 3: y_n, acc; writes y, its mirror, aux
)
SOLVE_START = {0: "p_n", 1: "2 p_n - p_(n-1) (linear extrapolation of the last two solutions)",
               2: "3 p_n - 3 p_(n-1) + p_(n-2) (quadratic extrapolation of the last three solutions)"}
ERR_FIELDS = {"u_zonal(w)": "w", "v": "v", "w_radial(u)": "u", "T": "t", "CH4": "ch4", "p": "p_dyn"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tol", type=float, default=1e-8, help="relative residual of the Poisson solve")
    ap.add_argument("--inner", type=int, default=1, help="FP32 smoothing iterations per refinement (mixed)")
    ap.add_argument("--mode", choices=["both", "fp64", "mixed"], default="both")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-errors", action="store_true", help="skip the 1/10/100-step mixed-vs-FP64 comparison")
    ap.add_argument("--cold-start", action="store_true", help="start every solve from p_n (= --warm 0)")
    ap.add_argument("--warm", type=int, choices=[0, 1, 2], default=2,
                    help="solve start: 0 p_n, 1 2 p_n - p_(n-1), 2 3 p_n - 3 p_(n-1) + p_(n-2) (default)")
    ap.add_argument("--cpu-one-thread", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--config1-one-thread", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-events", action="store_true",
                    help="no HIP events in the timed region and no per-class pass (for runs under a profiler)")
    ap.add_argument("--adjoint-steps", type=int, default=50, help="window of the adjoint timing (0: skip)")
    ap.add_argument("--checkpoint-every", type=int, default=10)
    ap.add_argument("--turb-calls", type=int, default=50, help="turbulence-closure calls timed (0: skip)")
    ap.add_argument("--satadj-calls", type=int, default=20, help="saturation-adjustment calls timed (0: skip)")
    ap.add_argument("--refined", action="store_true",
                    help="the 81x361x721 workload instead: mixed vs FP64 over --refined-steps steps and the "
                         "mixed vs FP64 adjoint over --refined-window steps")
    ap.add_argument("--refined-steps", type=int, default=200)
    ap.add_argument("--refined-window", type=int, default=100)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device-to-device copy of known size (calibrates the HBM byte counters)")
    ap.add_argument("--config1-steps", type=int, default=50,
                    help="GPU steps of the reference's CPU configuration (RK4 + 20 SOR sweeps, omega 1.5); 0: skip")
    return ap.parse_args()


COPY_BYTES = 512 * 1024 * 1024               # per buffer: far beyond the 256 MB Infinity Cache


def copy_calibration():
    """The library's plain copy kernel (k_copy_probe) over a known number of bytes, for scaling the
    FETCH_SIZE / WRITE_SIZE counters."""
    import aturan_amd as A
    ms = A.core.copy_probe(COPY_BYTES, 3)
    print(json.dumps({"copy_calibration": {"kernel": "k_copy_probe", "bytes_read_per_launch": COPY_BYTES,
                                           "bytes_written_per_launch": COPY_BYTES, "launches": 4,
                                           "ms_per_launch": ms}}), flush=True)


def hbm_copy():
    """The HBM bandwidth a plain streaming copy reaches on this device (read + write bytes / time)."""
    import aturan_amd as A
    nbytes = 4 * COPY_BYTES
    ms = A.core.copy_probe(nbytes, 20)
    return {"kernel": "k_copy_probe", "bytes_per_buffer": nbytes, "ms_per_copy": round(ms, 4),
            "GBps": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}


TRAFFIC_FILE = "profiles/traffic_rk_stage.json"
EVENT_EVERY = 5   # the timed region brackets the kernels of one step in five with events


def traffic_from_profiles():
    """HBM bytes per stage launch as the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary gives them
    (not measured in this run), with the commit the summary was made at."""
    f = ROOT / TRAFFIC_FILE
    if f.exists():
        try:
            return json.loads(f.read_text())
        except ValueError:
            return None
    return None


def runtime_record():
    import aturan_amd as A
    from aturan_amd import _lib as L
    built, running = A.core.runtime_versions()
    return {"libamdhip64": L.RUNTIME_PATH, "hip_built": built, "hip_running": running,
            "mapped_runtimes": L.mapped_runtimes()}


def check_upload(core, F):
    """The state read back from the device equals what was uploaded (a guard before timing)."""
    import numpy as np
    import aturan_amd as A
    for n in ("t", "w", "ch4", "nue", "p_stat"):
        if not np.array_equal(core.get_field(n), F[A.FIELD[n]]):
            raise RuntimeError(f"bench: field {n} read back from the device differs from the upload")


def initial_state(refine=1):
    """The seeded synthetic state with the domain BC applied once (by the library, on the GPU)."""
    import aturan_amd as A
    g = A.default_grid(refine)
    p = A.default_params(g)
    F = A.synthetic_state(g, p, 20261203)
    with A.UranusCore(g, p) as core:
        core.set_state(F)
        core.boundary_conditions()
        F = core.get_state()
    return g, p, F


PROFILE_CLASSES = ("rk_stage", "line", "baro", "residual", "hydro", "bc", "proj_src")


def timed_run(core, steps, warmup, opts, events=True, every=EVENT_EVERY):
    """The timed region: the stage kernels and the whole steps of every `every`-th step are bracketed
    by HIP events (for the roofline and the stream time per step); the other steps run without the
    events, whose markers add ~11 us between two kernels."""
    for _ in range(warmup):
        core.step(1, **opts)
    core.synchronize()
    if events:
        core.profile_begin("rk_stage,step", every=every)
    t0 = time.perf_counter()
    st = core.step(steps, **opts)
    core.synchronize()
    el = time.perf_counter() - t0
    if not events:
        return el, st, (0.0, 0), 0.0
    rk = core.profile_query("rk_stage")
    step_ms, step_n = core.profile_query("step")
    core.profile_end()
    return el, st, rk, (step_ms / step_n if step_n else 0.0)


def kernel_breakdown(core, steps, opts):
    """Milliseconds per step of each kernel class, every launch bracketed by events (a separate,
    untimed pass: the event pairs between kernels add their own gaps)."""
    core.profile_begin()
    core.step(steps, **opts)
    prof = {c: core.profile_query(c) for c in PROFILE_CLASSES}
    core.profile_end()
    return {k: round(v[0] / steps, 4) for k, v in prof.items()}


def rel_errors(a, b):
    import numpy as np
    out = {}
    for name, f in ERR_FIELDS.items():
        x, y = a[f][1:-1, 1:-1, :-1], b[f][1:-1, 1:-1, :-1]
        d = x - y
        out[name] = {"max": float(np.max(np.abs(d)) / max(np.max(np.abs(y)), 1e-300)),
                     "rms": float(np.sqrt(np.mean(d * d)) / max(np.sqrt(np.mean(y * y)), 1e-300))}
    return out


def mixed_errors(g, p, F, opts64, optsmx):
    """Mixed vs FP64 GPU after 1, 10 and 100 steps from the same state."""
    import aturan_amd as A
    res = {}
    with A.UranusCore(g, p, "fp64") as c64, A.UranusCore(g, p, "mixed") as cmx:
        c64.set_state(F)
        cmx.set_state(F)
        done = 0
        for n in (1, 10, 100):
            c64.step(n - done, **opts64)
            cmx.step(n - done, **optsmx)
            done = n
            a = {f: cmx.get_field(f) for f in set(ERR_FIELDS.values())}
            b = {f: c64.get_field(f) for f in set(ERR_FIELDS.values())}
            res[str(n)] = rel_errors(a, b)
    return res


def turb_bytes(g) -> int:
    """Algorithmic HBM bytes of one Turbulence::run (k-omega SST, TURB_FREE 1) on the device: the source
    loop reads u, v, w, tke, dis and writes prod, nue and the two sources per interior cell; the clamp
    pass reads and writes tke and dis; the friction velocity, the edge columns, the seam and the lid
    are the small rest."""
    im, jm, km = g.im, g.jm, g.km
    n_int = (im - 2) * (jm - 2) * (km - 1)
    cols_int = (jm - 2) * (km - 1)
    b = n_int * (9 * 8 + 4 * 8)                   # k_turb_sources, k_turb_clamp (interior)
    b += cols_int * 2 * 16 + jm * km * 8 * 2      # clamp of the end cells; vel_star write + read
    b += (2 * km + jm - 2) * (im * 16 + 6 * 8)    # k_turb_edges
    b += 8 * im * jm * 16                         # k_turb_seam
    b += 8 * (jm + 2 * (im - 1)) * km * 16        # k_turb_lid
    return b


def turb_bench(g, p, F, calls, events=True):
    """Turbulence::run (k-omega SST, the reference's defaults) on the step's state: wall time per
    call (which includes refreshing the stage buffers of tke and dis) and the closure's kernels."""
    import aturan_amd as A
    with A.UranusCore(g, p, "fp64") as core:
        core.set_state(F)
        core.Turbulence_init()
        core.Turbulence_run()
        core.synchronize()
        t0 = time.perf_counter()
        for _ in range(calls):
            core.Turbulence_run()
        core.synchronize()
        wall = (time.perf_counter() - t0) / calls
        kms = None
        if events:
            core.profile_begin("turbulence")
            for _ in range(calls):
                core.Turbulence_run()
            kms = core.profile_query("turbulence")[0] / calls
            core.profile_end()
    b = turb_bytes(g)
    out = {"closure": "k-omega SST, TURB_FREE 1 (Turbulence<cUranusModel>::run)", "calls": calls,
           "ms_per_call": round(wall * 1e3, 4), "algorithmic_bytes_per_call": b}
    if kms:
        out.update(kernels_ms_per_call=round(kms, 4), kernels_GBps=round(b / (kms * 1e-3) / 1e9, 1),
                   kernels_frac=round(b / (kms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4))
    return out


def satadj_state(g, p, F):
    """The step's state with clouds: every gas's vapour log-uniform in 1e-9..1e-3 kg/m3 (sub- and
    supersaturated cells), its cloud and ice 1e-5..1e-3 kg/m3 in a quarter of the cells each and both
    in another quarter, NH4SH in half of them (seeded)."""
    import aturan_amd as A
    import numpy as np
    rng = np.random.default_rng(20261203)
    F = F.copy()
    shp = F.shape[1:]
    for gas in ("h2o", "h2s", "nh3", "ch4"):
        F[A.FIELD[gas]] = 10.0 ** rng.uniform(-9.0, -3.0, shp)
        sel = rng.integers(0, 4, shp)
        F[A.FIELD[gas + "_cloud"]] = np.where((sel == 1) | (sel == 3), 10.0 ** rng.uniform(-5.0, -3.0, shp), 0.0)
        F[A.FIELD[gas + "_ice"]] = np.where((sel == 2) | (sel == 3), 10.0 ** rng.uniform(-5.0, -3.0, shp), 0.0)
    F[A.FIELD["nh4sh"]] = np.where(rng.random(shp) < 0.5, 10.0 ** rng.uniform(-6.0, -3.0, shp), 0.0)
    return F


def satadj_bench(g, p, F, calls, events=True):
    """The loop's saturation-adjustment block (four gases + NH4SH equilibrium), each call from the
    same cloudy state re-uploaded outside the timed kernels (events) or included (wall)."""
    import aturan_amd as A
    Fs = satadj_state(g, p, F)
    nbytes = 30 * 8 * g.im * g.jm * g.km      # 16 fields read, 14 written
    out = {"block": "SaturationAdjustmentUran x 4 (H2O, H2S, NH3, CH4) + EquilibriumNH4SH(loop)",
           "state": "step state with seeded clouds (vapour 1e-9..1e-3 kg/m3, mixed phases)", "calls": calls,
           "algorithmic_bytes_per_call": nbytes}
    with A.UranusCore(g, p, "fp64") as core:
        total = 0.0
        kms = 0.0
        for _ in range(calls):
            core.set_state(Fs)
            core.synchronize()
            if events:
                core.profile_begin("satadj")
            t0 = time.perf_counter()
            core.SaturationAdjustment()
            core.synchronize()
            total += time.perf_counter() - t0
            if events:
                kms += core.profile_query("satadj")[0]
                core.profile_end()
    out["ms_per_call"] = round(total / calls * 1e3, 4)
    if events:
        out["kernel_ms_per_call"] = round(kms / calls, 4)
        out["cells_per_s"] = round(g.im * g.jm * g.km / (kms / calls * 1e-3), 1)
        out["GBps"] = round(nbytes / (kms / calls * 1e-3) / 1e9, 1)
    return out


def adjoint_bench(g, p, F, nsteps, K, tol, fwd_step_s, events=True):
    """Forward window (with checkpoints) and hand-written adjoint (recompute + reverse) timings."""
    import aturan_amd as A
    with A.UranusCore(g, p, "fp64") as core:
        core.set_state(F)
        core.window_forward(min(2, nsteps), checkpoint_every=min(2, K), rel_tol=tol)   # allocations, warm-up
        core.window_backward(1.0)
        core.window_rewind()
        core.synchronize()
        t0 = time.perf_counter()
        loss = core.window_forward(nsteps, checkpoint_every=K, rel_tol=tol)
        core.synchronize()
        t1 = time.perf_counter()
        g_t, g_nue = core.window_backward(1.0)
        core.synchronize()
        t2 = time.perf_counter()
        prof_ms = None
        if events:
            core.window_rewind()
            core.window_forward(nsteps, checkpoint_every=K, rel_tol=tol)
            core.profile_begin()
            core.window_backward(1.0)
            prof_ms = core.profile_query("adjoint")[0]
            core.profile_end()
    fwd, bwd = t1 - t0, t2 - t1
    return {"window_steps": nsteps, "checkpoint_every": K, "loss": loss, "forward_s": round(fwd, 4),
            "backward_s": round(bwd, 4), "backward_over_forward": round(bwd / fwd, 3),
            "backward_over_plain_steps": round(bwd / (nsteps * fwd_step_s), 3) if fwd_step_s else None,
            "adjoint_kernels_ms": round(prof_ms, 2) if prof_ms is not None else None,
            "fwd_plus_adjoint_steps_per_s": round(nsteps / (fwd + bwd), 3),
            "fwd_plus_adjoint_cell_updates_per_s": round(nsteps / (fwd + bwd) * (g.im - 2) * (g.jm - 2) * (g.km - 1), 1),
            "grad_t0_norm": float((g_t * g_t).sum() ** 0.5), "grad_nue": g_nue,
            "checked_by": "tests/test_gpu_adjoint.py::test_adjoint_full_shell_tolerance_solves (this window, "
                          "checkpoint every 10, solves to 1e-8: central FD along 4 directions and in nue, bar 1e-4)",
            "memory": f"checkpoints {(nsteps + K - 1) // K + 1} x 20 fields + tape {K} x 19 fields of "
                      f"{g.im}x{g.jm}x384 FP64"}


def refined_bench(args):
    """81x361x721 (dr/2, 0.5 deg): mixed vs FP64 forward over `refined_steps` steps, and the window
    adjoint (checkpoint every 10) in mixed vs FP64 over `refined_window` steps."""
    import numpy as np
    import aturan_amd as A
    g, p, F = initial_state(2)
    n_int = (g.im - 2) * (g.jm - 2) * (g.km - 1)
    warm = 0 if args.cold_start else args.warm
    opts64 = dict(mode="converge", rel_tol=args.tol, max_sweeps=60, warm_start=warm)
    optsmx = dict(mode="converge", rel_tol=args.tol, max_sweeps=60, sweeps=args.inner, warm_start=warm)
    out, fields = {}, {}

    def progress(msg):
        print(f"[refined] {msg}", file=sys.stderr, flush=True)

    progress("initial state ready")
    for prec, opts in (("fp64", opts64), ("mixed", optsmx)):
        with A.UranusCore(g, p, prec) as core:
            core.set_state(F)
            core.step(2, **opts)
            core.set_state(F)
            core.synchronize()
            t0 = time.perf_counter()
            st = core.step(args.refined_steps, **opts)
            core.synchronize()
            el = time.perf_counter() - t0
            fields[prec] = {f: core.get_field(f) for f in set(ERR_FIELDS.values())}
            progress(f"{prec}: {args.refined_steps} steps in {el:.2f} s")
        out[prec] = {"steps_per_s": round(args.refined_steps / el, 3), "ms_per_step": round(1e3 * el / args.refined_steps, 3),
                     "cell_updates_per_s": round(n_int * args.refined_steps / el, 1),
                     "solve_iterations_last_step": st.sweeps_done, "solve_rel_resid_last_step": st.rel_resid}
    out["mixed"]["speedup_vs_fp64"] = round(out["mixed"]["steps_per_s"] / out["fp64"]["steps_per_s"], 4)
    out["mixed"]["rel_error_vs_fp64"] = rel_errors(fields["mixed"], fields["fp64"])
    del fields
    grads = {}
    for prec in ("fp64", "mixed"):
        with A.UranusCore(g, p, prec) as core:
            core.set_state(F)
            core.synchronize()
            t0 = time.perf_counter()
            loss = core.window_forward(args.refined_window, checkpoint_every=10, rel_tol=1e-10)
            core.synchronize()
            t1 = time.perf_counter()
            gt, gn = core.window_backward(1.0)
            core.synchronize()
            t2 = time.perf_counter()
            progress(f"{prec} window: forward {t1 - t0:.2f} s, backward {t2 - t1:.2f} s")
        grads[prec] = (gt[1:-1, 1:-1, :-1], gn)
        out[prec]["adjoint"] = {"window_steps": args.refined_window, "checkpoint_every": 10, "loss": loss,
                                "forward_s": round(t1 - t0, 3), "backward_s": round(t2 - t1, 3),
                                "backward_over_forward": round((t2 - t1) / (t1 - t0), 3), "grad_nue": gn}
    a, b = grads["mixed"][0], grads["fp64"][0]
    out["mixed"]["adjoint"]["grad_t0_rel_error_vs_fp64"] = {
        "max": float(np.max(np.abs(a - b)) / np.max(np.abs(b))),
        "l2": float(np.linalg.norm(a - b) / np.linalg.norm(b))}
    out["mixed"]["adjoint"]["grad_nue_rel_error_vs_fp64"] = abs(grads["mixed"][1] - grads["fp64"][1]) / abs(grads["fp64"][1])
    line = {"metric": "RK4 steps/s at 81x361x721 (mixed FP32/FP64 GPU step, Poisson to 1e-8)",
            "value": out["mixed"]["steps_per_s"], "unit": "steps/s", "n_gpus": 1, "steps": args.refined_steps,
            "warmup": 2, "ms_per_step": out["mixed"]["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32/f64", "data": "synthetic",
            "config": {"workload": f"81x361x721 shell (dr 0.0125, 0.5 deg), seeded synthetic state: "
                                   f"{args.refined_steps} steps, Poisson to {args.tol:g}; window adjoint over "
                                   f"{args.refined_window} steps, checkpoint every 10"},
            "fp64": out["fp64"], "mixed": out["mixed"]}
    print(json.dumps(line), flush=True)


def cpu_model() -> str:
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads() -> int:
    """The OpenMP threads the oracle runs on (its OMP_NUM_THREADS, else the cores this process may use)."""
    try:
        return int(os.environ["OMP_NUM_THREADS"])
    except (KeyError, ValueError):
        return len(os.sched_getaffinity(0))


def cpu_baseline(g, p, F0, tol, budget_s, warm=2):
    """The oracle on the same workload (warm-started solves, as the GPU runs it): one warm-up step on a
    copy of the initial state sizes the sample, then one call of n steps from the initial state (about
    budget_s); and, in a child process pinned to one OpenMP thread, one step. Returns the record, n and
    the oracle's state after the n steps."""
    from oracle import binding as O
    Fw = F0.copy()
    t0 = time.perf_counter()
    O.step_converge(g, p, Fw, 1, 200, tol, warm)         # warm-up, on a copy
    per = time.perf_counter() - t0
    del Fw
    n = max(2, min(50, int(budget_s / max(per, 1e-3))))
    F = F0.copy()
    t0 = time.perf_counter()
    O.step_converge(g, p, F, n, 200, tol, warm)
    el = time.perf_counter() - t0
    cores = cpu_threads()
    one = None
    try:
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--cpu-one-thread", "--tol", str(tol)],
                           env=env, capture_output=True, text=True, timeout=240)
        one = json.loads(r.stdout.strip().splitlines()[-1])["steps_per_s"] if r.returncode == 0 else None
    except (subprocess.TimeoutExpired, ValueError, IndexError, KeyError):
        one = None
    rec = {"value": round(n / el, 4), "unit": "steps/s", "cores": cores, "kind": "port",
           "cpu": cpu_model(), "one_thread_steps_per_s": one,
           "sample": f"{n} steps (hydro + RK4 + BC + warm-started Poisson to {tol:g}) of the 41x181x361 synthetic "
                     f"state in one call from the initial state, FP64 OpenMP oracle (C, -O2), after 1 warm-up step "
                     f"on a copy; one_thread: one step after one warm-up step with OMP_NUM_THREADS=1"}
    return rec, n, F


def fp64_vs_oracle(g, p, F0, Fo, n, opts):
    """The FP64 GPU step run n times from the initial state in one call (the oracle's warm-start history
    is that of one call) against the oracle's state after its n steps: per-field max |x - x_o| / max |x_o|."""
    import numpy as np
    import aturan_amd as A
    with A.UranusCore(g, p, "fp64") as core:
        core.set_state(F0)
        core.step(n, **opts)
        out = {}
        for name in ("t", "u", "v", "w", "ch4", "p_dyn", "p_hydro", "tke", "dis", "nh4sh"):
            a, b = core.get_field(name), Fo[A.FIELD[name]]
            s = float(np.max(np.abs(b)))
            out[name] = float(np.max(np.abs(a - b))) / (s if s > 0 else 1.0)
    return {"steps": n, "max_rel": out, "worst": max(out.values()), "bar": 1e-12}


def config1(g, p, F0, gpu_steps, cpu_steps=5):
    """The reference's CPU configuration: one RK4 step (hydro prep, four stages, BC and species clamp),
    then 20 red-black SOR sweeps at omega = 1.5. GPU: gpu_steps such steps in one call after 3 warm-up
    steps; CPU: the oracle on the host's cores, one warm-up step on a copy, then cpu_steps steps from
    the initial state, each timed alone (the median is reported), and the same on one thread in a child
    process (one warm-up step, the median of 3); the GPU's state after cpu_steps steps from the initial
    state against the oracle's; and the compiled reference's own step."""
    import numpy as np
    import aturan_amd as A
    from oracle import binding as O
    opts = dict(mode="reference", sweeps=20, omega=1.5)
    n_int = (g.im - 2) * (g.jm - 2) * (g.km - 1)
    with A.UranusCore(g, p, "fp64") as core:
        core.set_state(F0)
        core.step(cpu_steps, **opts)
        G = {n: core.get_field(n) for n in ("t", "u", "v", "w", "ch4", "p_dyn")}
        core.step(3, **opts)
        core.synchronize()
        t0 = time.perf_counter()
        core.step(gpu_steps, **opts)
        core.synchronize()
        el = time.perf_counter() - t0
    Fw = F0.copy()
    O.step(g, p, Fw, 1, 20, 1.5)
    del Fw
    F = F0.copy()
    times = []
    for _ in range(cpu_steps):
        t0 = time.perf_counter()
        O.step(g, p, F, 1, 20, 1.5)
        times.append(time.perf_counter() - t0)
    cel = sorted(times)[len(times) // 2]
    one = None
    try:
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--config1-one-thread"], env=env,
                           capture_output=True, text=True, timeout=240)
        one = json.loads(r.stdout.strip().splitlines()[-1])["steps_per_s"] if r.returncode == 0 else None
    except (subprocess.TimeoutExpired, ValueError, IndexError, KeyError):
        one = None
    err = {}
    for n, a in G.items():
        b = F[A.FIELD[n]]
        s = float(np.max(np.abs(b)))
        err[n] = float(np.max(np.abs(a - b))) / (s if s > 0 else 1.0)
    ref = reference_config1(g, F0)
    return {"workload": "41x181x361: hydro prep + RK4 + BC + species clamp + 20 red-black SOR sweeps, omega 1.5",
            "cpu_reference": ref,
            "gpu": {"steps": gpu_steps, "steps_per_s": round(gpu_steps / el, 2),
                    "ms_per_step": round(1e3 * el / gpu_steps, 4),
                    "cell_updates_per_s": round(n_int * gpu_steps / el, 1)},
            "cpu": {"steps": cpu_steps, "steps_per_s": round(1.0 / cel, 4), "seconds_per_step": [round(x, 4) for x in times],
                    "statistic": "median of single-step calls after one warm-up step", "cores": cpu_threads(),
                    "one_thread_steps_per_s": one, "kind": "port", "cpu": cpu_model()},
            "gpu_vs_oracle_after_cpu_steps": {"max_rel": err, "worst": max(err.values())}}


REF_HARNESS = ROOT / "oracle" / "_ref" / "ref_harness"


def reference_config1(g, F0, steps=2):
    """The compiled reference (oracle/_ref/ref_harness, built from the reference's own sources by
    oracle/build_ref.sh at -O2) running config 1's step as its loop does (restoreVar, RungeKuttaUran,
    BC, clampNegativeSpecies, PressureSolver with ATURAN_PRESS_SWEEPS=20; the reference's sweep has no
    relaxation factor, so omega is 1) on the same state, on the host's cores; the harness times each
    step. None when the harness is not built or its grid differs."""
    import tempfile
    import numpy as np
    if not REF_HARNESS.exists():
        return None
    threads = cpu_threads()
    with tempfile.TemporaryDirectory() as td:
        fin = Path(td) / "in.bin"
        with open(fin, "wb") as f:
            np.array([g.im, g.jm, g.km, F0.shape[0]], dtype=np.int32).tofile(f)
            np.ascontiguousarray(F0, dtype=np.float64).tofile(f)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        r = subprocess.run([str(REF_HARNESS), str(fin), "/dev/null", "step", str(steps), "20"], capture_output=True,
                           text=True, env=env, timeout=600)
    if r.returncode != 0:
        return {"error": r.stderr[-300:]}
    secs = [float(ln.split()[2]) for ln in r.stderr.splitlines() if ln.startswith("ref_step_seconds")]
    if not secs:
        return None
    best = min(secs[1:] or secs)      # after the first step, which pays the first touch of the arrays
    return {"steps": len(secs), "seconds_per_step": [round(x, 4) for x in secs], "steps_per_s": round(1.0 / best, 4),
            "cores": threads, "kind": "reference", "omega": 1.0, "build": "g++ -O2 -fno-fast-math -fopenmp",
            "cpu": cpu_model()}


def config1_one_thread():
    """Config 1's step in the oracle on this process's OpenMP threads (run with OMP_NUM_THREADS=1):
    one warm-up step, then the median of 3 single-step calls."""
    import aturan_amd as A
    from oracle import binding as O
    g = A.default_grid(1)
    p = A.default_params(g)
    F = A.synthetic_state(g, p, 20261203)
    O.bc(g, p, F)
    O.step(g, p, F, 1, 20, 1.5)
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        O.step(g, p, F, 1, 20, 1.5)
        ts.append(time.perf_counter() - t0)
    print(json.dumps({"steps_per_s": round(1.0 / sorted(ts)[1], 4)}), flush=True)


def cpu_one_thread(tol):
    import aturan_amd as A
    from oracle import binding as O
    g = A.default_grid(1)
    p = A.default_params(g)
    F = A.synthetic_state(g, p, 20261203)
    O.bc(g, p, F)
    O.step_converge(g, p, F, 1, 200, tol, 2)
    t0 = time.perf_counter()
    O.step_converge(g, p, F, 1, 200, tol, 2)
    print(json.dumps({"steps_per_s": round(1.0 / (time.perf_counter() - t0), 4)}), flush=True)


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.config1_one_thread:
        config1_one_thread()
        return
    if args.cpu_one_thread:
        cpu_one_thread(args.tol)
        return
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus N>1 is not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import aturan_amd as A

    if args.refined:
        refined_bench(args)
        return
    g, p, F = initial_state()
    warm = 0 if args.cold_start else args.warm
    opts64 = dict(mode="converge", rel_tol=args.tol, max_sweeps=60, warm_start=warm)
    optsmx = dict(mode="converge", rel_tol=args.tol, max_sweeps=60, sweeps=args.inner, warm_start=warm)
    line = {"metric": "RK4 steps/s at 41x181x361 (plain FP64 GPU step, Poisson to 1e-8)", "value": None,
            "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": None,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic"}
    cfg = {"workload": f"41x181x361 shell, seeded synthetic state: hydro prep + RK4 + BC + species clamp + red-black "
                       f"line/barotropic Poisson solve to relative residual {args.tol:g}, {args.steps} steps",
           "solve_start": SOLVE_START[warm]}
    line["runtime"] = runtime_record()
    copy = hbm_copy()
    progress(f"copy kernel {copy['GBps']} GB/s")
    if args.mode in ("both", "fp64"):
        with A.UranusCore(g, p, "fp64") as core:
            core.set_state(F)
            check_upload(core, F)
            el, st, (rk_ms, rk_n), gpu_ms = timed_run(core, args.steps, args.warmup, opts64, not args.no_events)
            breakdown = None if args.no_events else kernel_breakdown(core, min(args.steps, 20), opts64)
        sps = args.steps / el
        progress(f"fp64 {sps:.1f} steps/s")
        # the sampled steps' four stage launches: the mean algorithmic bytes of a stage over its mean duration
        achieved = sum(RK_PASSES_FP64) / 4 * 8 * N_INT / (rk_ms / rk_n * 1e-3) / 1e9 if rk_ms > 0 else 0.0
        line.update(value=round(sps, 3), ms_per_step=round(1e3 * el / args.steps, 4))
        cfg.update(cell_updates_per_s=round(sps * N_INT, 1), solve_iterations_last_step=st.sweeps_done,
                   solve_rel_resid_last_step=st.rel_resid)
        tr = traffic_from_profiles()
        line["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                            "frac": round(achieved / PEAK_HBM_GBS, 4),
                            "traffic": tr.get("bytes_per_launch") if tr else None,
                            "traffic_source": (f"{TRAFFIC_FILE}: rocprofv3 FETCH_SIZE and WRITE_SIZE of "
                                               f"{tr.get('kernel')}, scaled by the copy calibration; committed "
                                               f"summary, not measured in this run") if tr else None,
                            "kernel": "k_rk_stage_ring<STAGE> (fused RHS + RK fold, FP64, LDS ring)",
                            "avg_stage_ms": round(rk_ms / rk_n, 5) if rk_n else None, "launches": rk_n,
                            "events": f"stage kernels of every {EVENT_EVERY}th timed step",
                            "bytes_per_stage": sum(RK_PASSES_FP64) * 8 * N_INT / 4,
                            "copy_GBps": copy["GBps"], "frac_of_copy": round(achieved / copy["GBps"], 4)}
        line["kernels_ms_per_step"] = breakdown
        line["gpu_ms_per_step"] = round(gpu_ms, 4)   # stream time of each step (events), gaps included
    if args.mode in ("both", "mixed"):
        with A.UranusCore(g, p, "mixed") as core:
            core.set_state(F)
            check_upload(core, F)
            el, st, (rk_ms, rk_n), gpu_ms_mx = timed_run(core, args.steps, args.warmup, optsmx, not args.no_events)
            breakdown = None if args.no_events else kernel_breakdown(core, min(args.steps, 20), optsmx)
        sps = args.steps / el
        progress(f"mixed {sps:.1f} steps/s")
        achieved = sum(RK_BYTES_MIXED) / 4 * N_INT / (rk_ms / rk_n * 1e-3) / 1e9 if rk_ms > 0 else 0.0
        mixed = {"steps_per_s": round(sps, 3), "ms_per_step": round(1e3 * el / args.steps, 4),
                 "cell_updates_per_s": round(sps * N_INT, 1),
                 "speedup_vs_fp64": round(sps / line["value"], 4) if line["value"] else None,
                 "rk_stage_GBps": round(achieved, 1), "rk_stage_frac": round(achieved / PEAK_HBM_GBS, 4),
                 "solve_iterations_last_step": st.sweeps_done, "solve_refinements_last_step": st.outer_iters,
                 "solve_rel_resid_last_step": st.rel_resid,
                 "gpu_ms_per_step": round(gpu_ms_mx, 4), "kernels_ms_per_step": breakdown}
        if args.mode == "mixed":
            line.update(value=round(sps, 3), ms_per_step=mixed["ms_per_step"], dtype="f32/f64",
                        metric="RK4 steps/s at 41x181x361 (mixed FP32/FP64 GPU step, Poisson to 1e-8)")
        if not args.no_errors and args.mode == "both":
            mixed["rel_error_vs_fp64"] = mixed_errors(g, p, F, opts64, optsmx)
            progress("mixed vs fp64 errors done")
        line["mixed"] = mixed
    if args.adjoint_steps > 0 and args.mode in ("both", "fp64"):
        line["adjoint"] = adjoint_bench(g, p, F, args.adjoint_steps, args.checkpoint_every, args.tol,
                                        line["ms_per_step"] / 1e3 if line["ms_per_step"] else None, not args.no_events)
        progress("adjoint done")
    if args.turb_calls > 0 and args.mode in ("both", "fp64"):
        line["turbulence"] = turb_bench(g, p, F, args.turb_calls, not args.no_events)
    if args.satadj_calls > 0 and args.mode in ("both", "fp64"):
        line["saturation_adjustment"] = satadj_bench(g, p, F, args.satadj_calls, not args.no_events)
        progress("closure and saturation adjustment done")
    line["config"] = cfg
    if not args.no_cpu:
        rec, n, Fo = cpu_baseline(g, p, F, args.tol, args.cpu_seconds, warm)
        line["cpu_baseline"] = rec
        progress(f"cpu baseline {rec['value']} steps/s over {n} steps")
        if args.mode in ("both", "fp64"):
            line["fp64_vs_oracle"] = fp64_vs_oracle(g, p, F, Fo, n, opts64)
            progress(f"fp64 vs oracle after {n} steps: worst {line['fp64_vs_oracle']['worst']:.3e}")
        del Fo
        if args.config1_steps > 0:
            line["config1"] = config1(g, p, F, args.config1_steps)
            progress("config 1 done")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
