"""Benchmark of the dynamical-core step on one MI355X.

A step is one outer time step of the benchmark loop on the 41x181x361 shell: hydrostatic prep,
the four fused RK4 stages, the domain BC and the red-black Poisson solve, with every input already
resident in HBM. Prints ONE JSON line (rank 0) with the throughput in steps/s, the roofline of the
dominant kernel (the fused RK stage) measured with HIP events on the library's own stream, and the
CPU oracle timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBS = 8000.0          # MI355X HBM3E, vendor figure
N_INT = 39 * 179 * 360         # integrated interior cells
# Algorithmic bytes of the four fused stages: array passes per interior cell (stage 0: 80,
# stages 1-2: 118 each, stage 3: 102) at 8 B per FP64 value.
RK_PASSES = (80, 118, 118, 102)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--sweeps", type=int, default=20)
    ap.add_argument("--omega", type=float, default=1.5)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device-to-device copy of known size (calibrates the HBM byte counters)")
    return ap.parse_args()


def copy_calibration():
    """A plain copy kernel moving a known number of bytes, for scaling FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = 64 * 1024 * 1024                      # 512 MiB per FP64 buffer: far beyond any cache
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read_per_launch": n * 8, "bytes_written_per_launch": n * 8,
                                           "launches": 3}}), flush=True)


def traffic_from_profiles():
    f = ROOT / "profiles" / "traffic_rk_stage.json"
    if f.exists():
        try:
            return json.loads(f.read_text()).get("bytes_per_launch")
        except Exception:
            return None
    return None


def cpu_baseline(sweeps, omega, budget_s):
    import numpy as np  # noqa: F401
    import aturan_amd as A
    from oracle import binding as O
    g = A.default_grid(1)
    p = A.default_params(g)
    F = A.synthetic_state(g, p, 20261203)
    O.bc(g, p, F)
    O.step(g, p, F, 1, sweeps, omega)           # warm-up
    n, t0 = 0, time.perf_counter()
    while True:
        O.step(g, p, F, 1, sweeps, omega)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 20:
            break
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": n / el, "unit": "steps/s", "cores": cores, "kind": "port",
            "sample": f"{n} steps (hydro + RK4 + BC + {sweeps} red-black sweeps at omega {omega}) of the "
                      f"41x181x361 synthetic state, FP64 OpenMP oracle"}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus N>1 is not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np  # noqa: F401
    import torch
    import aturan_amd as A
    from oracle import binding as O

    g = A.default_grid(1)
    p = A.default_params(g)
    F = A.synthetic_state(g, p, 20261203)
    O.bc(g, p, F)
    core = A.UranusCore(g, p)
    core.set_state(F)
    for _ in range(args.warmup):
        core.step(1, "reference", sweeps=args.sweeps, omega=args.omega)
    core.synchronize()
    torch.cuda.synchronize()
    core.profile_begin()
    t0 = time.perf_counter()
    st = core.step(args.steps, "reference", sweeps=args.sweeps, omega=args.omega)
    core.synchronize()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    rk_ms, rk_n = core.profile_query("rk_stage")
    rb_ms, rb_n = core.profile_query("rb_pass")
    core.profile_end()
    steps_per_s = args.steps / el
    # algorithmic bytes of all stage launches / their summed duration = bytes per launch / avg duration
    bytes_all = sum(RK_PASSES) * 8 * N_INT * (rk_n // 4)
    achieved = bytes_all / (rk_ms * 1e-3) / 1e9 if rk_ms > 0 else 0.0
    line = {
        "metric": "RK4 steps/s at 41x181x361 (plain FP64 GPU step)",
        "value": round(steps_per_s, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * el / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"41x181x361 shell: hydro prep + RK4 + BC + {args.sweeps} red-black sweeps "
                               f"omega {args.omega}, FP64",
                   "cell_updates_per_s": round(steps_per_s * N_INT, 1),
                   "last_rel_resid": st.rel_resid},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic_from_profiles(),
                     "kernel": "k_rk_stage (fused RHS + RK fold)",
                     "avg_launch_ms": round(rk_ms / max(rk_n, 1), 5)},
        "kernels": {"rb_pass_avg_ms": round(rb_ms / max(rb_n, 1), 5), "rb_pass_launches": rb_n},
    }
    core.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.sweeps, args.omega, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
