"""Writes the golden fixtures of tests/golden/ from the compiled reference (oracle/_ref/ref_harness).

Run in the build container only (it needs /root/reference to build the harness):
    python tests/golden/make_golden.py

Input of every case: the seeded synthetic state of the library, enriched so that every term of the
path is exercised (all 13 species, every precipitation source, the reaction mass fluxes, the NH4SH
flux limiter, tke/dis sources, a non-zero p_dyn, Q_precip large enough to hit the +-0.5 cap and an
eddy viscosity that is negative in places). The enrichment uses integer hashing only, so the test
that re-creates the input on another machine gets the same bits.

A fixture keeps, per case, every field at a fixed sample of cells (the radial, polar and seam
planes, the neighbours of the sin(theta) floor edge and the interior) plus exact (correctly rounded, as math.fsum) sums
and sums of squares of every field, and the buoyancy level means.
"""
from __future__ import annotations

import math
import subprocess
from concurrent.futures import ThreadPoolExecutor
from fractions import Fraction
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

import aturan_amd as A  # noqa: E402

SEED = 20261203
SAMPLE_I = [0, 1, 2, 20, 30, 38, 39, 40]
SAMPLE_J = [0, 1, 2, 23, 24, 90, 156, 157, 178, 179, 180]
SAMPLE_K = [0, 1, 2, 180, 358, 359, 360]
CASES = {
    # name: (harness mode, steps, sweeps, apply BC (and the species clamp) to the input first)
    # the BC alone (params.species_clamp = 0): its radial and polar extrapolations leave negative
    # species values in the corner rows, which the loop's clamp then zeroes
    "bc": ("bc", 1, 1, False),
    # the BC then clampNegativeSpecies on an input with negative and NaN species values
    "bc_clamp": ("bc_clamp", 1, 1, False),
    "rk4": ("rk4", 1, 1, True),
    "pressure_1": ("pressure", 1, 1, True),
    "pressure_20": ("pressure", 1, 20, True),
    # the solver's PRESS_LINE / PRESS_BAROTROPIC options (PRESS_ENV), on one thread
    "pressure_line_1": ("pressure", 1, 1, True),
    "pressure_line_20": ("pressure", 1, 20, True),
    "pressure_baro_1": ("pressure", 1, 1, True),
    "pressure_baro_20": ("pressure", 1, 20, True),
    "pressure_lb_20": ("pressure", 1, 20, True),
    "step_2x3": ("step", 2, 3, True),
    "rk4_uran": ("rk4_uran", 1, 1, True),
    "step_uran_2": ("step_uran", 2, 1, True),
    "damp_t": ("damp_t", 1, 1, True),
    "damp_uvw_2": ("damp_uvw", 1, 2, True),
    # the turbulence closure (planet/Turbulence.h) under the knobs of TURB_ENV
    "turb_run": ("turb_run", 1, 1, False),
    "turb_init": ("turb_init", 1, 1, False),
    "turb_run_komega": ("turb_run", 1, 1, False),
    "turb_run_keps": ("turb_run", 1, 1, False),
    "turb_run_wall": ("turb_run", 1, 1, False),
    "turb_init_wall_keps": ("turb_init", 1, 1, False),
    # the loop's saturation adjustment (four gases) and NH4SH equilibrium on the satadj input
    "satadj": ("satadj", 1, 1, False),
    "satadj_mix": ("satadj", 1, 1, False),
}
# The reference reads the closure's knobs from its environment (planet/Turbulence.h:125-176,
# planet/ATPhys.h env_for): ATURAN_TURB_MODEL, ATURAN_TURB_FREE, ATURAN_ABL_TAPER, ATURAN_TURB_CURV.
TURB_ENV = {
    "turb_run": {},
    "turb_init": {},
    "turb_run_komega": {"ATURAN_TURB_MODEL": "k_omega"},
    "turb_run_keps": {"ATURAN_TURB_MODEL": "k_epsilon"},
    "turb_run_wall": {"ATURAN_TURB_FREE": "0", "ATURAN_ABL_TAPER": "1", "ATURAN_TURB_CURV": "0"},
    "turb_init_wall_keps": {"ATURAN_TURB_FREE": "0", "ATURAN_TURB_MODEL": "k_epsilon"},
}
# The red-black solver's two options (planet/PressureSolver.h:298-300), read from the environment:
# the radial line pass and the barotropic column-mean correction after the pointwise passes of every
# sweep. PRESS_LINE holds the interior mean of p with an OpenMP reduction whose order follows the
# thread count, so these cases run on one thread: the serial (i, j, k) order.
PRESS_ENV = {
    "pressure_line_1": {"ATURAN_PRESS_LINE": "1"},
    "pressure_line_20": {"ATURAN_PRESS_LINE": "1"},
    "pressure_baro_1": {"ATURAN_PRESS_BAROTROPIC": "1"},
    "pressure_baro_20": {"ATURAN_PRESS_BAROTROPIC": "1"},
    "pressure_lb_20": {"ATURAN_PRESS_LINE": "1", "ATURAN_PRESS_BAROTROPIC": "1"},
}
# The same options as the library's solve-mode bits (ATU_SOLVE_PRESS_LINE, ATU_SOLVE_PRESS_BAROTROPIC).
PRESS_FLAGS = {n: (0x100 if "ATURAN_PRESS_LINE" in e else 0) | (0x200 if "ATURAN_PRESS_BAROTROPIC" in e else 0)
               for n, e in PRESS_ENV.items()}
# The same knobs as keyword arguments of the library's and the oracle's closure calls.
TURB_KNOBS = {
    "turb_run": {},
    "turb_init": {},
    "turb_run_komega": {"model": "k_omega"},
    "turb_run_keps": {"model": "k_epsilon"},
    "turb_run_wall": {"free_flow": False, "abl_taper": True, "curvature": False},
    "turb_init_wall_keps": {"free_flow": False, "model": "k_epsilon"},
}


def hash_uniform(field: int, shape, salt: int) -> np.ndarray:
    """Uniforms in [0, 1) from a SplitMix64-style mix of (salt, field, flat index), integer only."""
    n = int(np.prod(shape))
    x = np.arange(n, dtype=np.uint64) + np.uint64((salt * 1000003 + field * 7919) & 0xFFFFFFFF) * np.uint64(1 << 32)
    with np.errstate(over="ignore"):
        x = x * np.uint64(0x9E3779B97F4A7C15)
        x ^= x >> np.uint64(30)
        x = x * np.uint64(0xBF58476D1CE4E5B9)
        x ^= x >> np.uint64(27)
        x = x * np.uint64(0x94D049BB133111EB)
        x ^= x >> np.uint64(31)
    return ((x >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)).reshape(shape)


def enriched_state(grid, params, seed: int = SEED) -> np.ndarray:
    F = A.synthetic_state(grid, params, seed)
    shp = F.shape[1:]
    Fi = A.FIELD
    for n in A.FIELD_NAMES[Fi["ch4_cloud"]:Fi["nh4sh"] + 1]:
        F[Fi[n]] = 1e-3 * (0.5 + hash_uniform(Fi[n], shp, seed))
    F[Fi["ch4"]] *= 1.0 + 0.05 * (hash_uniform(Fi["ch4"], shp, seed) - 0.5)
    F[Fi["tke"]] = 1e-4 * (0.5 + hash_uniform(Fi["tke"], shp, seed))
    F[Fi["dis"]] = 1.78 * (0.5 + hash_uniform(Fi["dis"], shp, seed))
    F[Fi["nue"]] = 1.65e-4 * (hash_uniform(Fi["nue"], shp, seed) - 0.1)       # ~10 % negative
    F[Fi["Q_precip"]] = 1600.0 * (hash_uniform(Fi["Q_precip"], shp, seed) - 0.5)   # |precip_t| capped above ~537
    for n in A.FIELD_NAMES[Fi["S_precip_ch4"]:Fi["dis_source"] + 1]:
        F[Fi[n]] = 1e-3 * (hash_uniform(Fi[n], shp, seed) - 0.5)
    F[Fi["p_dyn"]] = 1e-3 * (hash_uniform(Fi["p_dyn"], shp, seed) - 0.5)
    for n in ("aux_u", "aux_v", "aux_w"):
        F[Fi[n]] = 0.1 * (hash_uniform(Fi[n], shp, seed) - 0.5)
    return F


def exact_sum(a: np.ndarray) -> float:
    """The correctly rounded sum of float64 values (math.fsum's result), vectorised: each value is an
    integer mantissa times a power of two; the mantissas, split into 27- and 26-bit halves, are summed
    per exponent by bincount (every partial sum is an integer below 2**53, so exact in float64) and the
    per-exponent totals combined as Python integers."""
    x = np.ascontiguousarray(a, dtype=np.float64).ravel()
    if x.size == 0:
        return 0.0
    if not np.isfinite(x).all():
        return math.fsum(x.tolist())
    m, e = np.frexp(x)
    mi = (m * 9007199254740992.0).astype(np.int64)   # m * 2**53: the 53-bit integer mantissa
    hi = mi >> 26
    lo = mi - (hi << 26)
    e = e.astype(np.int64) - 53
    e0 = int(e.min())
    idx = e - e0
    hs = np.bincount(idx, weights=hi.astype(np.float64))
    ls = np.bincount(idx, weights=lo.astype(np.float64))
    total = 0
    for b in np.flatnonzero((hs != 0) | (ls != 0)).tolist():
        total += ((int(hs[b]) << 26) + int(ls[b])) << b
    return float(Fraction(total) * Fraction(2) ** e0) if e0 < 0 else float(total << e0)


def exact_sums(arrs, square: bool = False) -> np.ndarray:
    """exact_sum of each array (or of its squares), the arrays on 8 threads (numpy releases the GIL)."""
    with ThreadPoolExecutor(8) as ex:
        return np.array(list(ex.map(lambda a: exact_sum(np.square(a) if square else a), list(arrs))))


def summarize(F: np.ndarray, level: np.ndarray) -> dict:
    ii, jj, kk = np.ix_(SAMPLE_I, SAMPLE_J, SAMPLE_K)
    sums = exact_sums(F)
    sq = exact_sums(F, square=True)
    return {"samples": F[:, ii, jj, kk].copy(), "sums": sums, "sumsq": sq, "level": level.copy()}


def summarize_turb(T3: np.ndarray, vel_star: np.ndarray) -> dict:
    """The closure's own arrays (prod, tken, disn and the [jm][km] friction velocity) as summarize()."""
    ii, jj, kk = np.ix_(SAMPLE_I, SAMPLE_J, SAMPLE_K)
    arrs = list(T3) + [vel_star]
    return {"turb_samples": T3[:, ii, jj, kk].copy(),
            "turb_vs_samples": vel_star[np.ix_(SAMPLE_J, SAMPLE_K)].copy(),
            "turb_sums": exact_sums(arrs),
            "turb_sumsq": exact_sums(arrs, square=True)}


def turb_input(grid, seed: int = SEED) -> tuple[np.ndarray, np.ndarray]:
    """Seeded prod, tken, disn and vel_star inputs of the closure cases (integer hashing only)."""
    shp = (grid.im, grid.jm, grid.km)
    T3 = np.stack([1e-3 * hash_uniform(60 + n, shp, seed) for n in range(3)])
    vs = 0.05 + 0.1 * hash_uniform(63, (grid.jm, grid.km), seed)
    return T3, vs


def write_turb_bin(path: Path, T3: np.ndarray, vs: np.ndarray) -> None:
    with open(path, "ab") as f:
        f.write(np.ascontiguousarray(T3).tobytes())
        f.write(np.ascontiguousarray(vs).tobytes())


def write_bin(path: Path, F: np.ndarray) -> None:
    hdr = np.array([F.shape[1], F.shape[2], F.shape[3], F.shape[0]], dtype=np.int32)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.ascontiguousarray(F).tobytes())


def read_bin(path: Path, nf: int, shape) -> tuple[np.ndarray, np.ndarray]:
    raw = np.fromfile(path, dtype=np.uint8)
    body = raw[16:].view(np.float64)
    n3 = int(np.prod(shape))
    F = body[: nf * n3].reshape((nf,) + tuple(shape)).copy()
    level = body[nf * n3: nf * n3 + shape[0]].copy()
    return F, level


def read_turb_tail(path: Path, nf: int, shape) -> tuple[np.ndarray, np.ndarray]:
    body = np.fromfile(path, dtype=np.uint8)[16:].view(np.float64)
    n3 = int(np.prod(shape))
    o = nf * n3 + shape[0]
    T3 = body[o: o + 3 * n3].reshape((3,) + tuple(shape)).copy()
    vs = body[o + 3 * n3: o + 3 * n3 + shape[1] * shape[2]].reshape(shape[1], shape[2]).copy()
    return T3, vs


def satadj_mix(grid, params, F: np.ndarray, seed: int = SEED) -> None:
    """The saturation-adjustment input with every branch reached: temperatures spread +-40 % around
    the profile (cells above the melting and below the ice-onset temperature of each gas), vapour
    from 1e-9 to 1e-3 kg/m3 (sub- and supersaturated cells), and a quarter of the cells of each gas
    with its cloud, a quarter with its ice, a quarter with both and a quarter with neither."""
    Fi = A.FIELD
    shp = F.shape[1:]
    F[Fi["t"]] *= 1.0 + 0.8 * (hash_uniform(70, shp, seed) - 0.5)
    for n, gas in enumerate(("h2o", "h2s", "nh3", "ch4")):
        F[Fi[gas]] = 10.0 ** (-9.0 + 6.0 * hash_uniform(71 + n, shp, seed))
        sel = (hash_uniform(75 + n, shp, seed) * 4.0).astype(np.int64)
        F[Fi[gas + "_cloud"]] *= (sel == 1) | (sel == 3)
        F[Fi[gas + "_ice"]] *= (sel == 2) | (sel == 3)
    F[Fi["nh4sh"]] *= hash_uniform(79, shp, seed) < 0.5


def species_negatives(F: np.ndarray, seed: int = SEED) -> None:
    """The clamp input: in each of the 13 species fields about 8 % of the cells negative (-1e-4 to 0)
    and 2 % NaN, spread over the interior and every boundary plane (integer hashing only)."""
    Fi = A.FIELD
    shp = F.shape[1:]
    for f in range(Fi["ch4"], Fi["nh4sh"] + 1):
        h = hash_uniform(90 + f, shp, seed)
        F[f] = np.where(h < 0.08, -1e-4 * hash_uniform(120 + f, shp, seed), F[f])
        F[f] = np.where((h >= 0.08) & (h < 0.10), np.nan, F[f])


def case_input(name: str, grid, params) -> np.ndarray:
    from oracle import binding as O
    F = enriched_state(grid, params)
    if name == "satadj_mix":
        satadj_mix(grid, params, F)
    if name == "bc_clamp":
        species_negatives(F)
    if CASES[name][3]:
        O.bc(grid, params, F)       # with the species clamp of the default parameters
    return F


def main() -> None:
    r = subprocess.run(["bash", str(ROOT / "oracle" / "build_ref.sh")], capture_output=True, text=True)
    if r.returncode != 0:
        raise SystemExit(f"reference build failed:\n{r.stdout}\n{r.stderr}")
    harness = ROOT / "oracle" / "_ref" / "ref_harness"
    g = A.default_grid(1)
    p = A.default_params(g)
    shape = (g.im, g.jm, g.km)
    with tempfile.TemporaryDirectory() as td:
        names = sys.argv[1:] or list(CASES)
        for name in names:
            mode, steps, sweeps, _ = CASES[name]
            F = case_input(name, g, p)
            fin, fout = Path(td) / "in.bin", Path(td) / "out.bin"
            write_bin(fin, F)
            if name in TURB_ENV:
                write_turb_bin(fin, *turb_input(g))
            env = {"OMP_NUM_THREADS": "1" if name in PRESS_ENV else "8", "PATH": "/usr/bin",
                   **TURB_ENV.get(name, {}), **PRESS_ENV.get(name, {})}
            rr = subprocess.run([str(harness), str(fin), str(fout), mode, str(steps), str(sweeps)],
                                capture_output=True, text=True, env=env)
            if rr.returncode != 0:
                raise SystemExit(f"harness {name} failed:\n{rr.stderr[-2000:]}")
            consts = [ln for ln in rr.stderr.splitlines() if ln.startswith(("ref_harness:", "ref_satadj:"))]
            G, level = read_bin(fout, A.NFIELDS, shape)
            s = summarize(G, level)
            if name in TURB_ENV:
                s.update(summarize_turb(*read_turb_tail(fout, A.NFIELDS, shape)))
            np.savez_compressed(HERE / f"ref_{name}.npz", **s, mode=mode, steps=steps, sweeps=sweeps,
                                consts=np.array(consts))
            print(f"wrote ref_{name}.npz ({consts[0] if consts else ''})", flush=True)


if __name__ == "__main__":
    main()
