"""Pins the oracle's Poisson solves by what the mathematics fixes, independently of its code.

* The converging solve's answer solves the folded-wall system of the red-black sweep: checked on a
  tiny shell against a dense least-squares solve of that system, assembled here in NumPy.
* A converged p is a fixed point of one reference red-black sweep (planet/PressureSolver.h:233-235
  states this of the exact solution) when the source is consistent.
* SOR: one sweep at omega equals (1 - omega) p + omega p_GS cell by cell for the first colour.
* The converging solve reaches 1e-8 within a handful of iterations on the full shell.
"""
import numpy as np
import pytest

import aturan_amd as A

C43, C13 = 4.0 / 3.0, 1.0 / 3.0


def tiny_grid(im=6, jm=9, km=11):
    g = A.default_grid(1)
    g.im, g.jm, g.km = im, jm, km
    g.the0 = 60.0 * np.pi / 180.0          # rows away from the pole, both sine branches exercised
    return g


def weights(g, p):
    """num1, num2[i], num3[i,j], den[i,j] of the 7-point operator, from the grid alone."""
    radz = g.r0 + g.dr * np.arange(g.im)
    thez = g.the0 + g.dthe * np.arange(g.jm)
    rm = radz + (p.R_metric_km / p.L_atm - 1.0)
    s = np.maximum(np.sin(thez), p.sinthe_min)
    num1 = 1.0 / g.dr ** 2
    num2 = 1.0 / rm ** 2 / g.dthe ** 2
    num3 = (1.0 / rm ** 2)[:, None] / s[None, :] ** 2 / g.dphi ** 2
    den = 2 * num1 + 2 * num2[:, None] + 2 * num3
    return num1, num2, num3, den


def folded_matrix(g, p):
    """Dense A of the interior unknowns with p_wall = c43 p1 - c13 p2 substituted in r and theta."""
    ni, nj, N = g.im - 2, g.jm - 2, g.km - 1
    num1, num2, num3, den = weights(g, p)
    idx = lambda i, j, k: ((i - 1) * nj + (j - 1)) * N + (k % N)
    n = ni * nj * N
    M = np.zeros((n, n))

    def add(row, i, j, k, w):
        # neighbour (i, j, k) of the interior, folding walls
        if i == 0:
            add(row, 1, j, k, w * C43); add(row, 2, j, k, -w * C13); return
        if i == g.im - 1:
            add(row, g.im - 2, j, k, w * C43); add(row, g.im - 3, j, k, -w * C13); return
        if j == 0:
            add(row, i, 1, k, w * C43); add(row, i, 2, k, -w * C13); return
        if j == g.jm - 1:
            add(row, i, g.jm - 2, k, w * C43); add(row, i, g.jm - 3, k, -w * C13); return
        M[row, idx(i, j, k)] += w

    for i in range(1, g.im - 1):
        for j in range(1, g.jm - 1):
            for k in range(N):
                r = idx(i, j, k)
                M[r, r] -= den[i, j]
                add(r, i + 1, j, k, num1); add(r, i - 1, j, k, num1)
                add(r, i, j + 1, k, num2[i]); add(r, i, j - 1, k, num2[i])
                add(r, i, j, k + 1, num3[i, j]); add(r, i, j, k - 1 + N, num3[i, j])
    return M


def tiny_state(g, p, seed=3):
    rng = np.random.default_rng(seed)
    F = np.zeros((A.NFIELDS, g.im, g.jm, g.km))
    for n in ("aux_u", "aux_v", "aux_w", "p_dyn"):
        F[A.FIELD[n]] = 1e-3 * rng.standard_normal((g.im, g.jm, g.km))
    return F


def test_converged_solution_solves_the_folded_system(oracle):
    g = tiny_grid()
    p = A.default_params(g)
    F = tiny_state(g, p)
    it, rel, src = oracle.pressure_converge(g, p, F, 200, 1e-13)
    assert rel <= 1e-12, (it, rel)
    ni, nj, N = g.im - 2, g.jm - 2, g.km - 1
    b = src[1:-1, 1:-1, :N].ravel()
    y = np.outer(np.r_[1.5, np.ones(ni - 2), 1.5], np.r_[1.5, np.ones(nj - 2), 1.5])
    yv = np.repeat(y[:, :, None], N, axis=2).ravel()
    b = b - (yv @ b) / yv.sum()                         # the consistent source
    M = folded_matrix(g, p)
    assert np.abs(yv @ M).max() <= 1e-9 * np.abs(M).max()   # y is the left null vector
    pd = F[A.FIELD["p_dyn"]]
    x = pd[1:-1, 1:-1, :N].ravel()
    assert np.linalg.norm(M @ x - b) <= 1e-10 * np.linalg.norm(b)
    xs = np.linalg.lstsq(M, b, rcond=None)[0]
    xs += x.mean() - xs.mean()
    assert np.abs(x - xs).max() <= 1e-8 * np.abs(xs).max()
    # walls are the folds of the interior, and the seam the copy of k = 0
    assert np.allclose(pd[0, 1:-1, :N], C43 * pd[1, 1:-1, :N] - C13 * pd[2, 1:-1, :N], rtol=0, atol=1e-15)
    assert np.allclose(pd[1:-1, 0, :N], C43 * pd[1:-1, 1, :N] - C13 * pd[1:-1, 2, :N], rtol=0, atol=1e-15)
    assert np.array_equal(pd[:, :, N], pd[:, :, 0])


def test_converged_solution_is_a_fixed_point_of_the_reference_sweep(oracle):
    g = tiny_grid()
    p = A.default_params(g)
    F = tiny_state(g, p)
    F[A.FIELD["aux_u"]] = 0.0          # div = d(aux_w)/dphi only: its ring sums vanish, the
    F[A.FIELD["aux_v"]] = 0.0          # source is consistent without a shift
    oracle.pressure_converge(g, p, F, 200, 1e-13)
    p_star = F[A.FIELD["p_dyn"]].copy()
    oracle.pressure(g, p, F, 1, 1.0)
    change = np.abs(F[A.FIELD["p_dyn"]] - p_star).max()
    assert change <= 1e-10 * np.abs(p_star).max(), change


def test_sor_is_a_relaxed_gauss_seidel_sweep(oracle):
    g = tiny_grid()
    p = A.default_params(g)
    F0 = tiny_state(g, p)
    F1, F2 = F0.copy(), F0.copy()
    oracle.pressure(g, p, F1, 1, 1.0)
    oracle.pressure(g, p, F2, 1, 1.5)
    # the first colour sees only old neighbours: there SOR = (1 - w) p_old + w p_GS exactly
    pd0, pg, ps = (F[A.FIELD["p_dyn"]] for F in (F0, F1, F2))
    i, j, k = np.meshgrid(np.arange(g.im), np.arange(g.jm), np.arange(g.km), indexing="ij")
    red = ((i + j + k) % 2 == 0) & (i > 0) & (i < g.im - 1) & (j > 0) & (j < g.jm - 1) & (k < g.km - 1)
    assert np.allclose(ps[red], (1 - 1.5) * pd0[red] + 1.5 * pg[red], rtol=1e-14, atol=1e-18)


@pytest.mark.slow
def test_converging_solve_reaches_1e8_on_the_full_shell(oracle):
    g = A.default_grid(1)
    p = A.default_params(g)
    F = A.synthetic_state(g, p, 20261203)
    oracle.bc(g, p, F)
    oracle.hydro(g, p, F)
    oracle.rk4(g, p, F)
    it, rel, _ = oracle.pressure_converge(g, p, F, 30, 1e-8)
    assert rel <= 1e-8 and it <= 8, (it, rel)
