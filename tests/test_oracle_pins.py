"""Pins of two oracle functions against what the mathematics fixes, by brute
force on small grids (dense matrices assembled column by column from the
oracle's own f64 fine operator, which the dense-solve, DCT and symmetry tests
of test_oracle.py pin on its own):

- the Oettli-Prager backward error of delta3d_mobile.rs:243-271 is
  max_c |r_c| / (|b_c| + (|A| |p|)_c) over wet rows, |A| the elementwise
  absolute value of the assembled operator;
- the Galerkin coarse operators (the oracle's coarse_op = 1 extension) are
  P^T A P / 16 level by level, P the piecewise-constant prolongation onto the
  wet (active) children, at a free surface with ghost faces and around solids.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

f32 = np.float32


def geometry(nx, ny, nz, seed, solids=True, homogeneous=False):
    """A small tank: surface heights varying over the columns (ghost faces at
    several levels and fractions of a cell), a few solid cells."""
    rng = np.random.default_rng(seed)
    dx = 0.5
    eta = (f32(nz * dx * 0.55) + f32(nz * dx * 0.3) * rng.uniform(-1, 1, nx * ny).astype(f32)).astype(f32)
    solid = None
    if solids:
        solid = np.zeros(nx * ny * nz, np.uint8)
        solid[rng.choice(nx * ny * nz, size=max(1, nx * ny * nz // 12), replace=False)] = 1
    return po.Problem(nx, ny, nz, dx, eta, rho=1025.0, g_eff=9.81, rest=float(nz * dx) / 2,
                      homogeneous_ghost=homogeneous, solid=solid)


def dense_fine(prob):
    n = prob.cells
    A = np.zeros((n, n))
    e = np.zeros(n)
    for c in range(n):
        e[c] = 1.0
        A[:, c] = po.apply64(prob, e)
        e[c] = 0.0
    return A


def dense_coarse(prob, level, opts):
    nx, ny, nz = prob.nx, prob.ny, prob.nz
    for _ in range(level + 1):
        nx, ny, nz = nx // 2, ny // 2, nz // 2
    n = nx * ny * nz
    A = np.zeros((n, n))
    e = np.zeros(n)
    for c in range(n):
        e[c] = 1.0
        A[:, c] = po.coarse_apply(prob, level, e, opts)
        e[c] = 0.0
    return A, (nx, ny, nz)


def prolongation(fine_shape, active):
    """P[f, c] = 1 when fine cell f is active and its parent is c."""
    fx, fy, fz = fine_shape
    cx, cy, cz = fx // 2, fy // 2, fz // 2
    P = np.zeros((fx * fy * fz, cx * cy * cz))
    for k in range(fz):
        for j in range(fy):
            for i in range(fx):
                f = (k * fy + j) * fx + i
                if active[f]:
                    P[f, ((k // 2) * cy + j // 2) * cx + i // 2] = 1.0
    return P


@pytest.mark.parametrize("shape,seed", [((6, 5, 7), 1), ((4, 4, 6), 2), ((5, 3, 7), 3), ((6, 5, 7), 4)])
def test_backward_error_is_the_dense_oettli_prager_formula(shape, seed):
    prob = geometry(*shape, seed=seed, solids=seed % 2 == 1)
    A = dense_fine(prob)
    kind, _ = po.classify(prob)
    wet = kind == 1
    assert wet.sum() > 4
    rng = np.random.default_rng(100 + seed)
    n = prob.cells
    p = rng.standard_normal(n)
    b = np.where(wet, rng.standard_normal(n), 0.0)
    r = np.where(wet, 1e-3 * rng.standard_normal(n), 0.0)
    scale = np.abs(b) + np.abs(A) @ np.abs(p)
    rows = wet & (scale > 0)
    expected = np.max(np.abs(r[rows]) / scale[rows])
    got = po.backward_error(prob, p, b, r, precision=64)
    assert got == pytest.approx(expected, rel=1e-12)
    # a row whose residual dominates moves the maximum to that row
    worst = int(np.flatnonzero(wet)[len(np.flatnonzero(wet)) // 2])
    r2 = r.copy()
    r2[worst] = 10.0 * scale[worst]
    assert po.backward_error(prob, p, b, r2, precision=64) == pytest.approx(10.0, rel=1e-12)


def test_backward_error_is_zero_for_an_exact_solution_and_ignores_dry_rows():
    prob = geometry(6, 5, 7, seed=5)
    kind, _ = po.classify(prob)
    n = prob.cells
    p = np.random.default_rng(6).standard_normal(n)
    b = np.where(kind == 1, 1.0, 0.0)
    r = np.where(kind == 1, 0.0, 1e6)  # residual only on dry rows: not counted
    assert po.backward_error(prob, p, b, r, precision=64) == 0.0


@pytest.mark.parametrize("shape,seed,solids", [((8, 8, 8), 1, True), ((16, 8, 8), 2, False), ((8, 8, 12), 3, True)])
def test_galerkin_coarse_operators_are_p_transpose_a_p(shape, seed, solids):
    prob = geometry(*shape, seed=seed, solids=solids)
    opts = po.options(multigrid=1, smoother=1, coarse_op=1)
    A = dense_fine(prob)
    kind, ghost = po.classify(prob)
    wet = kind == 1
    assert ghost.sum() > 0, "the case must have free-surface rows"
    # level 0 from the fine operator
    P0 = prolongation(shape, wet)
    A0, shape0 = dense_coarse(prob, 0, opts)
    G0 = P0.T @ A @ P0 / 16.0
    assert np.abs(G0).max() > 0
    np.testing.assert_allclose(A0, G0, rtol=1e-12, atol=1e-12 * np.abs(G0).max())
    # level 1 from level 0 (active coarse cells: those with a wet child)
    active0 = P0.sum(axis=0) > 0
    P1 = prolongation(shape0, active0)
    A1, _ = dense_coarse(prob, 1, opts)
    G1 = P1.T @ A0 @ P1 / 16.0
    np.testing.assert_allclose(A1, G1, rtol=1e-12, atol=1e-12 * np.abs(G1).max())


def test_galerkin_differs_from_the_rediscretised_operator_at_the_surface():
    """The pin has teeth: the re-discretised coarse operator (Dirichlet at a
    coarse half cell) is not P^T A P / 16 on the same free-surface case."""
    shape = (8, 8, 8)
    prob = geometry(*shape, seed=1)
    A = dense_fine(prob)
    kind, _ = po.classify(prob)
    P0 = prolongation(shape, kind == 1)
    G0 = P0.T @ A @ P0 / 16.0
    R0, _ = dense_coarse(prob, 0, po.options(multigrid=1, smoother=1, coarse_op=0))
    assert np.abs(R0 - G0).max() > 1e-3 * np.abs(G0).max()
