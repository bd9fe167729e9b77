"""The coupled step's projection path in the oracle (CPU): the projection
over a coupled surface (delta3d_mobile.rs:92-147 with surface_coupled, the
terms of prepare_background3, delta3d_coupling.rs:97-189) and
refine_mobile3 (delta3d_coupling.rs:226-249).

Pins, none of which restates the ghost formulas:
- the up-ghost term: a coupled problem whose column term is
  rho g (total - eta) is the uncoupled problem over the total surface;
- the side-ghost terms: with every ghost value equal to a constant c and no
  velocity, p = c solves the system, so a term read from the wrong face (the
  test fills every other face with large noise) moves p off c;
- refine: for scale * k1 = -1 the corrected faces of an unconverged
  pressure p1 carry its residual, so p1 plus the refinement is the solution
  (A^-1 b) whatever p1 was;
- the step's rule: a converged projection is not refined.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

RHO, G = 1025.0, 9.81
DT = 0.01
SCALE = np.float32(-RHO / DT)
K1 = np.float32(DT / RHO)


def coupled_case(nx=12, ny=10, nz=8, dx=0.1, seed=3):
    """A tank with its own surface eta (the perturbation) and a total surface
    eta + background elevation that lifts and drops columns by up to 1.5
    cells, so that the total surface's wetness differs from eta's."""
    rng = np.random.default_rng(seed)
    rest = 4.0 * dx
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)  # (ny, nx), column j*nx + i
    eta = (rest + 0.3 * dx * np.sin(2 * np.pi * X) * np.cos(np.pi * Y)).astype(np.float32).ravel()
    zeta = (1.5 * dx * np.sin(2 * np.pi * (X + 0.7 * Y))).astype(np.float32).ravel()
    total = (eta + zeta).astype(np.float32)
    prob = po.Problem(nx, ny, nz, dx, eta, rho=RHO, g_eff=G, rest=rest, homogeneous_ghost=False)
    return prob, total, rng


def wet(prob, heights):
    k = np.arange(prob.nz)[:, None, None]
    h = heights.reshape(prob.ny, prob.nx)[None]
    return ((k.astype(np.float32) + np.float32(0.5)) * np.float32(prob.dx) < h)  # (nz, ny, nx)


def random_faces(prob, rng, scale=1.0):
    nu, nv, nw = prob.face_counts
    return [(scale * rng.standard_normal(n)).astype(np.float32) for n in (nu, nv, nw)]


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_up_term_equals_uncoupled_total_surface():
    prob, total, rng = coupled_case()
    assert (wet(prob, total) != wet(prob, prob.eta)).any()  # the coupling changes wetness
    us, vs, ws = random_faces(prob, rng)
    bg_up = (RHO * G * (total.astype(np.float64) - prob.eta.astype(np.float64))).astype(np.float32)
    rep_c, _, _, _, _, p_c = po.project_coupled(prob, us, vs, ws, SCALE, K1, total, bg_up=bg_up, precision=64)
    flat = po.Problem(prob.nx, prob.ny, prob.nz, prob.dx, total, rho=RHO, g_eff=G, rest=prob.rest,
                      homogeneous_ghost=False)
    rep_u, _, _, _, _, p_u = po.project(flat, us, vs, ws, SCALE, K1, precision=64)
    assert not rep_c["degraded"] and not rep_u["degraded"]
    assert rel(p_c, p_u) < 1e-6
    # the own surface instead of the total one is a different problem
    rep_o, _, _, _, _, p_o = po.project(prob, us, vs, ws, SCALE, K1, precision=64)
    assert rel(p_o, p_u) > 1e-3


def constant_ghost_terms(prob, total, c, rng):
    """Background terms that make every ghost value c: bg_up per column and
    bg_x / bg_y on the side-ghost faces between a wet and a dry cell of the
    total surface; every other face holds noise of size 1e4."""
    nx, ny, nz, dx = prob.nx, prob.ny, prob.nz, np.float32(prob.dx)
    rg = np.float64(RHO) * np.float64(G)
    bg_up = (c - rg * (prob.eta.astype(np.float64) - np.float64(np.float32(prob.rest)))).astype(np.float32)
    nu, nv, _ = prob.face_counts
    bg_x = (1e4 * rng.standard_normal(nu)).astype(np.float32)
    bg_y = (1e4 * rng.standard_normal(nv)).astype(np.float32)
    W = wet(prob, total)
    used = 0
    for k in range(nz):
        zc = (np.float32(k) + np.float32(0.5)) * dx
        side = (c - rg * (np.float64(zc) - np.float64(np.float32(prob.rest))))
        for j in range(ny):
            for i in range(1, nx):
                if W[k, j, i - 1] != W[k, j, i]:
                    bg_x[(k * ny + j) * (nx + 1) + i] = side
                    used += 1
        for j in range(1, ny):
            for i in range(nx):
                if W[k, j - 1, i] != W[k, j, i]:
                    bg_y[(k * (ny + 1) + j) * nx + i] = side
                    used += 1
    assert used > 10
    return bg_up, bg_x, bg_y


def test_side_terms_constant_ghost_value_gives_constant_pressure():
    prob, total, rng = coupled_case(seed=5)
    c = 1234.5
    bg_up, bg_x, bg_y = constant_ghost_terms(prob, total, c, rng)
    nu, nv, nw = prob.face_counts
    z = [np.zeros(n, np.float32) for n in (nu, nv, nw)]
    rep, _, u, v, w, p64 = po.project_coupled(prob, *z, SCALE, K1, total, bg_up, bg_x, bg_y, precision=64)
    assert not rep["degraded"]
    W = wet(prob, total).ravel()
    assert np.abs(p64[W] - c).max() < 1e-6 * c
    assert np.abs(p64[~W]).max() == 0.0
    assert max(np.abs(u).max(), np.abs(v).max(), np.abs(w).max()) < 1e-6
    # the same terms moved to the faces' other index (min instead of max)
    # break it: the noise is read
    bg_x2 = np.roll(bg_x, -1)
    rep2, _, _, _, _, p2 = po.project_coupled(prob, *z, SCALE, K1, total, bg_up, bg_x2, bg_y, precision=64)
    assert np.abs(p2[W] - c).max() > 1e-2 * c


@pytest.mark.parametrize("precision", [64, 32])
def test_refine_of_an_unconverged_pressure_is_the_solution(precision):
    prob, total, rng = coupled_case(seed=7)
    us, vs, ws = random_faces(prob, rng)
    bg_up = (50.0 * rng.standard_normal(prob.nx * prob.ny)).astype(np.float32)
    nu, nv, _ = prob.face_counts
    bg_x = (50.0 * rng.standard_normal(nu)).astype(np.float32)
    bg_y = (50.0 * rng.standard_normal(nv)).astype(np.float32)
    args = (prob, us, vs, ws, SCALE, K1, total, bg_up, bg_x, bg_y)
    full = po.project_coupled(*args, precision=64)
    assert not full[0]["degraded"]
    first = po.project_coupled(*args, max_iters=2, precision=64)
    assert first[0]["degraded"] and rel(first[5], full[5]) > 1e-3
    refined = po.project_coupled(*args, max_iters=2, refine=po.REFINE_ALWAYS, refine_iters=20000,
                                 precision=precision)
    rep = refined[0]
    assert rep["refinements"] == 1 and not rep["degraded"]
    p = refined[5] if precision == 64 else refined[1].astype(np.float64)
    # both solves stop at a relative residual of 1e-6
    tol = 1e-5 if precision == 64 else 1e-4
    assert rel(p, full[5]) < tol
    for x, y in zip(refined[2:5], full[2:5]):
        assert rel(x.astype(np.float64), y.astype(np.float64)) < 10 * tol


def test_step_rule_skips_a_converged_projection():
    prob, total, rng = coupled_case(seed=13)
    us, vs, ws = random_faces(prob, rng)
    rep = po.project_coupled(prob, us, vs, ws, SCALE, K1, total, refine=po.REFINE_STEP)[0]
    assert rep["refinements"] == 0 and not rep["degraded"]


def test_uncoupled_call_unchanged():
    """No coupling arrays: the coupled entry is the plain projection, bit for bit."""
    prob, _, rng = coupled_case(seed=17)
    us, vs, ws = random_faces(prob, rng)
    a = po.project(prob, us, vs, ws, SCALE, K1)
    b = po.project_coupled(prob, us, vs, ws, SCALE, K1, None)
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)
