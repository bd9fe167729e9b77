"""Edge cases of the projection on the card against the oracle: a single
cell, a one-cell-thick slab (ny = 1, the reference's 2D case), odd sizes
that take the generic (non-quad) path, a domain with no wet cell and one
whose every cell is wet under a closed lid. Pressure within 1e-4 rel. L2 of
the f64 oracle, exact zeros where the oracle has them."""
import numpy as np
import pytest

from fluidisim_amd import _lib
from fluidisim_amd.volume import Domain3, Volume3
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
f32 = np.float32


def noise(n, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n).astype(f32)


def case(nx, ny, nz, dx, eta, seed, closed_lid=False, **opts):
    d = Domain3(nx, ny, nz, dx)
    vol = Volume3.configure(d, rho=1025.0, g_eff=9.81, **opts)
    vol.set_surface_mode(False, closed_lid)
    vol.set_free_surface(np.asarray(eta, f32), float(nz * dx) / 2)
    nu, nv, nw = d.faces
    faces = [noise(n, seed + q) for q, n in enumerate((nu, nv, nw))]
    vol.set_velocity(*faces)
    prob = po.Problem(nx, ny, nz, dx, np.asarray(eta, f32), rho=1025.0, g_eff=9.81, rest=float(nz * dx) / 2,
                      homogeneous_ghost=False, closed_lid=closed_lid)
    return vol, prob, faces


def project_both(vol, prob, faces, opts):
    scale, k1 = float(f32(-1025.0 / 1e-3)), float(f32(1e-3 / 1025.0))
    rep = vol.project_mobile3(scale, k1, 4000, warm=False)
    oref, *_, p64 = po.project(prob, *faces, scale, k1, 4000, opts, precision=64)
    return rep, oref, vol.pressure().cpu().numpy().astype(np.float64), p64


def close(p, p64):
    n = np.linalg.norm(p64)
    if n == 0:
        return np.all(p == 0)
    return np.linalg.norm(p - p64) / n <= 1e-4


@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 1, 6), (5, 3, 7), (32, 1, 24), (13, 9, 11)])
def test_small_and_odd_shapes_match_the_oracle(shape):
    nx, ny, nz = shape
    dx = 0.25
    c = np.arange(nx * ny)
    eta = (f32(nz * dx * 0.6) + f32(0.3 * dx) * np.sin(c.astype(f32))).astype(f32)
    vol, prob, faces = case(nx, ny, nz, dx, eta, seed=nx + 7 * ny + 31 * nz)
    rep, oref, p, p64 = project_both(vol, prob, faces, po.options())
    print(shape, rep.iterations, oref["iterations"], rep.residual)
    assert not rep.degraded and rep.divergence_plain <= 1e-5
    assert close(p, p64)
    vol.close()


def test_no_wet_cell_projects_to_nothing():
    """Every column below half a cell: no wet row, b = 0, no iteration, p = 0
    and the faces come back as given (delta3d_mobile.rs:481-495)."""
    nx, ny, nz, dx = 8, 4, 6, 0.5
    vol, prob, faces = case(nx, ny, nz, dx, np.full(nx * ny, 0.2, f32), seed=3)
    rep, oref, p, p64 = project_both(vol, prob, faces, po.options())
    assert rep.iterations == 0 and oref["iterations"] == 0 and not rep.degraded
    assert np.all(p == 0) and np.all(p64 == 0)
    vol.close()


def test_full_closed_box_is_a_pure_neumann_problem_with_the_gauge():
    """Every cell wet under a closed lid: the operator is singular; the gauge
    removes the mean of b and p (the oracle's extension), and the pressure is
    the f64 oracle's within 1e-4."""
    nx, ny, nz, dx = 12, 8, 10, 0.5
    vol, prob, faces = case(nx, ny, nz, dx, np.full(nx * ny, nz * dx + 1.0, f32), seed=9, closed_lid=True)
    rep, oref, p, p64 = project_both(vol, prob, faces, po.options())
    print(rep, oref)
    assert not rep.degraded and rep.divergence_plain <= 1e-5
    assert close(p, p64)
    vol.close()
