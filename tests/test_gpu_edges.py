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


def test_quad_and_per_cell_kernels_agree_bit_for_bit():
    """The correction and divergence-metric kernels on quads against the
    per-cell ones: the same cold-start projection with 16-byte-aligned output
    faces and with faces offset by one float (which the host sends down the
    per-cell kernels) gives the same faces, pressure and report bit for bit."""
    import ctypes as C
    import torch
    from fluidisim_amd import workloads as W
    from fluidisim_amd.volume import _ptr

    wl = W.tank(64, 32, 48, seed=11)
    vol = W.build_volume(wl)
    L = _lib.lib()

    def run(offset):
        outs = []
        for n in wl.domain.faces:
            buf = torch.zeros(n + 1, dtype=torch.float32, device="cuda")
            outs.append(buf[offset:offset + n])
        p = torch.zeros(wl.domain.cells, dtype=torch.float32, device="cuda")
        rep = _lib.Report()
        vol._upload_geometry()
        vol._sync()
        st = L.fsp_project_device(vol._h, _ptr(vol.us), _ptr(vol.vs), _ptr(vol.ws), _ptr(p), 0,
                                  *[_ptr(o) for o in outs], wl.scale, wl.k1, 1e-6, 2000, C.byref(rep))
        assert st == 0
        return [o.cpu().numpy() for o in outs], p.cpu().numpy(), rep

    faces_q, p_q, rep_q = run(0)
    faces_c, p_c, rep_c = run(1)
    assert rep_q.iterations == rep_c.iterations > 0
    # the divergence metrics (k_divmetric_q against k_divmetric) and the residual
    for f, _ in _lib.Report._fields_:
        if not f.startswith("ms_"):  # timings
            assert getattr(rep_q, f) == getattr(rep_c, f), f
    assert np.array_equal(p_q.view(np.uint32), p_c.view(np.uint32))
    for a, b in zip(faces_q, faces_c):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert np.abs(a).max() > 0
    vol.close()
