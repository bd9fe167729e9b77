"""The coupled step's projection path on the card: the projection over a
coupled surface (fsp_set_coupling: the total surface's heights and the
background's ghost terms, delta3d_mobile.rs:92-147 with surface_coupled) and
refine_mobile3 (fsp_refine_device, delta3d_coupling.rs:226-249), against the
oracle in f64 (pressure within 1e-4 relative L2 on wet cells; faces within
1e-5 of the largest speed of the oracle's correction of the card's pressure,
whose plain-row divergence is within 1e-5), a closed form (every ghost value c and no velocity: p = c) and the
refinement identity (an unconverged pressure plus its refinement is the
solution)."""
import numpy as np
import pytest

from fluidisim_amd import _lib
from fluidisim_amd.volume import Domain3, Error, ProjectionError, Volume3
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

RHO, G, DT = 1025.0, 9.81, 0.01
SCALE = float(np.float32(-RHO / DT))
K1 = float(np.float32(DT / RHO))

OPTIONS = {
    "rbgs_galerkin_fused": dict(),
    "rbgs_galerkin_single": dict(fine_sweeps=_lib.SWEEP_SINGLE),
    "jacobi_rediscretised": dict(smoother=_lib.SMOOTH_JACOBI, coarse_operator=_lib.COARSE_REDISCRETISED,
                                 coarse_sweeps=8),
    "jacobi_pcg": dict(preconditioner=_lib.PRECOND_JACOBI),
}


def scene(nx=64, ny=48, nz=32, dx=0.05, seed=3, terms=True):
    """A tank whose own surface eta carries a small wave and whose total
    surface (eta plus a background elevation of up to 1.5 cells) wets and
    dries cells eta does not; background ghost terms of a few hundred Pa;
    random predicted faces, zero on the walls."""
    rng = np.random.default_rng(seed)
    rest = 0.5 * nz * dx
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    eta = (rest + 0.3 * dx * np.sin(2 * np.pi * X) * np.cos(np.pi * Y)).astype(np.float32).ravel()
    total = (eta + (1.5 * dx * np.sin(2 * np.pi * (X + 0.7 * Y))).astype(np.float32).ravel()).astype(np.float32)
    nu, nv, nw = (nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)
    if terms:
        # smooth, as a background wave's pressure is
        zu = (np.arange(nz)[:, None, None] + 0.5) / nz
        xu = np.arange(nx + 1)[None, None, :] / nx
        yv = np.arange(ny + 1)[None, :, None] / ny
        bg_up = 300.0 * np.sin(2 * np.pi * (X - 0.3 * Y))
        bg_x = 200.0 * np.cos(np.pi * zu) * np.sin(2 * np.pi * xu) * np.ones((1, ny, 1))
        bg_y = 200.0 * np.cos(np.pi * zu) * np.cos(np.pi * yv) * np.ones((1, 1, nx))
        bg = [a.astype(np.float32).ravel() for a in (bg_up, bg_x, bg_y)]
    else:
        bg = [None, None, None]
    u = (0.2 * rng.standard_normal((nz, ny, nx + 1))).astype(np.float32)
    v = (0.2 * rng.standard_normal((nz, ny + 1, nx))).astype(np.float32)
    w = (0.2 * rng.standard_normal((nz + 1, ny, nx))).astype(np.float32)
    u[:, :, [0, nx]] = 0
    v[:, [0, ny], :] = 0
    w[[0, nz]] = 0
    prob = po.Problem(nx, ny, nz, dx, eta, rho=RHO, g_eff=G, rest=rest, homogeneous_ghost=False)
    return prob, total, bg, (u.ravel(), v.ravel(), w.ravel())


def volume(prob, **opts):
    o = dict(preconditioner=_lib.PRECOND_MULTIGRID, smoother=_lib.SMOOTH_RBGS, max_levels=-1, pre_sweeps=2,
             post_sweeps=2, coarse_sweeps=2, coarse_operator=_lib.COARSE_GALERKIN)
    o.update(opts)
    vol = Volume3.configure(Domain3(prob.nx, prob.ny, prob.nz, prob.dx), rho=RHO, g_eff=G, **o)
    vol.set_free_surface(prob.eta, prob.rest)
    return vol


def fields(vol):
    return (vol.pressure().cpu().numpy().astype(np.float64), vol.velocity_u().cpu().numpy(),
            vol.velocity_v().cpu().numpy(), vol.velocity_w().cpu().numpy())


def wet_cells(prob, heights):
    k = np.arange(prob.nz)[:, None, None].astype(np.float32)
    return ((k + np.float32(0.5)) * np.float32(prob.dx) < heights.reshape(prob.ny, prob.nx)[None]).ravel()


def compare(prob, total, bg, faces, card, p64, what):
    """Pressure against the f64 oracle on wet cells; faces against the
    oracle's coupled correction of the card's own pressure (a face to a ghost
    of fraction theta carries a pressure difference times k1 / (theta dx),
    theta down to 1e-3, so the faces follow from the pressure, not from the
    tolerance on it); the reference's plain-row metric of those faces."""
    p, u, v, w = card
    wet = wet_cells(prob, total)
    err = np.linalg.norm(p[wet] - p64[wet]) / np.linalg.norm(p64[wet])
    ur, vr, wr, _, plain = po.correct(prob, *faces, p.astype(np.float32), K1, coupling=(total, *bg))
    umax = max(np.abs(a).max() for a in (ur, vr, wr))
    ferr = max(np.abs(a.astype(np.float64) - b.astype(np.float64)).max() for a, b in ((u, ur), (v, vr), (w, wr)))
    print(f"[{what}] pressure rel L2 {err:.2e}; faces {ferr / umax:.2e} of max speed from the oracle's "
          f"correction of the card's pressure, plain divergence {plain:.2e}", flush=True)
    assert err < 1e-4
    assert ferr <= 1e-5 * umax
    assert plain <= 1e-5
    assert np.all(p[~wet] == 0.0)


@pytest.mark.parametrize("name", list(OPTIONS))
def test_coupled_projection_matches_the_f64_oracle(name):
    prob, total, bg, faces = scene()
    # the ghosts' 1/theta up to 1000 make the system stiff: the oracle's
    # solution to 1e-10 is the reference, the card's to 1e-7 (Jacobi-PCG's
    # error at 1e-6 sits at the 1e-4 bound)
    rep_o, _, _, _, _, p64 = po.project_coupled(prob, *faces, SCALE, K1, total, *bg, opts=po.options(rel_tol=1e-10),
                                                precision=64)
    vol = volume(prob, **OPTIONS[name])
    vol.set_coupling(total, *bg)
    vol.set_velocity(*faces)
    rep = vol.project_mobile3(SCALE, K1, 2000, rel_tol=1e-7, warm=False)
    print(f"[coupled {name}] card {rep.iterations} it, residual {rep.residual:.2e}, plain divergence "
          f"{rep.divergence_plain:.2e}; oracle {rep_o['iterations']} it", flush=True)
    assert not rep.degraded and rep.refinements == 0
    assert rep.divergence_plain <= 1e-5
    compare(prob, total, bg, faces, fields(vol), p64, f"coupled {name}")
    vol.close()


def test_coupled_up_term_is_the_uncoupled_total_surface():
    """Column terms rho g (total - eta) over eta give the problem of the total
    surface uncoupled: the same pressure on the card."""
    prob, total, _, faces = scene(seed=5, terms=False)
    bg_up = (RHO * G * (total.astype(np.float64) - prob.eta.astype(np.float64))).astype(np.float32)
    vol = volume(prob)
    vol.set_coupling(total, bg_up)
    vol.set_velocity(*faces)
    rep = vol.project_mobile3(SCALE, K1, 2000, warm=False)
    p_c = fields(vol)[0]
    vol.close()
    flat = po.Problem(prob.nx, prob.ny, prob.nz, prob.dx, total, rho=RHO, g_eff=G, rest=prob.rest,
                      homogeneous_ghost=False)
    vol = volume(flat)
    vol.set_velocity(*faces)
    rep_u = vol.project_mobile3(SCALE, K1, 2000, warm=False)
    p_u = fields(vol)[0]
    vol.close()
    assert not rep.degraded and not rep_u.degraded
    err = np.linalg.norm(p_c - p_u) / np.linalg.norm(p_u)
    print(f"[coupled up term] rel L2 to the uncoupled total surface {err:.2e}", flush=True)
    assert err < 1e-5


def test_coupled_constant_ghost_values_give_constant_pressure():
    prob, total, _, _ = scene(seed=7, terms=False)
    nx, ny, nz = prob.nx, prob.ny, prob.nz
    c = 2345.0
    rng = np.random.default_rng(7)
    rg = np.float64(RHO) * np.float64(G)
    rest = np.float64(np.float32(prob.rest))
    bg_up = (c - rg * (prob.eta.astype(np.float64) - rest)).astype(np.float32)
    nu, nv, nw = (nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)
    bg_x = (1e4 * rng.standard_normal((nz, ny, nx + 1))).astype(np.float32)
    bg_y = (1e4 * rng.standard_normal((nz, ny + 1, nx))).astype(np.float32)
    W = wet_cells(prob, total).reshape(nz, ny, nx)
    zc = (np.arange(nz, dtype=np.float32) + np.float32(0.5)) * np.float32(prob.dx)
    side = (c - rg * (zc.astype(np.float64) - rest)).astype(np.float32)
    gx = W[:, :, :-1] != W[:, :, 1:]
    gy = W[:, :-1, :] != W[:, 1:, :]
    bg_x[:, :, 1:nx][gx] = np.broadcast_to(side[:, None, None], gx.shape)[gx]
    bg_y[:, 1:ny, :][gy] = np.broadcast_to(side[:, None, None], gy.shape)[gy]
    assert gx.sum() > 100 and gy.sum() > 100
    vol = volume(prob)
    vol.set_coupling(total, bg_up, bg_x.ravel(), bg_y.ravel())
    vol.set_velocity(np.zeros(nu, np.float32), np.zeros(nv, np.float32), np.zeros(nw, np.float32))
    rep = vol.project_mobile3(SCALE, K1, 2000, warm=False)
    p, u, v, w = fields(vol)
    vol.close()
    wet = W.ravel()
    dev = np.abs(p[wet] - c).max()
    speed = max(np.abs(a).max() for a in (u, v, w))
    print(f"[coupled closed form] {rep.iterations} it, max |p - c| {dev:.2e}, max speed {speed:.2e}", flush=True)
    assert not rep.degraded
    assert dev < 1e-6 * c
    # a ghost face's speed is k1 (c - p) / (theta dx), theta down to 1e-3
    assert speed < K1 / (1e-3 * prob.dx) * 1e-6 * c


def test_refine_of_an_unconverged_projection_on_the_card():
    prob, total, bg, faces = scene(seed=9)
    p64 = po.project_coupled(prob, *faces, SCALE, K1, total, *bg, opts=po.options(rel_tol=1e-10), precision=64)[5]
    vol = volume(prob)
    vol.set_coupling(total, *bg)
    vol.set_velocity(*faces)
    first = vol.project_mobile3(SCALE, K1, 2, warm=False)
    p1 = fields(vol)[0]
    wet = wet_cells(prob, total)
    err1 = np.linalg.norm(p1[wet] - p64[wet]) / np.linalg.norm(p64[wet])
    assert first.degraded and err1 > 1e-4
    rep = vol.refine_mobile3(SCALE, K1, 2000)
    print(f"[refine] first solve 2 it (pressure error {err1:.2e}); refinement {rep.iterations} it, residual "
          f"{rep.residual:.2e}", flush=True)
    assert rep.refinements == 1 and not rep.degraded
    compare(prob, total, bg, faces, fields(vol), p64, "refine")
    vol.close()


def test_projection_path_of_the_step_refines_only_after_a_degraded_floor():
    prob, total, bg, faces = scene(seed=11)
    vol = volume(prob)
    vol.set_coupling(total, *bg)
    vol.set_velocity(*faces)
    rep = vol.project_perturbation(SCALE, K1, 2000)
    assert rep.refinements == 0 and not rep.degraded
    vol.set_velocity(*faces)
    with pytest.raises(ProjectionError) as e:
        vol.project_perturbation(SCALE, K1, 2, warm=False)  # degraded without a floor stop
    assert e.value.error == Error.Convergence
    vol.close()


def test_coupling_off_is_the_uncoupled_projection_bit_for_bit():
    prob, total, bg, faces = scene(seed=13)
    ref = volume(prob)
    ref.set_velocity(*faces)
    ref.project_mobile3(SCALE, K1, 2000, warm=False)
    want = fields(ref)
    ref.close()
    vol = volume(prob)
    vol.set_coupling(total, *bg)
    vol.set_velocity(*faces)
    vol.project_mobile3(SCALE, K1, 2000, warm=False)
    vol.set_coupling(None)
    vol.set_velocity(*faces)
    vol.project_mobile3(SCALE, K1, 2000, warm=False)
    got = fields(vol)
    vol.close()
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


def test_coupling_refusals():
    prob, total, bg, faces = scene(nx=16, ny=12, nz=8, dx=0.1, seed=15)
    vol = volume(prob)
    with pytest.raises(ProjectionError) as e:
        vol.set_coupling(total[:-1])
    assert e.value.error == Error.Shape
    vol.set_coupling(total, *bg)
    with pytest.raises(ProjectionError) as e:
        vol.step_surface_mobile(1000, 100)
    assert e.value.error == Error.Domain
    labels = np.full(prob.cells, _lib.LABEL_WATER, np.uint8)
    vol.set_level_set(labels, np.full(prob.cells, -1.0, np.float32))
    with pytest.raises(ProjectionError) as e:
        vol.set_coupling(total)
    assert e.value.error == Error.Domain
    vol.close()
