"""The graded column on the card (fsp_enable_graded, delta3d_graded.rs): the
projection on P^T A P against the oracle's restatement in f64 (pressure
within 1e-5 relative L2, faces within 1e-5 of the largest speed of the
oracle's correction of the card's own pressure), plain and over a coupled
surface; every layer a node against the fine scheme; the guard and the
refusals; and the reference's mobile-step test (tests_delta3d.rs:2111-2130):
30 steps keep the volume, a surface below the cubic layers is refused with
the state restored bit for bit."""
import numpy as np
import pytest

from fluidisim_amd import _lib
from fluidisim_amd.volume import Domain3, Error, ProjectionError, Volume3
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

RHO, G, DT = 1025.0, 9.81, 0.002
SCALE = float(np.float32(-RHO / DT))
K1 = float(np.float32(DT / RHO))
NODES = [0, 6, 11, 15, 18, 20] + list(range(21, 32))


def scene(nx=64, ny=32, nz=32, dx=0.125, h=3.0, seed=3):
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    eta = (h + 0.02 * np.sin(2 * np.pi * X) * np.cos(np.pi * Y)).astype(np.float32).ravel()
    u = (0.1 * rng.standard_normal((nz, ny, nx + 1))).astype(np.float32)
    v = (0.1 * rng.standard_normal((nz, ny + 1, nx))).astype(np.float32)
    w = (0.1 * rng.standard_normal((nz + 1, ny, nx))).astype(np.float32)
    u[:, :, [0, nx]] = 0
    v[:, [0, ny], :] = 0
    w[[0, nz]] = 0
    prob = po.Problem(nx, ny, nz, dx, eta, rho=RHO, g_eff=G, rest=h, homogeneous_ghost=False)
    return prob, (u.ravel(), v.ravel(), w.ravel())


def coupling(prob):
    x = (np.arange(prob.nx) + 0.5) / prob.nx
    y = (np.arange(prob.ny) + 0.5) / prob.ny
    X, Y = np.meshgrid(x, y)
    total = (prob.eta + (0.5 * prob.dx * np.sin(2 * np.pi * (X + 0.3 * Y))).astype(np.float32).ravel())
    bg_up = (200.0 * np.cos(2 * np.pi * X)).astype(np.float32).ravel()
    nu, nv, _ = prob.face_counts
    bg_x = np.full(nu, 50.0, np.float32)
    bg_y = np.full(nv, -30.0, np.float32)
    return total.astype(np.float32), bg_up, bg_x, bg_y


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


def check(prob, faces, card, p64, what, cpl=None):
    p, u, v, w = card
    err = np.linalg.norm(p - p64) / np.linalg.norm(p64)
    ur, vr, wr, _, _ = po.correct(prob, *faces, p.astype(np.float32), K1, coupling=cpl)
    umax = max(np.abs(a).max() for a in (ur, vr, wr))
    ferr = max(np.abs(a.astype(np.float64) - b.astype(np.float64)).max() for a, b in ((u, ur), (v, vr), (w, wr)))
    print(f"[{what}] pressure rel L2 {err:.2e}; faces {ferr / umax:.2e} of max speed from the oracle's "
          f"correction of the card's pressure", flush=True)
    assert err < 1e-5
    assert ferr <= 1e-5 * umax


@pytest.mark.parametrize("coupled", [False, True])
def test_graded_projection_matches_the_f64_oracle(coupled):
    prob, faces = scene()
    cpl = coupling(prob) if coupled else None
    rep_o, _, _, _, _, p64 = po.project_graded(prob, *faces, SCALE, K1, NODES, coupling=cpl,
                                               opts=po.options(rel_tol=1e-9), precision=64)
    vol = volume(prob)
    vol.enable_graded(NODES)
    assert vol.pressure_unknowns_per_column() == len(NODES)
    if coupled:
        vol.set_coupling(*cpl)
    vol.set_velocity(*faces)
    rep = vol.project_mobile3(SCALE, K1, 4000, rel_tol=1e-7, warm=False)
    print(f"[graded coupled={coupled}] card {rep.iterations} it, residual {rep.residual:.2e}, plain divergence "
          f"{rep.divergence_plain:.2e}; oracle {rep_o['iterations']} it", flush=True)
    assert not rep.degraded and rep.divergence_plain <= 1e-5 and rep.backward_error == 0.0
    check(prob, faces, fields(vol), p64, f"graded coupled={coupled}", cpl)
    vol.close()


def test_graded_f32_report_follows_the_oracle():
    """At the reference's tolerance (1e-6) the card's report against the f32
    oracle's: iterations within 2 (the dot products' f64 order differs),
    both accepted, the same pressure within 1e-4."""
    prob, faces = scene(seed=5)
    rep_o, p_o, *_ = po.project_graded(prob, *faces, SCALE, K1, NODES)
    vol = volume(prob)
    vol.enable_graded(NODES)
    vol.set_velocity(*faces)
    rep = vol.project_mobile3(SCALE, K1, 4000, warm=False)
    p = fields(vol)[0]
    print(f"[graded f32] card {rep.iterations} it, oracle {rep_o['iterations']} it", flush=True)
    assert abs(rep.iterations - rep_o["iterations"]) <= 2
    assert not rep.degraded and not rep_o["degraded"]
    assert np.linalg.norm(p - p_o) / np.linalg.norm(p_o) < 1e-4
    vol.close()


def test_every_layer_as_a_node_is_the_fine_scheme():
    prob, faces = scene(nx=32, ny=16, nz=16, dx=0.25, h=3.0, seed=7)
    rep_o, _, _, _, _, p64 = po.project(prob, *faces, SCALE, K1, opts=po.options(multigrid=0, rel_tol=1e-9),
                                        precision=64)
    vol = volume(prob)
    vol.enable_graded(list(range(prob.nz)))
    vol.set_velocity(*faces)
    rep = vol.project_mobile3(SCALE, K1, 4000, rel_tol=1e-7, warm=False)
    assert not rep.degraded
    check(prob, faces, fields(vol), p64, "graded, every layer")
    vol.close()


def test_graded_guard_and_refusals():
    prob, faces = scene(nx=16, ny=8, nz=32, dx=0.125, h=2.7)  # layer 21 (centre 2.6875 m) dry in places
    vol = volume(prob)
    for bad in ([1, 5, 31], [0, 5, 30], [0, 5, 5, 31]):
        with pytest.raises(ProjectionError) as e:
            vol.enable_graded(bad)
        assert e.value.error == Error.Shape
    assert vol.pressure_unknowns_per_column() == prob.nz
    vol.enable_graded(NODES)
    vol.set_velocity(*faces)
    vol.p.fill_(1.5)
    before = [t.clone() for t in (vol.p, vol.u, vol.v, vol.w)]
    with pytest.raises(ProjectionError) as e:
        vol.project_mobile3(SCALE, K1, 100)
    assert e.value.error == Error.Domain
    for a, b in zip(before, (vol.p, vol.u, vol.v, vol.w)):
        assert torch.equal(a, b)
    vol.enable_graded(None)
    assert vol.pressure_unknowns_per_column() == prob.nz
    vol.close()
    solid = np.zeros(prob.cells, np.uint8)
    solid[0] = 1
    vol = volume(prob)
    vol.set_solid(solid)
    with pytest.raises(ProjectionError) as e:
        vol.enable_graded(NODES)
    assert e.value.error == Error.Domain
    vol.close()


def test_graded_mobile_step_keeps_volume_and_refuses_a_low_surface():
    """tests_delta3d.rs:2111-2130: nodes graded below 3 m and cubic from layer
    12 up (32x16x24, dx 0.25), a sine surface at 4 m; 30 steps of 2 ms stay
    accepted with the plain divergence within 1e-5 and keep the perturbation
    volume to 1e-9; a flat surface at 3.1 m (layer 12 dry) is refused with
    u and the surface restored bit for bit."""
    nx, ny, nz, dx = 32, 16, 24, 0.25
    dom = Domain3(nx, ny, nz, dx)
    vol = Volume3.configure(dom, rho=1025.0, g_eff=9.81, preconditioner=_lib.PRECOND_JACOBI)
    i = np.arange(nx)
    eta = np.tile(4 + 0.01 * np.sin(2 * np.pi * (i + 0.5) * dx / 8), ny).astype(np.float32)
    vol.set_free_surface(eta, 4.0)
    vol.enable_graded([0, 5, 9] + list(range(12, 24)))
    assert vol.pressure_unknowns_per_column() == 15
    start, _ = vol.surface_scalars()
    for _ in range(30):
        r = vol.step_surface_mobile(2000, 20000)
        assert not r.degraded and r.divergence_plain <= 1e-5, r
    end, _ = vol.surface_scalars()
    print(f"[graded step] volume {start:.6e} -> {end:.6e}", flush=True)
    assert abs(end - start) < 1e-9
    vol.set_free_surface(np.full(nx * ny, 3.1, np.float32), 3.1)
    before_u = vol.u.clone()
    before_eta = vol.surface()[0].copy()
    with pytest.raises(ProjectionError) as e:
        vol.step_surface_mobile(2000, 20000)
    assert e.value.error == Error.Domain
    assert torch.equal(before_u, vol.u)
    assert np.array_equal(before_eta, vol.surface()[0])
    vol.close()
