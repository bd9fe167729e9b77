"""The surface step on the card (fsp_step_surface_mobile_device and its
stages) against the oracle: advection, extrapolation and transport bit for
bit on the same inputs, the step's scalars, whole-step trajectories within
the projection's tolerance, and the reference's rest, transposition and
refusal tests (tests_delta3d.rs:453-504) run through the C ABI."""
import dataclasses

import numpy as np
import pytest

from fluidisim_amd import _lib
from fluidisim_amd import workloads as W
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
f32 = np.float32


def problem(wl):
    d = wl.domain
    return po.Problem(d.nx, d.ny, d.nz, d.dx, wl.eta, rho=wl.rho, g_eff=wl.g_eff, rest=wl.rest,
                      homogeneous_ghost=wl.homogeneous_ghost, closed_lid=wl.closed_lid, solid=wl.solid)


def oracle_options(wl):
    o = wl.options
    return po.options(multigrid=1, smoother=o["smoother"], pre=o["pre_sweeps"], post=o["post_sweeps"],
                      coarse=o["coarse_sweeps"], coarse_op=o["coarse_operator"])


def random_faces(wl, seed, scale=0.5):
    rng = np.random.default_rng(seed)
    return [(rng.standard_normal(n) * scale).astype(f32) for n in wl.domain.faces]


def bits(a):
    return np.asarray(a, f32).view(np.uint32)


def host(t):
    return t.detach().cpu().numpy()


def test_advection_is_bit_identical_to_the_oracle():
    wl = W.slosh(64, 32, 64, seed=3)
    vol = W.build_volume(wl)
    u, v, w = random_faces(wl, 1)
    got = vol.advect_for_trials(u, v, w, 1e-3)
    want = po.advect(problem(wl), u, v, w, 1e-3)
    for name, g, o in zip("uvw", got, want):
        assert np.array_equal(bits(host(g)), bits(o)), name
    vol.close()


def test_extrapolation_is_bit_identical_to_the_oracle():
    wl = W.slosh(64, 32, 64, seed=4)
    wl.eta = (wl.eta + np.random.default_rng(2).uniform(-0.2, 0.2, wl.eta.size)).astype(f32)
    vol = W.build_volume(wl)
    u, v, w = random_faces(wl, 2)
    got = vol.extrapolate_for_trials(u, v, w)
    want = po.extrapolate(problem(wl), u, v, w)
    for name, g, o in zip("uvw", got, want):
        assert np.array_equal(bits(host(g)), bits(o)), name
        assert not np.array_equal(o, {"u": u, "v": v, "w": w}[name])  # the stage did something
    vol.close()


def test_transport_is_bit_identical_to_the_oracle():
    wl = W.slosh(64, 32, 64, seed=5)
    vol = W.build_volume(wl)
    u, v, _ = random_faces(wl, 3)
    fx, fy = vol.transport_for_trials(u, v, 0.02)
    eta, ero = vol.surface()
    oe, oro, ofx, ofy = po.transport(problem(wl), u, v, 0.02)
    assert np.array_equal(bits(host(fx)), bits(ofx)) and np.array_equal(bits(host(fy)), bits(ofy))
    assert np.array_equal(bits(eta), bits(oe)) and np.array_equal(bits(ero), bits(oro))
    vol.close()


def test_surface_scalars_match_the_oracle():
    wl = W.slosh(64, 32, 64, seed=6)
    vol = W.build_volume(wl)
    volume, ok = vol.surface_scalars()
    want = po.perturbation_volume(problem(wl))
    assert ok == po.mobile_in_bounds(problem(wl)) is True
    assert abs(volume - want) <= 1e-12 * max(1.0, abs(want)) + 1e-12
    wl.eta = np.full_like(wl.eta, 0.5 * wl.domain.dx)
    low = W.build_volume(wl)
    assert low.surface_scalars()[1] is False
    vol.close()
    low.close()


def test_steps_follow_the_oracle_trajectory():
    """Ten steps of the sloshing tank (spheres, warm-started projections)
    against the f32 oracle with the same V-cycle: heights within 2e-6 m,
    velocities and pressure within 1e-3 relative L2, the balance closed."""
    wl = W.slosh(64, 32, 64, seed=7, amplitude=0.03)
    vol = W.build_volume(wl)
    reps, bals = [], []
    for _ in range(10):
        reps.append(vol.step_surface_mobile(1000, 4000))
        bals.append(vol.balance)
    eta, ero = vol.surface()
    d = wl.domain
    out = po.step_surface(problem(wl), wl.us, wl.vs, wl.ws, np.zeros(d.cells, f32), 10, 1000, 4000,
                          oracle_options(wl))
    assert out["status"] == 0 and out["done"] == 10
    print("card iterations", [r.iterations for r in reps], "oracle", [r["iterations"] for r in out["reports"]])
    de = np.abs(eta - out["eta"]).max()
    errs = {k: np.linalg.norm(host(t) - out[k]) / max(np.linalg.norm(out[k]), 1e-30)
            for k, t in (("u", vol.u), ("v", vol.v), ("w", vol.w), ("p", vol.p))}
    print(f"max |eta - oracle| {de:.2e} m; rel L2 {errs}")
    assert de <= 2e-6
    assert all(e <= 1e-3 for e in errs.values()), errs
    assert all(not r.degraded for r in reps)
    scale = abs(out["balances"][0]["volume"]) + 1e-3
    for b, ob in zip(bals, out["balances"]):
        assert b["perturbation_in"] == 0.0 and abs(b["residual"]) <= 1e-9 * scale
        assert abs(b["volume"] - ob["volume"]) <= 1e-9 * scale
    vol.close()


def small_tank(nx, ny, nz, dx, eta, rest, **opts):
    from fluidisim_amd.volume import Domain3, Volume3
    vol = Volume3.configure(Domain3(nx, ny, nz, dx), rho=1025.0, g_eff=9.81, **opts)
    vol.set_surface_mode(False, False)
    vol.set_free_surface(np.asarray(eta, f32), rest)
    return vol


def test_rest_stays_at_rest_bit_for_bit():
    """tests_delta3d.rs:467-473 on the card."""
    vol = small_tank(12, 7, 12, 0.25, np.full(84, 2.013, f32), 2.013)
    for _ in range(20):
        r = vol.step_surface_mobile(1000, 4000)
        assert r.iterations == 0
    eta, _ = vol.surface()
    assert np.all(bits(eta) == f32(2.013).view(np.uint32))
    for t in (vol.u, vol.v, vol.w, vol.p):
        assert np.all(bits(host(t)) == 0)
    vol.close()


def test_transposition_symmetry_on_the_card():
    """tests_delta3d.rs:453-466: x and y swapped, 300 steps, within 2e-6 m."""
    c = np.arange(84)
    eta = (f32(2.0) + f32(0.05) * np.cos((f32(np.pi) * (c % 12).astype(f32) / f32(12.0)).astype(f32))).astype(f32)
    trans = eta.reshape(7, 12).T.copy().ravel()
    x = small_tank(12, 7, 12, 0.25, eta, 2.0)
    y = small_tank(7, 12, 12, 0.25, trans, 2.0)
    for _ in range(300):
        x.step_surface_mobile(1000, 4000)
        y.step_surface_mobile(1000, 4000)
    ex, ey = x.surface()[0].reshape(7, 12), y.surface()[0].reshape(12, 7)
    worst = np.abs(ex - ey.T).max()
    print(f"transposition x-y on the card: {worst:.2e} m")
    assert worst < 2e-6
    x.close()
    y.close()


def test_refusals_restore_the_state_bit_for_bit():
    """tests_delta3d.rs:477-490 through the C ABI: Convergence, Domain and
    NotFinite leave u, v, w, p and the surface as they were."""
    from fluidisim_amd.volume import Error, ProjectionError
    c = np.arange(80)
    eta = (f32(2.0) + f32(0.1) * np.sin((c.astype(f32) * f32(0.7)).astype(f32))).astype(f32)
    vol = small_tank(16, 5, 12, 0.25, eta, 2.0)
    for _ in range(5):
        vol.step_surface_mobile(1000, 4000)

    def state():
        e, r = vol.surface()
        return [bits(host(t)).copy() for t in (vol.u, vol.v, vol.w, vol.p)] + [bits(e).copy(), bits(r).copy()]

    before = state()
    with pytest.raises(ProjectionError) as err:
        vol.step_surface_mobile(1000, 0)
    assert err.value.error == Error.Convergence
    assert all(np.array_equal(a, b) for a, b in zip(state(), before))
    with pytest.raises(ProjectionError) as err:
        vol.step_surface_mobile(0, 4000)
    assert err.value.error == Error.NotFinite
    for height in (0.49, 2.76):
        vol.set_free_surface(np.full(80, height, f32), 2.0)
        before = state()
        with pytest.raises(ProjectionError) as err:
            vol.step_surface_mobile(1000, 4000)
        assert err.value.error == Error.Domain
        assert all(np.array_equal(a, b) for a, b in zip(state(), before))
    vol.close()
