"""The oracle's surface step (oracle.hpp step_surface_mobile and its stages,
delta3d_mobile.rs:345-394, :615-927) pinned by what the mathematics and the
reference's own tests fix: the reference's rest, transposition and refusal
tests (tests_delta3d.rs:453-504), closed forms of each advection term, a
hand-worked extrapolation column, the wetted discharge of a uniform flow and
the conservation of the perturbation volume."""
import dataclasses

import numpy as np
import pytest

from oracle import pyoracle as po

f32 = np.float32


def tank(nx, ny, nz, dx, eta, rest, solid=None):
    return po.Problem(nx, ny, nz, dx, np.asarray(eta, f32), rho=1025.0, g_eff=9.81, rest=rest,
                      homogeneous_ghost=False, solid=solid)


def zeros(prob):
    nu, nv, nw = prob.face_counts
    return [np.zeros(n, f32) for n in (nu, nv, nw, prob.cells)]


# ---------------------------------------------------------------- the reference's step tests
def test_rest_stays_at_rest_bit_for_bit():
    """tests_delta3d.rs:467-473: a flat surface at rest, no velocity: 0
    iterations, every height and every field unchanged to the bit."""
    prob = tank(12, 7, 12, 0.25, np.full(84, 2.013, f32), 2.013)
    out = po.step_surface(prob, *zeros(prob), 50, 1000)
    assert out["status"] == 0 and out["done"] == 50
    assert all(r["iterations"] == 0 for r in out["reports"])
    assert np.all(out["eta"].view(np.uint32) == f32(2.013).view(np.uint32))
    for k in ("u", "v", "w", "p"):
        assert np.all(out[k].view(np.uint32) == 0)


def transposed_pair():
    c = np.arange(84)
    eta = (f32(2.0) + f32(0.05) * np.cos((f32(np.pi) * (c % 12).astype(f32) / f32(12.0)).astype(f32))).astype(f32)
    trans = np.zeros(84, f32)
    for j in range(7):
        for i in range(12):
            trans[i * 7 + j] = eta[j * 12 + i]
    return tank(12, 7, 12, 0.25, eta, 2.0), tank(7, 12, 12, 0.25, trans, 2.0)


def test_transposition_symmetry_over_300_steps():
    """tests_delta3d.rs:453-466: the same tank with x and y swapped keeps
    the transposed surface within 2e-6 m over 300 steps."""
    px, py = transposed_pair()
    ox = po.step_surface(px, *zeros(px), 300, 1000)
    oy = po.step_surface(py, *zeros(py), 300, 1000)
    assert ox["status"] == 0 and oy["status"] == 0 and ox["done"] == oy["done"] == 300
    worst = np.abs(ox["eta"].reshape(7, 12) - oy["eta"].reshape(12, 7).T).max()
    print(f"transposition x-y: {worst:.2e} m")
    assert worst < 2e-6


def test_refusals_restore_the_state_bit_for_bit():
    """tests_delta3d.rs:477-490: Convergence (no iterations allowed), Domain
    (surface below the floor or above the top) and NotFinite (zero duration)
    leave u, v, w, p, eta and eta_roundoff as they were."""
    c = np.arange(80)
    eta = (f32(2.0) + f32(0.1) * np.sin((c.astype(f32) * f32(0.7)).astype(f32))).astype(f32)
    prob = tank(16, 5, 12, 0.25, eta, 2.0)
    warm = po.step_surface(prob, *zeros(prob), 5, 1000)
    assert warm["status"] == 0
    state = dict(warm)
    moved = dataclasses.replace(prob, eta=state["eta"], eta_roundoff=state["eta_roundoff"])
    out = po.step_surface(moved, state["u"], state["v"], state["w"], state["p"], 1, 1000, max_iters=0)
    assert out["status"] == po.ERR_CONVERGENCE
    for k in ("u", "v", "w", "p", "eta", "eta_roundoff"):
        assert np.array_equal(out[k].view(np.uint32), state[k].view(np.uint32)), k
    for height in (0.49, 2.76):
        flat = tank(16, 5, 12, 0.25, np.full(80, height, f32), 2.0)
        out = po.step_surface(flat, state["u"], state["v"], state["w"], state["p"], 1, 1000)
        assert out["status"] == po.ERR_DOMAIN and out["done"] == 0
        assert np.all(out["eta"] == f32(height)) and np.array_equal(out["u"], state["u"])
    out = po.step_surface(moved, state["u"], state["v"], state["w"], state["p"], 1, 0)
    assert out["status"] == po.ERR_NOT_FINITE


def test_the_guard_after_transport_is_exercised_and_restores():
    """tests_delta3d.rs:491-503: a crest that reaches the highest admitted
    height during the run is refused (Domain) with the state of the step
    before it intact."""
    i = np.arange(16)
    eta = (2.0 + 0.125 * np.cos(np.pi * (i + 0.5) / 16.0)).astype(f32)
    prob = tank(16, 1, 18, 0.125, eta, 2.0)
    u, v, w, p = zeros(prob)
    cur = prob
    for step in range(1604):
        out = po.step_surface(cur, u, v, w, p, 1, 1000)
        if out["status"] != 0:
            assert out["status"] == po.ERR_DOMAIN, out["status"]
            assert np.array_equal(out["eta"], cur.eta) and np.array_equal(out["u"], u) and np.array_equal(out["p"], p)
            break
        u, v, w, p = out["u"], out["v"], out["w"], out["p"]
        cur = dataclasses.replace(cur, eta=out["eta"], eta_roundoff=out["eta_roundoff"])
    else:
        pytest.fail("the guard after transport was never exercised")


def test_the_step_conserves_the_perturbation_volume():
    """delta3d_mobile.rs:861-877: closed walls, so the volume change equals
    the (zero) wall discharge: the balance residual is roundoff."""
    px, _ = transposed_pair()
    out = po.step_surface(px, *zeros(px), 40, 1000)
    assert out["status"] == 0
    scale = 0.05 * 84 * 0.25 * 0.25
    for b in out["balances"]:
        assert b["perturbation_in"] == 0.0
        assert abs(b["residual"]) < 1e-9 * scale


# ---------------------------------------------------------------- advection: closed forms per term
def box(nx=8, ny=6, nz=7, dx=0.5):
    return tank(nx, ny, nz, dx, np.full(nx * ny, (nz - 1) * dx, f32), (nz - 1) * dx)


def face_coords(prob):
    nx, ny, nz = prob.nx, prob.ny, prob.nz
    ku, ju, iu = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx + 1), indexing="ij")
    kv, jv, iv = np.meshgrid(np.arange(nz), np.arange(ny + 1), np.arange(nx), indexing="ij")
    kw, jw, iw = np.meshgrid(np.arange(nz + 1), np.arange(ny), np.arange(nx), indexing="ij")
    return (iu, ju, ku), (iv, jv, kv), (iw, jw, kw)


def test_uniform_flow_is_not_advected():
    """Every centred difference of a constant field is 0: us = u bit for bit."""
    prob = box()
    nu, nv, nw = prob.face_counts
    u, v, w = np.full(nu, 0.7, f32), np.full(nv, -0.3, f32), np.full(nw, 0.2, f32)
    us, vs, ws = po.advect(prob, u, v, w, 0.01)
    assert np.array_equal(us, u) and np.array_equal(vs, v) and np.array_equal(ws, w)


def test_each_advection_term_has_its_closed_form():
    """u = a x: u_t = -u a (self term); u = b y with v = V: -V b (transverse);
    u = c z with w = W: -W c (vertical); the same for v and w by symmetry.
    Checked on interior faces (the wall rows fall back to one-sided values)."""
    prob = box()
    dx, dt = 0.5, 0.01
    (iu, ju, ku), (iv, jv, kv), (iw, jw, kw) = face_coords(prob)
    nx, ny, nz = prob.nx, prob.ny, prob.nz
    a = 0.3
    inner_u = (iu > 0) & (iu < nx) & (ju > 0) & (ju < ny - 1) & (ku > 0) & (ku < nz - 1)
    # self term
    u = (a * iu * dx).astype(f32).ravel()
    us, _, _ = po.advect(prob, u, np.zeros(prob.face_counts[1], f32), np.zeros(prob.face_counts[2], f32), dt)
    want = u * (1 - dt * a)
    m = inner_u.ravel()
    assert np.allclose(us[m], want[m], rtol=2e-6, atol=1e-7)
    # transverse term: u = b y, v = V
    b, V = 0.4, 0.8
    u = (b * ju * dx).astype(f32).ravel()
    v = np.full(prob.face_counts[1], V, f32)
    us, vs, _ = po.advect(prob, u, v, np.zeros(prob.face_counts[2], f32), dt)
    assert np.allclose(us[m], (u - dt * V * b)[m], rtol=2e-6, atol=1e-6)
    # vertical term: u = c z, w = W
    c, Wv = -0.6, 0.5
    u = (c * ku * dx).astype(f32).ravel()
    w = np.full(prob.face_counts[2], Wv, f32)
    us, _, ws = po.advect(prob, u, np.zeros(prob.face_counts[1], f32), w, dt)
    assert np.allclose(us[m], (u - dt * Wv * c)[m], rtol=2e-6, atol=1e-6)
    # w faces: w = a z (self term), with u = U (no transverse change)
    inner_w = ((iw > 0) & (iw < nx - 1) & (jw > 0) & (jw < ny - 1) & (kw > 0) & (kw < nz)).ravel()
    w = (a * kw * dx).astype(f32).ravel()
    _, _, ws = po.advect(prob, np.full(prob.face_counts[0], 0.9, f32), np.zeros(prob.face_counts[1], f32), w, dt)
    assert np.allclose(ws[inner_w], (w * (1 - dt * a))[inner_w], rtol=2e-6, atol=1e-6)
    # v faces: v = a y (self term)
    inner_v = ((iv > 0) & (iv < nx - 1) & (jv > 0) & (jv < ny) & (kv > 0) & (kv < nz - 1)).ravel()
    v = (a * jv * dx).astype(f32).ravel()
    _, vs, _ = po.advect(prob, np.zeros(prob.face_counts[0], f32), v, np.zeros(prob.face_counts[2], f32), dt)
    assert np.allclose(vs[inner_v], (v * (1 - dt * a))[inner_v], rtol=2e-6, atol=1e-6)


def test_walls_closed_faces_and_the_top_are_not_advected():
    rng = np.random.default_rng(3)
    nx, ny, nz, dx = 8, 6, 7, 0.5
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[2:4, 2:4, 3:5] = 1
    prob = tank(nx, ny, nz, dx, np.full(nx * ny, (nz - 1) * dx, f32), (nz - 1) * dx, solid=solid.ravel())
    nu, nv, nw = prob.face_counts
    u, v, w = (rng.standard_normal(n).astype(f32) for n in (nu, nv, nw))
    us, vs, ws = po.advect(prob, u, v, w, 0.01)
    (iu, ju, ku), (iv, jv, kv), (iw, jw, kw) = face_coords(prob)
    S = solid.astype(bool)
    closed_u = (iu == 0) | (iu == nx)
    closed_u[:, :, 1:nx] |= S[:, :, :-1] | S[:, :, 1:]
    closed_w = (kw == 0) | (kw == nz)
    closed_w[1:nz] |= S[:-1] | S[1:]
    assert np.array_equal(us[closed_u.ravel()], u[closed_u.ravel()])
    assert np.array_equal(ws[closed_w.ravel()], w[closed_w.ravel()])
    assert not np.array_equal(us[~closed_u.ravel()], u[~closed_u.ravel()])


# ---------------------------------------------------------------- extrapolation
def test_extrapolation_of_a_hand_worked_column():
    """nx = 2, ny = 1, nz = 6, dx = 1, surface at 2.6: cells 0-2 are wet
    ((k + 0.5) < 2.6); w faces 1-3 are solved (above a wet cell), faces 4-6
    take face 3's value; u faces of column i = 1 between two wet cells (k <=
    2) are solved, those above take the value at k = 2. Face i = 0 and the
    bottom w face are never touched."""
    prob = tank(2, 1, 6, 1.0, np.array([2.6, 2.6], f32), 2.0)
    nu, nv, nw = prob.face_counts
    u = np.arange(nu, dtype=f32) + 1
    v = np.arange(nv, dtype=f32) + 100
    w = np.arange(nw, dtype=f32) + 200
    eu, ev, ew = po.extrapolate(prob, u, v, w)
    fu = lambda i, k: k * 3 + i
    fw = lambda i, k: k * 2 + i
    want_u, want_w = u.copy(), w.copy()
    for k in (3, 4, 5):
        want_u[fu(1, k)] = u[fu(1, 2)]
    for i in (0, 1):
        for k in (4, 5, 6):
            want_w[fw(i, k)] = w[fw(i, 3)]
    assert np.array_equal(eu, want_u) and np.array_equal(ew, want_w) and np.array_equal(ev, v)


def test_extrapolation_is_idempotent_and_keeps_solved_faces():
    rng = np.random.default_rng(5)
    nx, ny, nz, dx = 9, 7, 10, 0.25
    eta = (1.4 + 0.6 * rng.random(nx * ny)).astype(f32)
    solid = (rng.random(nx * ny * nz) > 0.93).astype(np.uint8)
    prob = tank(nx, ny, nz, dx, eta, 1.5, solid=solid)
    nu, nv, nw = prob.face_counts
    u, v, w = (rng.standard_normal(n).astype(f32) for n in (nu, nv, nw))
    once = po.extrapolate(prob, u, v, w)
    twice = po.extrapolate(prob, *once)
    for a, b in zip(once, twice):
        assert np.array_equal(a, b)
    # the faces the extrapolation changed are all above the water
    changed = once[2] != w
    kw = np.repeat(np.arange(nz + 1), nx * ny)
    col_eta = np.tile(eta, nz + 1)
    assert np.all((kw[changed] - 0.5) * dx >= col_eta[changed] - 1e-6 - dx)


# ---------------------------------------------------------------- transport
def test_uniform_flow_discharge_and_heights():
    """Uniform u = U over a flat surface H: every inner x discharge is U H
    (the wetted fractions of a column sum to H / dx); only the first and last
    columns change height, by -+ factor U H."""
    nx, ny, nz, dx = 10, 3, 8, 0.5
    H = 2.3
    prob = tank(nx, ny, nz, dx, np.full(nx * ny, H, f32), H)
    nu, nv, nw = prob.face_counts
    U = 0.4
    eta, ero, fx, fy = po.transport(prob, np.full(nu, U, f32), np.zeros(nv, f32), 0.01)
    fx = fx.reshape(ny, nx + 1)
    assert np.all(fx[:, 0] == 0) and np.all(fx[:, nx] == 0) and np.all(fy == 0)
    assert np.allclose(fx[:, 1:nx], U * H, rtol=1e-6)
    e = eta.reshape(ny, nx) - ero.reshape(ny, nx)
    assert np.allclose(e[:, 0], H - 0.01 * U * H, rtol=0, atol=1e-6)
    assert np.allclose(e[:, -1], H + 0.01 * U * H, rtol=0, atol=1e-6)
    assert np.allclose(e[:, 1:-1], H, rtol=0, atol=1e-6)


def test_transport_of_still_water_changes_nothing_and_conserves_volume():
    rng = np.random.default_rng(7)
    nx, ny, nz, dx = 9, 7, 10, 0.25
    eta = (1.4 + 0.6 * rng.random(nx * ny)).astype(f32)
    prob = tank(nx, ny, nz, dx, eta, 1.5)
    nu, nv, nw = prob.face_counts
    e0, r0, _, _ = po.transport(prob, np.zeros(nu, f32), np.zeros(nv, f32), 0.01)
    assert np.array_equal(e0, eta) and np.all(r0 == 0)
    u, v = rng.standard_normal(nu).astype(f32), rng.standard_normal(nv).astype(f32)
    e1, r1, _, _ = po.transport(prob, u, v, 0.01)
    before = po.perturbation_volume(prob)
    after = po.perturbation_volume(dataclasses.replace(prob, eta=e1, eta_roundoff=r1))
    assert abs(after - before) < 1e-10


def test_bounds_and_volume_known_answers():
    prob = tank(4, 3, 10, 0.5, np.full(12, 2.0, f32), 1.5)
    assert po.mobile_in_bounds(prob)
    assert po.perturbation_volume(prob) == pytest.approx(12 * 0.5 * 0.25)
    assert not po.mobile_in_bounds(tank(4, 3, 10, 0.5, np.full(12, 0.99, f32), 1.5))  # below 2 dx
    assert not po.mobile_in_bounds(tank(4, 3, 10, 0.5, np.full(12, 4.51, f32), 1.5))  # above (nz-1) dx
    solid = np.zeros((10, 3, 4), np.uint8)
    solid[2, 1, 1] = 1  # floor of that column: 1.5 m, so the surface needs 2.5 m
    assert not po.mobile_in_bounds(tank(4, 3, 10, 0.5, np.full(12, 2.0, f32), 1.5, solid=solid.ravel()))
    assert po.mobile_in_bounds(tank(4, 3, 10, 0.5, np.full(12, 2.5, f32), 1.5, solid=solid.ravel()))
