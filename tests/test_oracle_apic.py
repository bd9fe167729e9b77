"""The oracle's APIC-3D projection (apic3d.rs:887-1080) pinned without the
reference: the hydrostatic closed form of a pool under gravity (any wrong
sign, scale or ghost fraction breaks it), a symmetric positive operator on
random labels and level sets, and the conjugate gradient reaching the dense
solution of that operator."""
import numpy as np
import pytest

from oracle import pyoracle as po


def pool(nx, ny, nz, dx, h):
    """Water below the flat surface z = h, air above; phi = z_c - h."""
    zc = (np.arange(nz) + 0.5) * dx
    phi = np.broadcast_to((zc - h)[:, None, None], (nz, ny, nx)).astype(np.float32).ravel()
    label = np.where(phi < 0, po.APIC_WATER, po.APIC_AIR).astype(np.uint8)
    return label, phi


def gravity_faces(nx, ny, nz, g_dt):
    """The predicted faces after gravity and the walls (apic3d.rs:757-768):
    every w face -g dt, then the domain's bottom and lid faces 0; u, v 0."""
    u = np.zeros((nx + 1) * ny * nz, np.float32)
    v = np.zeros(nx * (ny + 1) * nz, np.float32)
    w = np.full((nz + 1, ny, nx), -g_dt, np.float32)
    w[0] = 0
    w[nz] = 0
    return u, v, w.ravel()


@pytest.mark.parametrize("precision", [32, 64])
def test_pool_at_rest_is_hydrostatic(precision):
    """Gravity on a flat pool: p = rho g (h - z_c) in the water, every water
    face back to 0, the divergence metric 0 (closed form of the ghost-fluid
    pressure with theta = (h - z_top)/dx at the surface)."""
    nx, ny, nz, dx, rho, g, dt = 6, 5, 12, 0.05, 1000.0, 9.81, 0.004
    h = 6.3 * dx
    label, phi = pool(nx, ny, nz, dx, h)
    u, v, w = gravity_faces(nx, ny, nz, np.float32(g * dt))
    p, u2, v2, w2, it, res, div = po.apic_project(nx, ny, nz, dx, rho, dt, label, phi, u, v, w, precision)
    zc = np.broadcast_to(((np.arange(nz) + 0.5) * dx)[:, None, None], (nz, ny, nx)).ravel()
    water = label == po.APIC_WATER
    want = rho * g * (h - zc[water])
    # the conjugate gradient stops at |r| <= 1e-6 |b| (apic3d.rs:716)
    assert np.abs(p[water] - want).max() <= 1e-5 * want.max()
    assert np.all(p[~water] == 0)
    wz = w2.reshape(nz + 1, ny, nx)
    top = int(np.count_nonzero(water) // (nx * ny))  # water layers
    assert np.abs(wz[1:top + 1]).max() <= 1e-4 * g * dt
    assert it > 0 and res <= 1e-6


def test_operator_is_symmetric_and_positive():
    """The ghost-fluid operator is symmetric (the air fractions sit on the
    diagonal) and positive on water cells, for random labels and level sets."""
    rng = np.random.default_rng(4)
    nx, ny, nz = 5, 4, 6
    n = nx * ny * nz
    label = rng.choice([po.APIC_AIR, po.APIC_WATER, po.APIC_SOLID], size=n, p=[0.3, 0.6, 0.1]).astype(np.uint8)
    phi = np.where(label == po.APIC_WATER, -rng.uniform(0.01, 1, n), rng.uniform(0.01, 1, n)).astype(np.float32)
    A = np.stack([po.apic_apply(nx, ny, nz, label, phi, e) for e in np.eye(n)], axis=1)
    water = label == po.APIC_WATER
    Aw = A[np.ix_(water, water)]
    assert np.abs(A[~water]).max() == 0
    assert np.abs(Aw - Aw.T).max() <= 1e-12 * np.abs(Aw).max()
    assert np.linalg.eigvalsh(Aw).min() > 0


def test_conjugate_gradient_reaches_the_dense_solution():
    """A droplet over a pool with a solid block and random predicted faces:
    the f64 projection's pressure is A^-1 b of the oracle's own operator and
    right-hand side (b = -(rho dx^2/dt) div u* / dx on water cells), and the
    corrected field is divergence-free on the inner water cells."""
    rng = np.random.default_rng(7)
    nx, ny, nz, dx, rho, dt = 6, 6, 8, 0.1, 1000.0, 0.01
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dx
    z = (np.arange(nz) + 0.5) * dx
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    phi_pool = Z - 0.25
    phi_drop = np.sqrt((X - 0.3) ** 2 + (Y - 0.3) ** 2 + (Z - 0.6) ** 2) - 0.15
    phi = np.minimum(phi_pool, phi_drop).astype(np.float32).ravel()
    label = np.where(phi < 0, po.APIC_WATER, po.APIC_AIR).astype(np.uint8)
    label[((X < 0.2) & (Y < 0.2) & (Z < 0.2)).ravel()] = po.APIC_SOLID
    u, v, w = (rng.uniform(-1, 1, m).astype(np.float32) for m in ((nx + 1) * ny * nz, nx * (ny + 1) * nz,
                                                                   nx * ny * (nz + 1)))
    p, u2, v2, w2, it, res, div = po.apic_project(nx, ny, nz, dx, rho, dt, label, phi, u, v, w, 64)
    n = nx * ny * nz
    water = label == po.APIC_WATER
    A = np.stack([po.apic_apply(nx, ny, nz, label, phi, e) for e in np.eye(n)], axis=1)[np.ix_(water, water)]
    uu = u.astype(np.float64).reshape(nz, ny, nx + 1)
    vv = v.astype(np.float64).reshape(nz, ny + 1, nx)
    ww = w.astype(np.float64).reshape(nz + 1, ny, nx)
    dv = (uu[:, :, 1:] - uu[:, :, :-1]) + (vv[:, 1:, :] - vv[:, :-1, :]) + (ww[1:] - ww[:-1])
    b = (-rho * dx * dx / dt) * dv.ravel()[water] / dx
    p_dense = np.linalg.solve(A, b)
    # the conjugate gradient stops at |r| <= 1e-6 |b| (apic3d.rs:716)
    assert np.abs(p[water] - p_dense).max() <= 1e-5 * np.abs(p_dense).max()
    assert res <= 1e-6 and div <= 1e-5
