"""The projection workloads (synthetic, seeded), in reference orientation.

Sizes are given as (X, Y, Z) with Y vertical and mapped to the reference's
Domain3 {nx: X, ny: Z, nz: Y} (z up). Every workload has dx = 1 and
rho = dt = 1, so scale = -rho/dt = -1 and k1 = dt/rho = 1 (b = -div u*).

  closed_box    all fluid, closed box (Neumann on all six faces), faces
                i.i.d. U[-1,1]; 3-level V-cycle, 2 RBGS pre/post sweeps.
  tank          free surface h(x,z) = Y/2 + 4 sines, air (p = 0) above,
                8 solid spheres, 16-mode smooth velocity of RMS 1;
                multigrid over every coarse level the grid allows.
Both use Galerkin coarse operators (P^T A P / 16).

  slosh         the surface step's workload (physical units): an 8 m wide
                tank, depth 2/3 of its height, a surface raised by a
                cosine mode, solid spheres well below the surface, water at
                rest; rho = 1025, g = 9.81, surface ghosts from the heights.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .volume import Domain3, generate


@dataclass
class Workload:
    name: str
    domain: Domain3
    eta: np.ndarray
    solid: np.ndarray | None
    us: np.ndarray
    vs: np.ndarray
    ws: np.ndarray
    homogeneous_ghost: bool
    closed_lid: bool
    scale: float = -1.0
    k1: float = 1.0
    rho: float = 1.0
    g_eff: float = 9.81
    rest: float = 0.0
    options: dict = field(default_factory=dict)
    seed: int = 0
    # face apertures in [0, 1] (the reference's cut geometry), or None: derived
    open_u: np.ndarray | None = None
    open_v: np.ndarray | None = None
    open_w: np.ndarray | None = None

    def describe(self):
        d = self.domain
        return {"workload": self.name, "nx": d.nx, "ny": d.ny, "nz": d.nz, "cells": d.cells, "seed": self.seed}


def closed_box(n: int = 64, seed: int = 1, levels: int = 2) -> Workload:
    eta, solid, us, vs, ws = generate(0, n, n, n, seed)
    return Workload(f"closed_box_{n}", Domain3(n, n, n, 1.0), eta, None, us, vs, ws,
                    homogeneous_ghost=True, closed_lid=True, rest=float(n),
                    options=dict(preconditioner=_lib.PRECOND_MULTIGRID, smoother=_lib.SMOOTH_RBGS,
                                 max_levels=levels, pre_sweeps=2, post_sweeps=2, coarse_sweeps=8,
                                 coarse_operator=_lib.COARSE_GALERKIN),
                    seed=seed)


def tank(X: int = 128, Y: int = 128, Z: int = 128, seed: int = 1) -> Workload:
    nx, ny, nz = X, Z, Y
    eta, solid, us, vs, ws = generate(1, nx, ny, nz, seed)
    return Workload(f"tank_{X}x{Y}x{Z}", Domain3(nx, ny, nz, 1.0), eta, solid, us, vs, ws,
                    homogeneous_ghost=True, closed_lid=True, rest=float(nz) / 2,
                    options=dict(preconditioner=_lib.PRECOND_MULTIGRID, smoother=_lib.SMOOTH_RBGS,
                                 max_levels=-1, pre_sweeps=2, post_sweeps=2, coarse_sweeps=8,
                                 coarse_operator=_lib.COARSE_GALERKIN),
                    seed=seed)


def build_volume(wl: Workload, **option_overrides):
    """A configured Volume3 holding the workload's geometry and predicted faces."""
    from .volume import Volume3
    opts = dict(wl.options)
    opts.update(option_overrides)
    vol = Volume3.configure(wl.domain, rho=wl.rho, g_eff=wl.g_eff, **opts)
    vol.set_surface_mode(wl.homogeneous_ghost, wl.closed_lid)
    vol.set_solid(wl.solid, wl.open_u, wl.open_v, wl.open_w)
    vol.set_free_surface(wl.eta, wl.rest)
    vol.set_velocity(wl.us, wl.vs, wl.ws)
    return vol


PATCH_EDGES = (32, 48, 64, 96, 128)
PATCH_WEIGHTS = (0.3, 0.25, 0.2, 0.15, 0.1)


def patches(count: int = 48, seed: int = 1):
    """Independent delta patches: edge lengths drawn per axis from 32..128
    with weights 0.3/0.25/0.2/0.15/0.1, each the free-surface tank generator
    at its own size and seed."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(count):
        X, Y, Z = (int(v) for v in rng.choice(PATCH_EDGES, size=3, p=PATCH_WEIGHTS))
        out.append(tank(X, Y, Z, seed=seed * 1000 + q))
    return out


def slosh(X: int = 64, Y: int = 32, Z: int = 64, seed: int = 1, spheres: int = 4, amplitude: float = 0.02) -> Workload:
    """Tank of X x Z columns, Y cells high (reference nx = X, ny = Z, nz = Y),
    dx = 8 / X m; rest level at 2/3 of the height, surface rest + amplitude *
    H * cos(pi x / Lx) cos(pi y / Ly) with H the rest depth; `spheres` solid
    spheres of radius 2-4 cells whose tops stay below half the rest depth (so
    the surface stays far above every column floor); zero velocity. All
    arithmetic of the heights in f32."""
    nx, ny, nz = X, Z, Y
    f32 = np.float32
    dx = f32(8.0) / f32(nx)
    rest = f32(nz) * dx * f32(2.0) / f32(3.0)
    i = (np.arange(nx, dtype=f32) + f32(0.5)) / f32(nx)
    j = (np.arange(ny, dtype=f32) + f32(0.5)) / f32(ny)
    mode = np.outer(np.cos(f32(np.pi) * j), np.cos(f32(np.pi) * i)).astype(f32)
    eta = (rest + f32(amplitude) * rest * mode).astype(f32).ravel()
    rng = np.random.default_rng(seed)
    solid = np.zeros((nz, ny, nx), np.uint8)
    kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    top_max = float(rest) / 2.0
    for _ in range(spheres):
        r = rng.uniform(2.0, 4.0)
        cx, cy = rng.uniform(r + 1, nx - r - 1), rng.uniform(r + 1, ny - r - 1)
        cz = rng.uniform(r + 1, max(r + 1.5, top_max / float(dx) - r - 1))
        solid[(ii + 0.5 - cx) ** 2 + (jj + 0.5 - cy) ** 2 + (kk + 0.5 - cz) ** 2 <= r * r] = 1
    nu, nv, nw = Domain3(nx, ny, nz, float(dx)).faces
    z = lambda n: np.zeros(n, np.float32)
    return Workload(f"slosh_{X}x{Y}x{Z}", Domain3(nx, ny, nz, float(dx)), eta, solid.ravel(), z(nu), z(nv), z(nw),
                    homogeneous_ghost=False, closed_lid=False, scale=float(np.float32(-1025.0 / 1e-3)),
                    k1=float(np.float32(1e-3 / 1025.0)), rho=1025.0, g_eff=9.81, rest=float(rest),
                    options=dict(preconditioner=_lib.PRECOND_MULTIGRID, smoother=_lib.SMOOTH_RBGS,
                                 max_levels=-1, pre_sweeps=2, post_sweeps=2, coarse_sweeps=8,
                                 coarse_operator=_lib.COARSE_GALERKIN),
                    seed=seed)
