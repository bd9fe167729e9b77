"""Writes the golden fixtures of tests/golden/ (small .npz files).

Each fixture holds its inputs (geometry and predicted faces, so it is
self-contained), the generator settings and seed that produced them, and the
outputs of the f64 oracle (pressure, corrected faces, iteration count) or of
a closed form. Run from the repository root:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.fft

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fluidisim_amd import workloads as W  # noqa: E402
from oracle import pyoracle as po  # noqa: E402


def problem(wl):
    d = wl.domain
    return po.Problem(d.nx, d.ny, d.nz, d.dx, wl.eta, rho=wl.rho, g_eff=wl.g_eff, rest=wl.rest,
                      homogeneous_ghost=wl.homogeneous_ghost, closed_lid=wl.closed_lid, solid=wl.solid)


def opts(wl):
    o = wl.options
    return po.options(multigrid=1, smoother=o["smoother"], max_levels=o["max_levels"], pre=o["pre_sweeps"],
                      post=o["post_sweeps"], coarse=o["coarse_sweeps"], coarse_op=o.get("coarse_operator", 0))


def save_workload_case(name, wl, generator):
    rep, p32, u, v, w, p64 = po.project(problem(wl), wl.us, wl.vs, wl.ws, wl.scale, wl.k1, 2000, opts(wl),
                                        precision=64)
    d = wl.domain
    np.savez_compressed(
        os.path.join(HERE, f"{name}.npz"),
        generator=np.array(generator), seed=wl.seed, dims=np.array([d.nx, d.ny, d.nz]), dx=d.dx,
        eta=wl.eta, solid=wl.solid if wl.solid is not None else np.zeros(0, np.uint8),
        us=wl.us, vs=wl.vs, ws=wl.ws, scale=wl.scale, k1=wl.k1, rho=wl.rho, g_eff=wl.g_eff, rest=wl.rest,
        homogeneous=int(wl.homogeneous_ghost), closed_lid=int(wl.closed_lid),
        p=p64, u=u, v=v, w=w, iterations=rep["iterations"])
    print(name, rep["iterations"], rep["residual"])


def save_s385(nx):
    prob, scale, k1 = po.s385_problem(nx)
    z = [np.zeros(n, np.float32) for n in prob.face_counts]
    rep, p32, u, v, w = po.project(prob, *z, scale, k1, 20000, po.options(multigrid=1))
    np.savez_compressed(os.path.join(HERE, f"s385_{nx}.npz"), dims=np.array([prob.nx, prob.ny, prob.nz]),
                        dx=prob.dx, eta=prob.eta, scale=scale, k1=k1, p=p32, u=u, v=v, w=w,
                        iterations=rep["iterations"], published_iterations=9)
    print("s385", nx, rep["iterations"])


def save_dct_box():
    """All-Neumann box 12x10x8, dx = 0.5: the pressure by DCT-II (closed form)."""
    nx, ny, nz, dx = 12, 10, 8, 0.5
    rng = np.random.default_rng(2024)
    b = rng.standard_normal((nz, ny, nx))
    b -= b.mean()
    lam = np.zeros((nz, ny, nx))
    for axis, n in ((2, nx), (1, ny), (0, nz)):
        shape = [1, 1, 1]
        shape[axis] = n
        lam = lam + ((2 - 2 * np.cos(np.pi * np.arange(n) / n)) / dx ** 2).reshape(shape)
    lam[0, 0, 0] = 1.0
    ph = scipy.fft.dctn(b, type=2, norm="ortho") / lam
    ph[0, 0, 0] = 0.0
    p = scipy.fft.idctn(ph, type=2, norm="ortho")
    np.savez_compressed(os.path.join(HERE, "dct_box.npz"), dims=np.array([nx, ny, nz]), dx=dx, b=b.ravel(),
                        p=p.ravel())
    print("dct_box")


if __name__ == "__main__":
    save_workload_case("closed_box_12", W.closed_box(12, seed=1), "closed_box(12, seed=1)")
    save_workload_case("tank_16", W.tank(16, 16, 16, seed=2), "tank(16, 16, 16, seed=2)")
    save_workload_case("tank_24x16x20", W.tank(24, 16, 20, seed=3), "tank(24, 16, 20, seed=3)")
    save_s385(32)
    save_dct_box()
