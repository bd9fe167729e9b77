"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py): the seeded generator still produces their
inputs bit for bit, the oracle still produces their outputs, and the DCT
closed form satisfies the oracle's operator."""
import os

import numpy as np
import pytest

from fluidisim_amd import workloads as W
from oracle import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

CASES = {
    "closed_box_12": lambda: W.closed_box(12, seed=1),
    "tank_16": lambda: W.tank(16, 16, 16, seed=2),
    "tank_24x16x20": lambda: W.tank(24, 16, 20, seed=3),
}


def load(name):
    return np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)


def problem_from(g):
    nx, ny, nz = (int(v) for v in g["dims"])
    solid = g["solid"] if g["solid"].size else None
    return po.Problem(nx, ny, nz, float(g["dx"]), g["eta"], rho=float(g["rho"]), g_eff=float(g["g_eff"]),
                      rest=float(g["rest"]), homogeneous_ghost=bool(g["homogeneous"]),
                      closed_lid=bool(g["closed_lid"]), solid=solid)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("name", list(CASES))
def test_generator_reproduces_the_fixture_inputs(name):
    g = load(name)
    wl = CASES[name]()
    assert np.array_equal(bits(wl.eta), bits(g["eta"]))
    for a, key in ((wl.us, "us"), (wl.vs, "vs"), (wl.ws, "ws")):
        assert np.array_equal(bits(a), bits(g[key]))
    if g["solid"].size:
        assert np.array_equal(wl.solid, g["solid"])


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_reproduces_the_fixture_outputs(name):
    g = load(name)
    wl = CASES[name]()
    o = wl.options
    opts = po.options(multigrid=1, smoother=o["smoother"], max_levels=o["max_levels"], pre=o["pre_sweeps"],
                      post=o["post_sweeps"], coarse=o["coarse_sweeps"], coarse_op=o.get("coarse_operator", 0))
    rep, p32, u, v, w, p64 = po.project(problem_from(g), g["us"], g["vs"], g["ws"], float(g["scale"]),
                                        float(g["k1"]), 2000, opts, precision=64)
    assert rep["iterations"] == int(g["iterations"])
    assert np.array_equal(p64, g["p"])
    assert np.array_equal(bits(u), bits(g["u"]))


def test_s385_fixture_has_the_published_iterations():
    g = load("s385_32")
    assert int(g["iterations"]) == int(g["published_iterations"]) == 9
    prob, scale, k1 = po.s385_problem(32)
    assert np.array_equal(bits(prob.eta), bits(g["eta"]))


def test_dct_box_closed_form_satisfies_the_operator():
    """A p = b with the oracle's own operator on an all-Neumann box."""
    g = load("dct_box")
    nx, ny, nz = (int(v) for v in g["dims"])
    dx = float(g["dx"])
    eta = np.full(nx * ny, nz * dx, np.float32)
    prob = po.Problem(nx, ny, nz, dx, eta, homogeneous_ghost=True, closed_lid=True)
    ap = po.apply64(prob, g["p"])
    assert np.abs(ap - g["b"]).max() <= 1e-9 * np.abs(g["b"]).max()


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_card_matches_the_fixture_pressure(name):
    """The HIP projection on the fixture's own inputs against its f64 pressure."""
    import torch

    g = load(name)
    wl = CASES[name]()
    vol = W.build_volume(wl)
    rep = vol.project_mobile3(float(g["scale"]), float(g["k1"]), 2000, warm=False)
    torch.cuda.synchronize()
    assert not rep.degraded and rep.residual <= 1e-6
    assert abs(rep.iterations - int(g["iterations"])) <= 1
    kind, _ = po.classify(problem_from(g))
    wet = kind == 1
    p = vol.pressure().cpu().numpy().astype(np.float64)[wet]
    ref = g["p"][wet]
    if bool(g["closed_lid"]):
        p, ref = p - p.mean(), ref - ref.mean()
    err = np.linalg.norm(p - ref) / np.linalg.norm(ref)
    print(f"{name}: iterations {rep.iterations} vs {int(g['iterations'])}, pressure rel L2 {err:.3e}")
    assert err <= 1e-4
