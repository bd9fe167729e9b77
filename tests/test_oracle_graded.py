"""The graded column in the oracle (CPU): delta3d_graded.rs, the pressure at
node layers only, linear between nodes, the projection on P^T A P.

Pins, from the reference's own tests and from the mathematics:
- P^T is the exact transpose of P (tests_delta3d.rs:1933-1939), P injects
  at the nodes and keeps constants;
- P^T A P is symmetric and positive (tests_delta3d.rs:1920-1932);
- the projection is the Galerkin solution: P x with x from a dense f64
  solve of (P^T A P) x = P^T b, P built here as the reference's frequency
  test builds it (tests_delta3d.rs:2005-2015) and A from the oracle's
  operator on unit vectors;
- with every layer a node it is the fine scheme (tests_delta3d.rs:2095-2109);
- the refusals (tests_delta3d.rs:1904-1918, 2111-2130).
"""
import numpy as np
import pytest

from oracle import pyoracle as po

RHO, G = 1025.0, 9.81


def mobile_case(nx=8, ny=4, nz=16, dx=0.25, h=3.6, seed=2, solid=None):
    rng = np.random.default_rng(seed)
    i = np.arange(nx)
    eta = np.tile(h + 0.02 * np.sin(2 * np.pi * (i + 0.5) / nx), ny).astype(np.float32)
    prob = po.Problem(nx, ny, nz, dx, eta, rho=RHO, g_eff=G, rest=h, homogeneous_ghost=False, solid=solid)
    nu, nv, nw = prob.face_counts
    u = (0.1 * rng.standard_normal((nz, ny, nx + 1))).astype(np.float32)
    v = (0.1 * rng.standard_normal((nz, ny + 1, nx))).astype(np.float32)
    w = (0.1 * rng.standard_normal((nz + 1, ny, nx))).astype(np.float32)
    u[:, :, [0, nx]] = 0
    v[:, [0, ny], :] = 0
    w[[0, nz]] = 0
    return prob, (u.ravel(), v.ravel(), w.ravel())


NODES = [0, 3, 7, 10, 12, 13, 14, 15]
SCALE, K1 = np.float32(-RHO / 0.002), np.float32(0.002 / RHO)


def dense_p(nx, ny, nz, nodes):
    """P (cells x graded), as the reference's graded frequency test builds
    its vertical part: (1 - t, t) on each segment, 1 on the top node."""
    cols = nx * ny
    r = len(nodes)
    P1 = np.zeros((nz, r))
    for s in range(r - 1):
        lo, hi = nodes[s], nodes[s + 1]
        for k in range(lo, hi):
            t = (k - lo) / (hi - lo)
            P1[k, s] = 1 - t
            P1[k, s + 1] = t
    P1[nodes[-1], r - 1] = 1
    return np.kron(P1, np.eye(cols))


@pytest.mark.parametrize("precision,tol", [(64, 1e-13), (32, 1e-5)])
def test_restriction_is_the_exact_transpose_of_the_prolongation(precision, tol):
    prob, _ = mobile_case()
    rng = np.random.default_rng(5)
    n = prob.nx * prob.ny * len(NODES)
    x = rng.uniform(-1, 1, n)
    f = rng.uniform(-1, 1, prob.cells)
    if precision == 32:
        x, f = x.astype(np.float32).astype(np.float64), f.astype(np.float32).astype(np.float64)
    Px = po.graded_op(prob, NODES, po.GRADED_EXPAND, x, precision)
    Ptf = po.graded_op(prob, NODES, po.GRADED_RESTRICT, f, precision)
    lhs, rhs = Px @ f, x @ Ptf
    assert abs(lhs - rhs) < tol * np.sqrt((Px @ Px) * (f @ f))
    # against the dense P
    P = dense_p(prob.nx, prob.ny, prob.nz, NODES)
    assert np.abs(Px - P @ x).max() < 10 * tol
    assert np.abs(Ptf - P.T @ f).max() < 100 * tol


def test_prolongation_injects_at_nodes_and_keeps_constants():
    prob, _ = mobile_case()
    cols, n = prob.nx * prob.ny, prob.nx * prob.ny * len(NODES)
    x = np.random.default_rng(7).uniform(-1, 1, n)
    fine = po.graded_op(prob, NODES, po.GRADED_EXPAND, x).reshape(prob.nz, cols)
    for s, k in enumerate(NODES):
        assert np.array_equal(fine[k], x[s * cols:(s + 1) * cols])
    ones = po.graded_op(prob, NODES, po.GRADED_EXPAND, np.ones(n))
    assert np.abs(ones - 1).max() < 1e-15
    # the weights are P^T of ones
    w = po.graded_op(prob, NODES, po.GRADED_WEIGHTS)
    assert np.allclose(w, dense_p(1, 1, prob.nz, NODES).sum(axis=0), rtol=0, atol=1e-14)


@pytest.mark.parametrize("precision,tol", [(64, 1e-12), (32, 1e-5)])
def test_graded_operator_is_symmetric_and_positive(precision, tol):
    prob, _ = mobile_case()
    rng = np.random.default_rng(11)
    n = prob.nx * prob.ny * len(NODES)
    x, y = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    if precision == 32:
        x, y = x.astype(np.float32).astype(np.float64), y.astype(np.float32).astype(np.float64)
    ax = po.graded_op(prob, NODES, po.GRADED_APPLY, x, precision)
    ay = po.graded_op(prob, NODES, po.GRADED_APPLY, y, precision)
    assert abs(x @ ay - y @ ax) < tol * np.sqrt((x @ x) * (ay @ ay))
    assert x @ ax > 0 and y @ ay > 0


def test_graded_projection_is_the_dense_galerkin_solution():
    prob, faces = mobile_case()
    cells = prob.cells
    A = np.zeros((cells, cells))
    for j in range(cells):
        e = np.zeros(cells)
        e[j] = 1.0
        A[:, j] = po.apply64(prob, e)
    P = dense_p(prob.nx, prob.ny, prob.nz, NODES)
    Ar = P.T @ A @ P
    b = po.graded_op(prob, NODES, po.GRADED_RHS, np.concatenate(faces).astype(np.float64), scale=float(SCALE))
    # dry nodes (zero rows) carry nothing: solve on the others
    live = np.abs(Ar).sum(axis=1) > 0
    x = np.zeros(len(b))
    x[live] = np.linalg.solve(Ar[np.ix_(live, live)], b[live])
    want = P @ x
    rep, _, _, _, _, p64 = po.project_graded(prob, *faces, SCALE, K1, NODES, opts=po.options(rel_tol=1e-9),
                                             precision=64)
    assert not rep["degraded"], rep
    err = np.linalg.norm(p64 - want) / np.linalg.norm(want)
    assert err < 1e-7, err
    # the fine solution of the same problem is a different pressure
    rep_f, _, _, _, _, pf = po.project(prob, *faces, SCALE, K1, opts=po.options(multigrid=0, rel_tol=1e-9),
                                       precision=64)
    assert np.linalg.norm(pf - want) / np.linalg.norm(want) > 1e-4


def test_every_layer_as_a_node_is_the_fine_scheme():
    prob, faces = mobile_case(seed=4)
    every = list(range(prob.nz))
    opts = po.options(multigrid=0, rel_tol=1e-9)
    rep_g, _, ug, vg, wg, pg = po.project_graded(prob, *faces, SCALE, K1, every, opts=opts, precision=64)
    rep_f, _, uf, vf, wf, pf = po.project(prob, *faces, SCALE, K1, opts=opts, precision=64)
    assert not rep_g["degraded"] and not rep_f["degraded"]
    assert np.linalg.norm(pg - pf) / np.linalg.norm(pf) < 1e-8
    for a, b in ((ug, uf), (vg, vf), (wg, wf)):
        assert np.abs(a - b).max() < 1e-6 * max(np.abs(b).max(), 1e-30)
    # the reference's mobile-step nodes on its own case: f32 projection accepted
    nx, ny, nz, dx = 32, 16, 24, 0.25
    eta = np.tile(4 + 0.01 * np.sin(2 * np.pi * (np.arange(nx) + 0.5) * dx / 8), ny).astype(np.float32)
    big = po.Problem(nx, ny, nz, dx, eta, rho=RHO, g_eff=G, rest=4.0, homogeneous_ghost=False)
    rng = np.random.default_rng(9)
    zs = [(0.05 * rng.standard_normal(n)).astype(np.float32) for n in big.face_counts]
    nodes = [0, 5, 9] + list(range(12, 24))
    rep, *_ = po.project_graded(big, *zs, SCALE, K1, nodes)
    assert not rep["degraded"] and rep["divergence_plain"] <= 1e-5, rep


def test_graded_refusals():
    prob, faces = mobile_case()
    for bad in ([1, 5, 15], [0, 5, 14], [0, 5, 5, 15]):
        with pytest.raises(RuntimeError, match="status 3"):
            po.graded_op(prob, bad, po.GRADED_WEIGHTS)
    solid = np.zeros(prob.cells, np.uint8)
    solid[0] = 1
    cut, _ = mobile_case(solid=solid)
    with pytest.raises(RuntimeError, match="status 1"):
        po.graded_op(cut, NODES, po.GRADED_WEIGHTS)
    # the lowest cubic node (12) and the layer above must be wet everywhere:
    # a surface at 3.3 m leaves layer 13 (centre 3.375 m) dry
    low, lfaces = mobile_case(h=3.3)
    with pytest.raises(RuntimeError, match="status 1"):
        po.project_graded(low, *lfaces, SCALE, K1, NODES)
