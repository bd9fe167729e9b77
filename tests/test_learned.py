"""The learned preconditioner: the CPU reference network (oracle/netref.py)
pinned by finite differences and a direct convolution, the card's MFMA
forward and backward passes against it, training, and flexible CG with the
network inside the projection (MGPCG fallback)."""
import numpy as np
import pytest

from fluidisim_amd import _lib
from fluidisim_amd import workloads as W
from oracle import netref as nr
from oracle import pyoracle as po


# ---------------------------------------------------------------- CPU: the reference network
def test_parameter_count_matches_the_library():
    assert nr.param_count() == 80729 == _lib.lib().fsp_net_param_count()


def test_reference_convolution_is_the_direct_sum():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((3, 4, 5, 2))
    W_ = rng.standard_normal((3, 27, 2))
    b = rng.standard_normal(3)
    got = nr.conv(x, W_, b)
    want = np.zeros((3, 4, 5, 3))
    for k in range(3):
        for j in range(4):
            for i in range(5):
                s = b.copy()
                for t, (dz, dy, dx) in enumerate(nr.TAPS):
                    kk, jj, ii = k + dz, j + dy, i + dx
                    if 0 <= kk < 3 and 0 <= jj < 4 and 0 <= ii < 5:
                        s += W_[:, t, :] @ x[kk, jj, ii]
                want[k, j, i] = s
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    th = nr.init(2)
    th += rng.standard_normal(th.size) * 0.01  # non-zero biases
    x = rng.standard_normal((4, 8, 8, 4))
    wet = (rng.random((4, 8, 8)) > 0.2).astype(float)
    t = rng.standard_normal((4, 8, 8))
    _, g, _ = nr.loss_and_grad(th, x, t, wet)
    eps = 1e-6
    for i in rng.choice(th.size, 40, replace=False):
        tp, tm = th.copy(), th.copy()
        tp[i] += eps
        tm[i] -= eps
        fd = (nr.loss_and_grad(tp, x, t, wet)[0] - nr.loss_and_grad(tm, x, t, wet)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-4 * max(abs(fd), abs(g[i]), 1e-3), (i, fd, g[i])


def test_reference_identity_network_passes_its_input_through():
    """Centre taps with identity channel maps pass channel 0 through."""
    th = np.zeros(nr.param_count())
    params = nr.unpack(th)
    for (W_, b), (name, level, ci, co, relu, skip) in zip(params, nr.LAYERS):
        for c in range(min(ci, co)):
            W_[c, 13, c] = 1.0
    x = np.abs(np.random.default_rng(3).standard_normal((4, 8, 8, 4)))
    y, _ = nr.forward(th, x)
    # channel 0 reaches d1a through the e1b skip; the coarse path adds a
    # non-negative term (ReLU outputs), so y >= x0, with equality where the
    # pooled-and-upsampled path is zero
    assert np.all(y >= x[..., 0] - 1e-12)
    th_noskip_path = np.zeros_like(th)
    for (W_, b), (name, level, ci, co, relu, skip) in zip(nr.unpack(th_noskip_path), nr.LAYERS):
        if level == 0:
            for c in range(min(ci, co)):
                W_[c, 13, c] = 1.0
    y0, _ = nr.forward(th_noskip_path, x)
    assert np.allclose(y0, x[..., 0], rtol=0, atol=1e-12)


# ---------------------------------------------------------------- GPU
torch = None


def _t():
    global torch
    if torch is None:
        import torch as t
        torch = t
    return torch


def f16_tensor(a):
    t = _t()
    return t.from_numpy(np.ascontiguousarray(a, np.float16).view(np.int16).ravel()).cuda()


def random_features(rng, nz, ny, nx):
    x = np.zeros((nz, ny, nx, 4), np.float32)
    wet = rng.random((nz, ny, nx)) > 0.15
    x[..., 0] = rng.standard_normal((nz, ny, nx)) * wet
    x[..., 1] = wet
    x[..., 2] = wet & (rng.random((nz, ny, nx)) > 0.9)
    x[..., 3] = ~wet & (rng.random((nz, ny, nx)) > 0.5)
    return x.astype(np.float16), wet


@pytest.mark.gpu
def test_mfma_operand_layout():
    t = _t()
    rng = np.random.default_rng(4)
    A = rng.standard_normal((16, 16)).astype(np.float16)
    B = rng.standard_normal((16, 16)).astype(np.float16)
    D = t.empty(256, dtype=t.float32, device="cuda")
    a, b = f16_tensor(A), f16_tensor(B)  # kept alive across the call
    t.cuda.synchronize()
    st = _lib.lib().fsp_net_mfma_probe(a.data_ptr(), b.data_ptr(), D.data_ptr())
    assert st == 0
    want = A.astype(np.float64) @ B.astype(np.float64)
    got = D.cpu().numpy().reshape(16, 16)
    assert np.abs(got - want).max() <= 1e-4 * np.abs(want).max()


@pytest.fixture(scope="module")
def net():
    from fluidisim_amd.learned import Network
    n = Network()
    yield n
    n.close()


SHAPE = (64, 8, 8)  # nx, ny, nz


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [SHAPE, (64, 12, 12)], ids=["whole-bricks", "partial-bricks"])
def test_forward_matches_the_reference(net, shape):
    """The card's forward pass against the float64 reference. (64, 12, 12)
    leaves the last bricks of every level partly outside the grid: 12, 6 and
    3 rows and planes at levels 0, 1 and 2 against bricks of 8 (4 for the
    upsampling kernels' planes)."""
    t = _t()
    rng = np.random.default_rng(5)
    th = nr.init(6).astype(np.float32)
    th[np.arange(th.size) % 7 == 0] += 0.01
    net.set_params(th)
    nx, ny, nz = shape
    x16, wet = random_features(rng, nz, ny, nx)
    y = net.forward(f16_tensor(x16), shape).cpu().numpy().reshape(nz, ny, nx)
    yr, _ = nr.forward(th.astype(np.float64), x16.astype(np.float64), round16=True)
    err = np.linalg.norm(y - yr) / np.linalg.norm(yr)
    print(f"forward rel L2 {err:.2e}")
    assert err <= 1e-2


@pytest.mark.gpu
def test_gradients_match_the_reference(net):
    t = _t()
    rng = np.random.default_rng(7)
    th = nr.init(8).astype(np.float32)
    net.set_params(th)
    nx, ny, nz = SHAPE
    x16, wet = random_features(rng, nz, ny, nx)
    target = (rng.standard_normal((nz, ny, nx)) * wet).astype(np.float32)
    loss, g = net.loss_grad(f16_tensor(x16), t.from_numpy(target.ravel()).cuda(),
                            t.from_numpy(wet.ravel().astype(np.uint8)).cuda(), SHAPE)
    lr, gr, _ = nr.loss_and_grad(th.astype(np.float64), x16.astype(np.float64), target.astype(np.float64),
                                 wet.astype(np.float64), round16=True)
    print(f"loss {loss:.6e} vs {lr:.6e}")
    assert abs(loss - lr) <= 1e-2 * lr
    cos = float(g @ gr / (np.linalg.norm(g) * np.linalg.norm(gr)))
    rel = float(np.linalg.norm(g - gr) / np.linalg.norm(gr))
    print(f"gradient cosine {cos:.6f}, rel L2 {rel:.3e}")
    assert cos >= 0.999 and rel <= 5e-2
    # every layer's gradient, not just the largest ones
    off = 0
    for name, level, ci, co, relu, skip in nr.LAYERS:
        n = co * 27 * ci + co
        a, b = g[off:off + n], gr[off:off + n]
        assert np.linalg.norm(a - b) <= 0.1 * np.linalg.norm(b) + 1e-6, name
        off += n


@pytest.mark.gpu
def test_training_steps_reduce_the_loss(net):
    t = _t()
    rng = np.random.default_rng(9)
    net.set_params(nr.init(10).astype(np.float32))
    nx, ny, nz = SHAPE
    x16, wet = random_features(rng, nz, ny, nx)
    target = (x16[..., 0].astype(np.float32) * 0.5).astype(np.float32)
    xt = f16_tensor(x16)
    tt = t.from_numpy(target.ravel()).cuda()
    wt = t.from_numpy(wet.ravel().astype(np.uint8)).cuda()
    losses = [net.train_step(xt, tt, wt, SHAPE, 3e-3)[0] for _ in range(60)]
    print(losses[0], losses[-1])
    assert losses[-1] < 0.3 * losses[0]


@pytest.mark.gpu
def test_features_follow_the_cell_classes(net):
    """The network input of a residual: (r sigma, wet, ghost row, solid) per
    cell in binary16 with sigma = sqrt(n_wet / |r|^2), checked against the
    oracle's classification; a residual whose address is not 16-byte aligned
    (one cell at a time) gives the same bits as an aligned one (four cells
    per thread)."""
    t = _t()
    wl = W.tank(64, 32, 64, seed=21)
    d = wl.domain
    prob = po.Problem(d.nx, d.ny, d.nz, d.dx, wl.eta, rho=wl.rho, g_eff=wl.g_eff, rest=wl.rest,
                      homogeneous_ghost=wl.homogeneous_ghost, closed_lid=wl.closed_lid, solid=wl.solid)
    kind, ghost = po.classify(prob)
    wet = kind == 1
    vol = W.build_volume(wl)
    r = np.random.default_rng(22).standard_normal(d.cells).astype(np.float32) * wet
    ra = t.tensor(r, device="cuda")
    buf = t.zeros(d.cells + 1, dtype=t.float32, device="cuda")
    buf[1:] = ra
    assert buf[1:].data_ptr() % 16 != 0
    xa, sa = net.features(vol, ra)
    xb, sb = net.features(vol, buf[1:])
    assert sa == sb and t.equal(xa, xb)
    sigma = np.float32(np.sqrt(wet.sum() / np.sum(r.astype(np.float64) ** 2)))
    assert abs(sa - sigma) <= 1e-6 * sigma
    x = xa.view(t.float16).view(-1, 4).cpu().numpy()
    assert np.array_equal(x[:, 1], wet.astype(np.float16))
    # r sigma rounded once to binary16 (the card's mixed fused multiply-add
    # rounds the exact product; the float64 product of two floats is exact)
    want = np.where(wet, (r.astype(np.float64) * np.float64(sa)).astype(np.float16), np.float16(0))
    bad = np.flatnonzero(x[:, 0] != want)
    for c in bad[:8]:
        print(f"cell {c}: r {r[c]!r} got {x[c, 0]!r} want {want[c]!r}")
    assert bad.size == 0
    assert np.array_equal(x[:, 2], ((ghost != 0) & wet).astype(np.float16))
    solid = np.zeros(d.cells, np.uint8) if wl.solid is None else np.asarray(wl.solid).ravel()
    assert np.array_equal(x[:, 3], (solid != 0).astype(np.float16))
    vol.close()


@pytest.mark.gpu
def test_learned_fcg_with_an_untrained_network_still_converges(net):
    """A network that is no preconditioner: the MGPCG fallback brings the
    projection to 1e-6 with the reference's pressure."""
    from fluidisim_amd.learned import init_params
    net.set_params(init_params(11))
    wl = W.tank(64, 32, 64, seed=12)
    vol = W.build_volume(wl, preconditioner=_lib.PRECOND_LEARNED, learned_budget=10)
    net.attach(vol)
    rep = vol.project_mobile3(wl.scale, wl.k1, 2000, warm=False)
    d = wl.domain
    prob = po.Problem(d.nx, d.ny, d.nz, d.dx, wl.eta, rho=wl.rho, g_eff=wl.g_eff, rest=wl.rest,
                      homogeneous_ghost=wl.homogeneous_ghost, closed_lid=wl.closed_lid, solid=wl.solid)
    o = wl.options
    oref, *_, p64 = po.project(prob, wl.us, wl.vs, wl.ws, wl.scale, wl.k1, 2000,
                               po.options(multigrid=1, smoother=o["smoother"], pre=o["pre_sweeps"],
                                          post=o["post_sweeps"], coarse=o["coarse_sweeps"],
                                          coarse_op=o.get("coarse_operator", 0)), precision=64)
    print(rep, oref)
    assert not rep.degraded and rep.residual <= 1e-6 and rep.divergence_plain <= 1e-5
    kind, _ = po.classify(prob)
    wet = kind == 1
    p = vol.pressure().cpu().numpy()[wet].astype(np.float64)
    err = np.linalg.norm(p - p64[wet]) / np.linalg.norm(p64[wet])
    print(f"pressure rel L2 {err:.2e}")
    assert err <= 1e-4
    vol.close()


@pytest.mark.gpu
def test_trained_network_beats_the_jacobi_preconditioner(net):
    """What the trained network gains: flexible CG with it needs fewer
    iterations to 1e-6 than CG with the reference's Jacobi preconditioner."""
    from fluidisim_amd.learned import init_params, train
    net.set_params(init_params(13))
    hist = train(net, size=(64, 32, 64), samples=12, epochs=25, lr=2e-3, seed0=2000, log=lambda *_: None)
    print("loss history", hist[0], hist[-1])
    assert hist[-1] < 0.5 * hist[0]
    wl = W.tank(64, 32, 64, seed=3000)  # not a training sample
    vj = W.build_volume(wl, preconditioner=_lib.PRECOND_JACOBI)
    rj = vj.project_mobile3(wl.scale, wl.k1, 5000, warm=False)
    vl = W.build_volume(wl, preconditioner=_lib.PRECOND_LEARNED, learned_budget=2000)
    net.attach(vl)
    rl = vl.project_mobile3(wl.scale, wl.k1, 5000, warm=False)
    print("jacobi", rj, "learned", rl)
    assert not rl.degraded and rl.residual <= 1e-6
    assert rl.iterations < rj.iterations
    vj.close()
    vl.close()


@pytest.mark.gpu
def test_network_then_vcycle_converges_like_mgpcg(net):
    """learned_vcycle = 1: the trained network's output is the initial guess
    of one V-cycle inside flexible CG. The projection reaches 1e-6 with the
    reference's pressure, in no more iterations than the plain MGPCG plus
    one (the V-cycle alone already gives the MGPCG rate)."""
    from fluidisim_amd.learned import WEIGHTS
    net.load(WEIGHTS)
    wl = W.tank(64, 32, 64, seed=14)
    vm = W.build_volume(wl)
    rm = vm.project_mobile3(wl.scale, wl.k1, 2000, warm=False)
    vl = W.build_volume(wl, preconditioner=_lib.PRECOND_LEARNED, learned_budget=200, learned_vcycle=1)
    net.attach(vl)
    rl = vl.project_mobile3(wl.scale, wl.k1, 2000, warm=False)
    print("mgpcg", rm, "network + V-cycle", rl)
    assert not rl.degraded and rl.residual <= 1e-6 and rl.divergence_plain <= 1e-5
    assert rl.iterations <= rm.iterations + 1
    d = wl.domain
    prob = po.Problem(d.nx, d.ny, d.nz, d.dx, wl.eta, rho=wl.rho, g_eff=wl.g_eff, rest=wl.rest,
                      homogeneous_ghost=wl.homogeneous_ghost, closed_lid=wl.closed_lid, solid=wl.solid)
    o = wl.options
    *_, p64 = po.project(prob, wl.us, wl.vs, wl.ws, wl.scale, wl.k1, 2000,
                         po.options(multigrid=1, smoother=o["smoother"], pre=o["pre_sweeps"],
                                    post=o["post_sweeps"], coarse=o["coarse_sweeps"],
                                    coarse_op=o.get("coarse_operator", 0)), precision=64)
    kind, _ = po.classify(prob)
    wet = kind == 1
    p = vl.pressure().cpu().numpy()[wet].astype(np.float64)
    err = np.linalg.norm(p - p64[wet]) / np.linalg.norm(p64[wet])
    print(f"pressure rel L2 {err:.2e}")
    assert err <= 1e-4
    vm.close()
    vl.close()
