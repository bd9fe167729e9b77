"""The library's contracts on the card: nothing is reserved after
fsp_create (the reference's I-06, every buffer reserved before seal(),
delta3d.rs:225-241), and a learned-phase graph stays valid when the shared
network's buffers move to another shape."""
import gc

import numpy as np
import pytest

from fluidisim_amd import _lib
from fluidisim_amd import workloads as W
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def free_bytes():
    torch.cuda.synchronize()
    return torch.cuda.mem_get_info()[0]


def test_projection_and_surface_step_reserve_no_device_memory():
    """Device memory in use (hipMemGetInfo) is the same before and after the
    first projection and the first surface step of a fresh solver. A solver
    of the same shape runs first, so that the runtime's own first-launch
    work (code objects) is not counted."""
    wl = W.slosh(64, 32, 64, seed=2)
    warm = W.build_volume(wl)
    warm.project_mobile3(wl.scale, wl.k1, 2000, warm=False)
    warm.step_surface_mobile(1000, 4000)
    warm.close()
    vol = W.build_volume(wl)
    rng = np.random.default_rng(4)
    vol.set_velocity(*(rng.uniform(-1e-2, 1e-2, n).astype(np.float32) for n in wl.domain.faces))
    vol._upload_geometry()
    # the count is device-wide: solvers of earlier tests that the collector
    # has yet to destroy must not give their memory back inside the window
    gc.collect()
    gc.disable()
    try:
        f0 = free_bytes()
        rep = vol.project_mobile3(wl.scale, wl.k1, 2000, warm=False)
        f1 = free_bytes()
        vol.step_surface_mobile(1000, 4000)
        f2 = free_bytes()
    finally:
        gc.enable()
    print(f"free bytes {f0} -> {f1} (projection, {rep.iterations} it) -> {f2} (surface step)", flush=True)
    assert rep.iterations > 0
    assert f0 == f1 == f2
    vol.close()


def test_learned_graph_follows_the_network_to_another_shape():
    """A projection with the network attached, then a forward pass of the
    same network at another shape (its buffers are reallocated), then the
    projection again: the learned phase is captured anew and the result is
    the f64 oracle's."""
    from fluidisim_amd.learned import WEIGHTS, Network
    net = Network()
    net.load(WEIGHTS)
    wl = W.tank(64, 32, 64, seed=21)
    vol = W.build_volume(wl, preconditioner=_lib.PRECOND_LEARNED, learned_budget=200, learned_vcycle=1)
    net.attach(vol)
    first = vol.project_mobile3(wl.scale, wl.k1, 2000, warm=False)
    p_first = vol.pressure().cpu().numpy().copy()
    # another shape: 128 x 32 x 64 cells of random binary16 features
    shape = (128, 32, 64)
    x = torch.randn(4 * shape[0] * shape[1] * shape[2], device="cuda").to(torch.float16).view(torch.int16)
    y = net.forward(x, shape)
    assert bool(torch.isfinite(y).all())
    second = vol.project_mobile3(wl.scale, wl.k1, 2000, warm=False)
    p_second = vol.pressure().cpu().numpy()
    d = wl.domain
    prob = po.Problem(d.nx, d.ny, d.nz, d.dx, wl.eta, rho=wl.rho, g_eff=wl.g_eff, rest=wl.rest,
                      homogeneous_ghost=wl.homogeneous_ghost, closed_lid=wl.closed_lid, solid=wl.solid)
    o = wl.options
    *_, p64 = po.project(prob, wl.us, wl.vs, wl.ws, wl.scale, wl.k1, 2000,
                         po.options(multigrid=1, smoother=o["smoother"], pre=o["pre_sweeps"], post=o["post_sweeps"],
                                    coarse=o["coarse_sweeps"], coarse_op=o.get("coarse_operator", 0)), precision=64)
    kind, _ = po.classify(prob)
    wet = kind == 1
    err = np.linalg.norm(p_second[wet] - p64[wet]) / np.linalg.norm(p64[wet])
    print(first, second, f"pressure rel L2 {err:.2e}", flush=True)
    assert not second.degraded and second.residual <= 1e-6
    assert second.iterations == first.iterations
    assert np.array_equal(p_first.view(np.uint32), p_second.view(np.uint32))
    assert err <= 1e-4
    vol.close()
    net.close()


def test_batch_refuses_solvers_that_share_a_network():
    from fluidisim_amd.learned import WEIGHTS, Network
    from fluidisim_amd.volume import ProjectionError, project_batch
    net = Network()
    net.load(WEIGHTS)
    wls = [W.tank(64, 32, 64, seed=s) for s in (31, 32)]
    vols = [W.build_volume(wl, preconditioner=_lib.PRECOND_LEARNED, learned_vcycle=1) for wl in wls]
    for vol in vols:
        net.attach(vol)
    with pytest.raises(ProjectionError):
        project_batch(vols, [wl.scale for wl in wls], [wl.k1 for wl in wls], 2000)
    for vol in vols:
        vol.close()
    net.close()
