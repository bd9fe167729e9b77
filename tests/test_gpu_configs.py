"""Every target configuration (closed box 64^3, tank 128^3, the 256x128x256
tank with 16-bit V-cycle storage and residual replacement, 48 patches, the
learned preconditioner) on the card at its stated size, against
the f64 oracle (the reference's MGPCG restated in C++, run in double), plus
the report's metrics against the oracle's on the same inputs and the
driver's acceptance paths forced from the host (the divergence-tightening
restart and the round-off floor stop of delta3d_mobile.rs:546-585).

Parity bar: pressure within 1e-4 relative L2 of the f64 oracle
on wet cells (mean-free on the pure-Neumann closed box), max |div u| (the
reference's normalised plain-row metric) within 1e-5, residual 1e-6.
"""
import ctypes as C
import sys
import time

import numpy as np
import pytest

from fluidisim_amd import _lib
from fluidisim_amd import workloads as W
from fluidisim_amd.volume import _ptr
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def problem(wl):
    d = wl.domain
    return po.Problem(d.nx, d.ny, d.nz, d.dx, wl.eta, rho=wl.rho, g_eff=wl.g_eff, rest=wl.rest,
                      homogeneous_ghost=wl.homogeneous_ghost, closed_lid=wl.closed_lid, solid=wl.solid)


def oracle_options(wl, **kw):
    o = wl.options
    return po.options(multigrid=1, smoother=o["smoother"], max_levels=o["max_levels"], pre=o["pre_sweeps"],
                      post=o["post_sweeps"], coarse=o["coarse_sweeps"], coarse_op=o.get("coarse_operator", 0), **kw)


def oracle64(wl, max_iters=2000, **kw):
    t0 = time.perf_counter()
    out = po.project(problem(wl), wl.us, wl.vs, wl.ws, wl.scale, wl.k1, max_iters, oracle_options(wl, **kw),
                     precision=64)
    say(f"[oracle f64] {wl.name}: {out[0]['iterations']} iterations in {time.perf_counter() - t0:.1f} s")
    return out


def rel_l2(wl, p_gpu, p_ref, mean_free=False):
    kind, _ = po.classify(problem(wl))
    wet = kind == 1
    a = np.asarray(p_gpu, np.float64)[wet]
    b = np.asarray(p_ref, np.float64)[wet]
    if mean_free:
        a, b = a - a.mean(), b - b.mean()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def host(t):
    return t.detach().cpu().numpy()


def check(wl, vol, rep, p64, what, mean_free=False, oref=None):
    err = rel_l2(wl, host(vol.pressure()), p64, mean_free)
    say(f"[{what}] {wl.name}: {rep.iterations} iterations (oracle {oref['iterations'] if oref else '-'}), "
        f"residual {rep.residual:.2e}, divergence {rep.divergence_plain:.2e}, pressure rel L2 vs f64 {err:.2e}")
    assert not rep.degraded and rep.residual <= 1e-6 and rep.divergence_plain <= 1e-5
    assert err <= 1e-4
    return err


def test_c0_closed_box_64_matches_the_f64_oracle():
    """64^3 all fluid in a closed box (pure Neumann, the gauge removes the
    mean), i.i.d. U[-1,1] faces, 3-level V-cycle with 2+2 red-black sweeps."""
    wl = W.closed_box(64, seed=1)
    vol = W.build_volume(wl)
    rep = vol.project_mobile3(wl.scale, wl.k1, 2000, warm=False)
    oref, _, _, _, _, p64 = oracle64(wl)
    check(wl, vol, rep, p64, "C0", mean_free=True, oref=oref)
    assert abs(rep.iterations - oref["iterations"]) <= 1
    vol.close()


def test_c1_tank_128_matches_the_f64_oracle():
    wl = W.tank(128, 128, 128, seed=1)
    vol = W.build_volume(wl)
    rep = vol.project_mobile3(wl.scale, wl.k1, 2000, warm=False)
    oref, _, _, _, _, p64 = oracle64(wl)
    check(wl, vol, rep, p64, "C1", oref=oref)
    assert abs(rep.iterations - oref["iterations"]) <= 1
    vol.close()


@pytest.fixture(scope="module")
def tank256():
    """The 256x128x256 tank (reference Domain3 {256, 256, 128}, 8.4 M cells)
    with the benchmark's settings, and the f64 oracle's projection of it."""
    wl = W.tank(256, 128, 256, seed=1)
    wl.options["coarse_sweeps"] = 2
    oref, _, _, _, _, p64 = oracle64(wl)
    return wl, oref, p64


def test_headline_plain_mgpcg_256x128x256_matches_the_f64_oracle(tank256):
    wl, oref, p64 = tank256
    vol = W.build_volume(wl)
    rep = vol.project_mobile3(wl.scale, wl.k1, 2000, warm=False)
    check(wl, vol, rep, p64, "plain MGPCG", oref=oref)
    assert abs(rep.iterations - oref["iterations"]) <= 1
    vol.close()


def test_c2_bf16_vcycle_with_residual_replacement_matches_the_f64_oracle(tank256):
    """C2: V-cycle vectors stored in bfloat16 with f32 arithmetic, f32 CG
    vectors, f64 dots and the true residual replacing the recurrence every
    50 iterations; the error against the f64 oracle is stated."""
    wl, oref, p64 = tank256
    vol = W.build_volume(wl, vcycle_storage=_lib.STORE_BF16, replace_every=50)
    rep = vol.project_mobile3(wl.scale, wl.k1, 2000, warm=False)
    err = check(wl, vol, rep, p64, "C2 bf16", oref=oref)
    assert rep.iterations <= oref["iterations"] + 2
    say(f"[C2] bf16 V-cycle storage, replacement every 50: {rep.iterations} iterations, "
        f"pressure error vs the f64 oracle {err:.2e}")
    vol.close()


def test_c4_learned_fcg_256x128x256_matches_the_f64_oracle(tank256):
    """C4: the committed network inside flexible CG (network output as the
    initial guess of one V-cycle), MGPCG fallback, at the full size."""
    from fluidisim_amd.learned import WEIGHTS, Network
    wl, oref, p64 = tank256
    net = Network()
    net.load(WEIGHTS)
    vol = W.build_volume(wl, preconditioner=_lib.PRECOND_LEARNED, learned_budget=200, learned_vcycle=1)
    net.attach(vol)
    rep = vol.project_mobile3(wl.scale, wl.k1, 2000, warm=False)
    check(wl, vol, rep, p64, "C4 learned FCG", oref=oref)
    vol.close()
    net.close()


def test_c3_48_patches_in_one_batch_each_match_the_f64_oracle():
    """C3: 48 independent patches (edges 32..128 per axis, weights
    .3/.25/.2/.15/.1), autotuned, projected together by fsp_project_batch;
    every patch against its own f64 oracle projection."""
    from fluidisim_amd.volume import project_batch
    wls = W.patches(48, seed=1)
    vols = [W.build_volume(wl) for wl in wls]
    for vol in vols:
        vol.autotune()
    reps = project_batch(vols, [wl.scale for wl in wls], [wl.k1 for wl in wls], 2000, concurrency=4)
    worst = 0.0
    for q, (wl, vol, rep) in enumerate(zip(wls, vols, reps)):
        oref, _, _, _, _, p64 = po.project(problem(wl), wl.us, wl.vs, wl.ws, wl.scale, wl.k1, 2000,
                                           oracle_options(wl), precision=64)
        err = rel_l2(wl, host(vol.pressure()), p64)
        worst = max(worst, err)
        say(f"[C3] patch {q} {wl.name}: {rep.iterations} iterations (oracle {oref['iterations']}), "
            f"divergence {rep.divergence_plain:.2e}, pressure rel L2 {err:.2e}")
        assert not rep.degraded and rep.residual <= 1e-6 and rep.divergence_plain <= 1e-5
        assert err <= 1e-4
    say(f"[C3] worst pressure rel L2 over {len(wls)} patches: {worst:.2e}")
    for vol in vols:
        vol.close()


# ---------------------------------------------------------------- the report against the oracle's
def test_divergence_metrics_equal_the_oracle_metric_of_the_same_faces():
    """The report's divergence and divergence_plain are the reference's
    metric (delta3d_mobile.rs:398-433) of the corrected faces the card
    returns, bit for bit: the oracle evaluates it on those faces."""
    for wl in (W.tank(64, 32, 48, seed=11), W.slosh(64, 32, 64, seed=2)):
        if wl.name.startswith("slosh"):
            # a moving field for the sloshing tank (at rest it projects to zero)
            d = wl.domain
            rng = np.random.default_rng(5)
            wl.us, wl.vs, wl.ws = (rng.uniform(-1e-2, 1e-2, n).astype(np.float32) for n in d.faces)
        vol = W.build_volume(wl)
        rep = vol.project_mobile3(wl.scale, wl.k1, 2000, warm=False)
        a, pl = po.divergence_metric(problem(wl), host(vol.velocity_u()), host(vol.velocity_v()),
                                     host(vol.velocity_w()))
        say(f"[metric] {wl.name}: card {rep.divergence:.6e} / {rep.divergence_plain:.6e}, oracle {a:.6e} / {pl:.6e}")
        assert rep.divergence > 0
        assert rep.divergence == a and rep.divergence_plain == pl
        vol.close()


@pytest.mark.parametrize("homogeneous", [True, False])
def test_backward_error_equals_the_oracle_on_a_given_pressure(homogeneous):
    """With max_iters = 0 the projection evaluates a given (warm-start)
    pressure: its backward error (delta3d_mobile.rs:243-271) against the f64
    oracle's on the same pressure, within the f32 rounding of b and r."""
    wl = W.tank(64, 32, 48, seed=12)
    wl.homogeneous_ghost = homogeneous
    prob = problem(wl)
    kind, _ = po.classify(prob)
    p0 = np.where(kind == 1, np.random.default_rng(3).uniform(-1, 1, wl.domain.cells), 0.0).astype(np.float32)
    vol = W.build_volume(wl)
    vol.p.copy_(torch.from_numpy(p0))
    rep = vol.project_mobile3(wl.scale, wl.k1, 0, warm=True)
    oref, *_ = po.project(prob, wl.us, wl.vs, wl.ws, wl.scale, wl.k1, 0, oracle_options(wl), p_warm=p0,
                          precision=64)
    say(f"[backward] card {rep.backward_error:.9e}, oracle {oref['backward_error']:.9e}, iterations "
        f"{rep.iterations}/{oref['iterations']}")
    assert rep.iterations == 0 and oref["iterations"] == 0
    assert rep.backward_error > 1e-3  # a pressure that is no solution
    assert rep.backward_error == pytest.approx(oref["backward_error"], rel=1e-5)
    assert rep.residual == pytest.approx(oref["residual"], rel=1e-5)
    vol.close()


# ---------------------------------------------------------------- forced acceptance paths
def test_divergence_tightening_restart_runs_on_the_card():
    """rel_tol = 1e-2: the residual target is met long before the plain-row
    divergence is below 1e-5, so the driver tightens the target
    (physical_target = actual * (1e-5 / reached)^2, delta3d_mobile.rs:556-568)
    and goes on; the oracle, given the same target, does the same."""
    wl = W.tank(64, 32, 64, seed=13)
    vol = W.build_volume(wl)
    rep = vol.project_mobile3(wl.scale, wl.k1, 2000, rel_tol=1e-2, warm=False)
    oref, *_, p64 = oracle64(wl, rel_tol=1e-2)
    err = rel_l2(wl, host(vol.pressure()), p64)
    say(f"[tighten] card {rep.iterations} it, residual {rep.residual:.2e}, divergence {rep.divergence_plain:.2e}; "
        f"oracle {oref['iterations']} it, residual {oref['residual']:.2e}; pressure rel L2 {err:.2e}")
    assert not rep.degraded and not oref["degraded"]
    assert rep.divergence_plain <= 1e-5 and oref["divergence_plain"] <= 1e-5
    # the target alone stops near 1e-2: the restart carried both far below it
    assert rep.residual < 1e-3 and oref["residual"] < 1e-3
    assert abs(rep.iterations - oref["iterations"]) <= 2
    assert err <= 1e-3


def test_round_off_floor_stop_runs_on_the_card():
    """rel_tol = 1e-20 cannot be met: the passes end every 5 iterations on a
    true-residual check (replace_every), the f64 pressure keeps improving
    until its backward error reaches the round-off certificate (or p repeats)
    and the driver stops on the floor (delta3d_mobile.rs:571-585), accepted.
    The f64 oracle floors on the same problem."""
    wl = W.tank(32, 32, 32, seed=3)
    vol = W.build_volume(wl, replace_every=5)
    rep = vol.project_mobile3(wl.scale, wl.k1, 400, rel_tol=1e-20, warm=False)
    oref, _, _, _, _, p64 = oracle64(wl, max_iters=400, rel_tol=1e-20)
    err = rel_l2(wl, host(vol.pressure()), p64)
    say(f"[floor] card {rep.iterations} it, floor {rep.floor}, backward error {rep.backward_error:.2e}, residual "
        f"{rep.residual:.2e}; oracle {oref['iterations']} it, floor {oref['floor']}; pressure rel L2 {err:.2e}")
    assert rep.floor and not rep.degraded and rep.iterations < 400
    assert oref["floor"] and not oref["degraded"]
    assert rep.residual < 1e-12
    assert err <= 1e-6  # both at the f64 solution; the card's p is its f32 rounding


def test_residual_replacement_keeps_the_projection():
    """replace_every = 3 on a tank that needs about 10 iterations: the true
    f64 residual replaces the recurrence three times; the result is the f64
    oracle's within the parity bar."""
    wl = W.tank(64, 48, 64, seed=7)
    vol = W.build_volume(wl, replace_every=3)
    rep = vol.project_mobile3(wl.scale, wl.k1, 2000, warm=False)
    oref, _, _, _, _, p64 = oracle64(wl)
    check(wl, vol, rep, p64, "replace every 3", oref=oref)
    assert rep.iterations <= oref["iterations"] + 3
    vol.close()
