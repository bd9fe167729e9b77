"""Benchmark: milliseconds per pressure projection to a 1e-6 relative true
residual on the 256x128x256 free-surface tank (plain MGPCG), one MI355X.

A step is one cold-start projection (p = 0) of the whole workload: geometry
rebuild, divergence and right-hand side, the preconditioned CG to 1e-6,
the gradient subtract and the acceptance metrics, with every input already
resident in HBM. Prints one JSON line; beside the headline it carries the
dominant kernel's roofline (live launch timing, HBM traffic from the
committed counter table), the CPU oracle's time for the same projection,
the bfloat16 V-cycle configuration with residual replacement, the learned
preconditioner in its two modes, the 48-patch batch, the surface step of the
sloshing tank and the host-pointer (PCIe-inclusive) projection.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms/projection to 1e-6 rel. residual, 256×128×256: plain MGPCG vs learned FCG"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per fine cell of one launch: every array the kernel
# must touch, read once and written once (u8 row code = 1 B, f32 = 4 B,
# f64 = 8 B, V-cycle value = zb B: 4 in f32 storage, 2 in f16/bf16; a
# coarse value is 1/8 of a fine one).
def bytes_per_cell(kernel, zb):
    table = {
        "k_q_brick<0>": 1 + zb + 4 + zb,           # red-black sweep, both colours: mask, z, r -> z'
        "k_q_brick<2>": 1 + zb + zb / 8 + 4 + zb,  # prolongation + sweep: mask, z, coarse x, r -> z'
        "k_q_gs<0>": 1 + zb + 4 + zb,              # red-black pass: mask, z, r -> z (whole quads)
        "k_q_gs_first": 1 + 4 + 4 + zb,            # first two passes from zero: mask, r, row_prec -> z
        "k_q_gs<2>": 1 + zb + zb / 8 + 4 + zb,     # prolongation + pass: mask, z, coarse x, r -> z'
        "k_q_spmv_dir": 1 + zb + 4 + 4 + 4,        # mask, z, d_old -> d_new, q
        "k_q_update": 8 + 8 + 4 + 4 + 4 + 4,       # p64 rw, d, q, r rw
        "k_q_restrict": 1 + zb + 4 + zb / 8,       # mask, z, r -> coarse r
        "k_residual": 1 + 8 + 4 + 4,               # mask, p64, b -> r
    }
    return table.get(kernel)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", default="256x128x256", help="X x Y x Z, Y vertical")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--kernel", default=None,
                    help="kernel timed live for the roofline (default: the fused fine sweep, or the single pass)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=3,
                    help="complete oracle projections timed for the CPU baseline (median)")
    ap.add_argument("--coarse-sweeps", type=int, default=2,
                    help="red-black sweeps on the coarsest level (2 keeps the 9 iterations of 8 on this tank)")
    ap.add_argument("--single-sweeps", action="store_true",
                    help="one launch per red-black colour on the fine level (else both colours in one launch)")
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16"],
                    help="V-cycle vector storage (arithmetic stays f32)")
    ap.add_argument("--weights", default=os.path.join(ROOT, "fluidisim_amd", "learned_weights.npz"),
                    help="trained network for the learned-FCG line (skipped when the file is absent)")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--batch", type=int, default=48, help="patches of the batched line (0: skip)")
    ap.add_argument("--batch-concurrency", default="1,4,8", help="concurrency levels timed for the batched line")
    ap.add_argument("--no-step", action="store_true", help="skip the surface-step line")
    ap.add_argument("--no-host", action="store_true", help="skip the host-pointer (PCIe-inclusive) line")
    ap.add_argument("--no-mixed", action="store_true", help="skip the 16-bit V-cycle storage line")
    ap.add_argument("--copy-probe-only", action="store_true",
                    help="run only the 256 MiB device copy (counter calibration) and exit")
    return ap.parse_args()


def copy_probe(torch, mib=256, reps=10):
    """Achievable HBM bandwidth: a plain device-to-device copy of `mib` MiB,
    timed with events; counts read + write bytes."""
    n = mib * 1024 * 1024 // 4
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return 2.0 * 4 * n / (ms * 1e-3) / 1e9


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    counter summary (FETCH_SIZE and WRITE_SIZE per dispatch, already scaled by
    the copy-kernel calibration); None when no summary covers the pair."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\|\s*`([^`]+)`\s*\|\s*`([^`]+)`\s*\|\s*([0-9.]+)\s*\|\s*([0-9.]+)\s*\|\s*([0-9]+)\s*\|", line)
        if m and m.group(1) == kernel and m.group(2) == workload:
            return int(m.group(5))
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(wl, runs):
    """The oracle (C++ restatement of the reference algorithm, f32 with the
    reference's chunk-64 f32 reductions, 1 thread) running the complete
    projection of the full workload `runs` times: the median wall time, and
    the iterations, residual and floor stop the runs reported (nothing is
    extrapolated). The f32 arithmetic floors above 1e-6 on this workload, so
    the FP64 oracle's complete projection to 1e-6 is timed once beside it as
    labelled context."""
    import platform
    from oracle import pyoracle as po
    d = wl.domain
    prob = po.Problem(d.nx, d.ny, d.nz, d.dx, wl.eta, rho=wl.rho, g_eff=wl.g_eff, rest=wl.rest,
                      homogeneous_ghost=wl.homogeneous_ghost, closed_lid=wl.closed_lid, solid=wl.solid)
    o = wl.options
    opts = po.options(multigrid=1, smoother=o["smoother"], max_levels=o["max_levels"], pre=o["pre_sweeps"],
                      post=o["post_sweeps"], coarse=o["coarse_sweeps"],
                      coarse_op=o.get("coarse_operator", 0))
    times = []
    for q in range(runs):
        t0 = time.perf_counter()
        rep, *_ = po.project(prob, wl.us, wl.vs, wl.ws, wl.scale, wl.k1, 2000, opts)
        times.append(time.perf_counter() - t0)
        log(f"[cpu] f32 oracle run {q}: {times[-1]:.2f} s, {rep['iterations']} iterations, residual "
            f"{rep['residual']:.2e}, floor {rep['floor']}")
    t0 = time.perf_counter()
    rep64, *_ = po.project(prob, wl.us, wl.vs, wl.ws, wl.scale, wl.k1, 2000, opts, precision=64)
    t64 = time.perf_counter() - t0
    log(f"[cpu] f64 oracle: {t64:.2f} s, {rep64['iterations']} iterations, residual {rep64['residual']:.2e}")
    ms = 1000.0 * float(np.median(times))
    cpu = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": round(ms, 1), "unit": "ms/projection", "cores": 1, "kind": "port",
        "sample": (f"C++ oracle (restatement of the reference MGPCG, same V-cycle settings, f32, chunk-64 f32 "
                   f"reductions, 1 thread): the complete cold-start projection of the full {d.nx}x{d.ny}x{d.nz} "
                   f"workload, median of {runs} runs ({sum(times):.1f} s in all)"),
        "runs_s": [round(t, 3) for t in times],
        "iterations": rep["iterations"], "residual": rep["residual"], "floor": rep["floor"],
        "degraded": rep["degraded"],
        "note": "the f32 arithmetic stops on its round-off floor above 1e-6 on this workload",
        "f64_context": {"ms": round(1000.0 * t64, 1), "iterations": rep64["iterations"],
                        "residual": rep64["residual"], "label": "FP64 oracle to 1e-6, 1 thread (context)"},
        "host_cpu": cpu, "host_threads": os.cpu_count(),
    }


def batch_line(args, torch, W):
    """Independent patches (edges 32..128 per axis), each autotuned, projected
    one after another and four at a time on their own streams."""
    from fluidisim_amd.volume import project_batch
    wls = W.patches(args.batch, seed=args.seed)
    vols = [W.build_volume(wl) for wl in wls]
    tiles = [vol.autotune()[0] for vol in vols]
    scales, k1s = [wl.scale for wl in wls], [wl.k1 for wl in wls]
    out = {"patches": len(wls), "cells": int(sum(wl.domain.cells for wl in wls)),
           "tiles": {str(t): tiles.count(t) for t in sorted(set(tiles))}}
    levels = [int(v) for v in args.batch_concurrency.split(",")]
    for conc in levels:
        project_batch(vols, scales, k1s, 2000, concurrency=conc)  # warm-up (graphs, caches)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(max(1, args.steps)):
            reps = project_batch(vols, scales, k1s, 2000, concurrency=conc)
        torch.cuda.synchronize()
        ms = 1000.0 * (time.perf_counter() - t0) / max(1, args.steps)
        out[f"ms_concurrency_{conc}"] = round(ms, 3)
        out["iterations_total"] = int(sum(r.iterations for r in reps))
        out["degraded"] = int(sum(r.degraded for r in reps))
        log(f"[bench] batch of {len(wls)} patches, concurrency {conc}: {ms:.3f} ms")
    best = min(levels, key=lambda c: out[f"ms_concurrency_{c}"])
    out["concurrency"] = best
    out["ms_per_batch"] = out[f"ms_concurrency_{best}"]
    out["mcells_per_s"] = round(out["cells"] / (out["ms_per_batch"] * 1e-3) / 1e6, 1)
    for vol in vols:
        vol.close()
    return out


def learned_fcg(wl, args, torch, _lib, W, vcycle):
    """The same projection with the trained network inside flexible CG
    (MGPCG fallback), timed like the plain one. vcycle: the network's output
    is the initial guess of one V-cycle (else learned_smooth sweeps follow)."""
    from fluidisim_amd.learned import Network
    net = Network()
    net.load(args.weights)
    vol = W.build_volume(wl, preconditioner=_lib.PRECOND_LEARNED, learned_budget=200, learned_vcycle=int(vcycle))
    net.attach(vol)

    def step():
        return vol.project_mobile3(wl.scale, wl.k1, 2000, rel_tol=1e-6, warm=False)

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = [step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    ms = 1000.0 * (time.perf_counter() - t0) / args.steps
    r = reps[-1]
    mode = "network + V-cycle" if vcycle else "network + red-black sweeps"
    log(f"[bench] learned FCG ({mode}): {ms:.3f} ms/projection, {r.iterations} iterations, "
        f"residual {r.residual:.2e}")
    vol.close()
    net.close()
    with np.load(args.weights, allow_pickle=False) as f:
        trained = {k: (f[k].tolist() if f[k].size < 16 else None) for k in f.files if k != "theta"}
    return {"preconditioner": mode, "ms_per_projection": round(ms, 4), "iterations": r.iterations,
            "residual": r.residual, "divergence_plain": r.divergence_plain, "degraded": r.degraded,
            "training": trained}


def host_pointer_line(vol, wl, steps, torch):
    """The same projection through the host-pointer entry point (fsp_project):
    the predicted faces copied in and p, u, v, w copied out over PCIe, from
    pageable numpy arrays and from pinned host arrays (torch pin_memory, the
    DMA rate); reported beside the device-resident value."""
    import ctypes as C
    from fluidisim_amd import _lib
    d = wl.domain
    nu, nv, nw = d.faces

    def arrays(pinned):
        def make(n, src=None):
            a = (torch.empty(n, dtype=torch.float32, pin_memory=True).numpy() if pinned
                 else np.empty(n, np.float32))
            a[:] = 0.0 if src is None else src
            return a
        return (make(nu, wl.us), make(nv, wl.vs), make(nw, wl.ws), make(d.cells), make(nu), make(nv), make(nw))

    def run(pinned):
        us, vs, ws, p, u, v, w = arrays(pinned)
        rep = _lib.Report()
        ms, transfer = [], []
        for q in range(steps + 1):
            t0 = time.perf_counter()
            st = _lib.lib().fsp_project(vol._h, us.ctypes.data, nu, vs.ctypes.data, nv, ws.ctypes.data, nw,
                                        p.ctypes.data, d.cells, 0, u.ctypes.data, nu, v.ctypes.data, nv,
                                        w.ctypes.data, nw, wl.scale, wl.k1, 1e-6, 2000, C.byref(rep))
            if st != 0:
                raise RuntimeError(f"fsp_project: status {st}")
            if q:  # the first call warms the path
                ms.append(1000.0 * (time.perf_counter() - t0))
                transfer.append(rep.ms_transfer)
        return {"ms_per_projection": round(float(np.mean(ms)), 4), "ms_transfer": round(float(np.mean(transfer)), 4)}

    out = {"bytes_moved": int(4 * (nu + nv + nw) * 2 + 4 * d.cells),
           "pageable": run(False), "pinned": run(True)}
    for kind in ("pageable", "pinned"):
        r = out[kind]
        log(f"[bench] host pointers ({kind}): {r['ms_per_projection']:.3f} ms/projection, of which "
            f"{r['ms_transfer']:.3f} ms PCIe transfer ({out['bytes_moved'] / r['ms_transfer'] / 1e6:.1f} GB/s)")
    return out


def mixed_precision_line(wl, vol, args, torch, _lib, W):
    """The 16-bit V-cycle configuration: bfloat16 storage of the V-cycle
    vectors (f32 arithmetic, f32 CG vectors, f64 pressure, residual and
    reductions) with the true residual replacing the recurrence every 50
    iterations, on the headline workload, timed like the headline. Its
    pressure is compared with the f32-storage projection's (relative L2); the
    comparison with the f64 oracle at this size is a test
    (test_c2_bf16_vcycle_with_residual_replacement_matches_the_f64_oracle).
    Binary16 storage is not timed: this workload's V-cycle values leave its
    range and the projection runs to its iteration limit, degraded."""
    p_ref = vol.pressure().double()
    out = {"replace_every": 50, "f16": "not timed: binary16 leaves its range on this workload (degraded)"}
    for name, storage in (("bf16", _lib.STORE_BF16),):
        v = W.build_volume(wl, vcycle_storage=storage, replace_every=50)

        def step():
            return v.project_mobile3(wl.scale, wl.k1, 2000, rel_tol=1e-6, warm=False)

        for _ in range(max(1, args.warmup)):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reps = [step() for _ in range(args.steps)]
        torch.cuda.synchronize()
        ms = 1000.0 * (time.perf_counter() - t0) / args.steps
        r = reps[-1]
        err = float(torch.linalg.norm(v.pressure().double() - p_ref) / torch.linalg.norm(p_ref))
        log(f"[bench] {name} V-cycle storage, replacement every 50: {ms:.3f} ms/projection, {r.iterations} "
            f"iterations, residual {r.residual:.2e}, pressure {err:.1e} from the f32-storage projection")
        out[name] = {"ms_per_projection": round(ms, 4), "iterations": r.iterations, "residual": r.residual,
                     "divergence_plain": r.divergence_plain, "degraded": r.degraded,
                     "p_rel_l2_vs_f32_storage": err}
        v.close()
    return out


def surface_step_line(args, torch, W):
    """The whole surface step (advect, project with a warm start,
    extrapolate, transport, balance and checks) on the sloshing tank of the
    same size, dt = 1 ms, timed over consecutive steps."""
    X, Y, Z = (int(v) for v in args.size.lower().split("x"))
    wl = W.slosh(X, Y, Z, seed=args.seed)
    vol = W.build_volume(wl)
    for _ in range(max(1, args.warmup)):
        vol.step_surface_mobile(1000, 4000)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = [vol.step_surface_mobile(1000, 4000) for _ in range(args.steps)]
    torch.cuda.synchronize()
    ms = 1000.0 * (time.perf_counter() - t0) / args.steps
    log(f"[bench] surface step ({wl.name}): {ms:.3f} ms/step, iterations {[r.iterations for r in reps]}")
    out = {"workload": wl.name, "cells": wl.domain.cells, "dt_us": 1000, "ms_per_step": round(ms, 4),
           "iterations": [r.iterations for r in reps], "balance_residual": vol.balance["residual"],
           "steps_before_timing": max(1, args.warmup)}
    vol.close()
    return out


def main():
    args = parse()
    if args.kernel is None:
        args.kernel = "k_q_gs<0>" if args.single_sweeps else "k_q_brick<0>"
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU design is "
              f"described, not built.", flush=True)
        sys.exit(2)
    import torch
    if args.copy_probe_only:
        print(json.dumps({"copy_probe_gbs": round(copy_probe(torch), 1)}), flush=True)
        return
    from fluidisim_amd import _lib
    from fluidisim_amd import workloads as W

    X, Y, Z = (int(v) for v in args.size.lower().split("x"))
    t0 = time.perf_counter()
    wl = W.tank(X, Y, Z, seed=args.seed)
    wl.options["coarse_sweeps"] = args.coarse_sweeps
    log(f"[bench] workload {wl.name}: {wl.domain.cells} cells, generated in {time.perf_counter() - t0:.1f} s")
    storage = {"f32": _lib.STORE_F32, "f16": _lib.STORE_F16, "bf16": _lib.STORE_BF16}[args.storage]
    sweeps = _lib.SWEEP_SINGLE if args.single_sweeps else _lib.SWEEP_FUSED
    vol = W.build_volume(wl, vcycle_storage=storage, fine_sweeps=sweeps)
    log(f"[bench] solver ready, {vol.multigrid_levels()} coarse levels")

    def step():
        return vol.project_mobile3(wl.scale, wl.k1, 2000, rel_tol=1e-6, warm=False)

    for s in range(args.warmup):
        r = step()
        log(f"[bench] warmup {s}: {r.iterations} it, residual {r.residual:.2e}, {r.ms_device:.2f} ms")
    L = _lib.lib()
    # Timed steps, uninstrumented (each CG iteration replays its HIP graph).
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = [step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = 1000.0 * (t1 - t0) / args.steps
    # The same steps again with the roofline kernel bracketed by HIP events on
    # the library's stream (event recording needs eager launches, so these
    # steps are not the timed ones).
    max_launches = 100000
    L.fsp_profile_begin(vol._h, args.kernel.encode(), max_launches)
    torch.cuda.synchronize()
    ti0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ti1 = time.perf_counter()
    launches, total = C.c_uint32(), C.c_double()
    durations = np.zeros(max_launches, np.float32)
    active = np.zeros(max_launches, np.int32)
    L.fsp_profile_end(vol._h, C.byref(launches), C.byref(total), durations.ctypes.data, active.ctypes.data,
                      max_launches)
    log(f"[bench] instrumented steps: {1000.0 * (ti1 - ti0) / args.steps:.3f} ms/projection")
    last = reps[-1]
    log(f"[bench] timed {args.steps} steps: {ms_per_step:.3f} ms/projection, {last.iterations} iterations, "
        f"residual {last.residual:.2e}, divergence {last.divergence_plain:.2e}, degraded {last.degraded}")

    # Roofline of the timed kernel: algorithmic bytes per launch / mean launch
    # duration (HIP events on the library's stream over the instrumented steps).
    # Launches queued after convergence return at once: only the launches that
    # started with the CG loop active (the flag the library records on the
    # device in stream order before each timed launch) are averaged.
    bpc = bytes_per_cell(args.kernel, 4 if args.storage == "f32" else 2)
    roofline = None
    if bpc is not None and launches.value > 0:
        m = min(launches.value, max_launches)
        d = durations[:m].astype(np.float64)
        eff = d[active[:m] != 0]
        avg_ms = float(eff.mean())
        bytes_launch = bpc * wl.domain.cells
        achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "instrumented_ms_per_step": round(1000.0 * (ti1 - ti0) / args.steps, 4),
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(args.kernel, wl.name),
                    "kernel": args.kernel, "copy_probe_gbs": round(copy_probe(torch), 1),
                    "launches": int(eff.size), "noop_launches": int(d.size - eff.size),
                    "launches_per_projection": round(eff.size / args.steps, 2),
                    "avg_us": round(avg_ms * 1000.0, 2), "bytes_per_launch": int(bytes_launch)}
        log(f"[bench] {args.kernel}: {eff.size} launches, {avg_ms * 1000:.1f} us avg, {achieved:.0f} GB/s")

    host = None if args.no_host else host_pointer_line(vol, wl, max(1, min(args.steps, 3)), torch)

    mixed = None if args.no_mixed else mixed_precision_line(wl, vol, args, torch, _lib, W)

    batched = batch_line(args, torch, W) if args.batch > 0 else None
    if batched:
        # the headline projection's rate, for comparison with the patches'
        batched["single_domain_mcells_per_s"] = round(wl.domain.cells / (ms_per_step * 1e-3) / 1e6, 1)

    learned = learned_sweeps = None
    if not args.no_learned and os.path.exists(args.weights):
        learned = learned_fcg(wl, args, torch, _lib, W, vcycle=True)
        learned_sweeps = learned_fcg(wl, args, torch, _lib, W, vcycle=False)

    step = None if args.no_step else surface_step_line(args, torch, W)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(wl, max(1, args.cpu_runs))

    line = {
        "metric": METRIC, "value": round(ms_per_step, 4), "unit": "ms/projection", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl.name, "domain": [wl.domain.nx, wl.domain.ny, wl.domain.nz],
                   "cells": wl.domain.cells, "solver": f"plain MGPCG (RBGS V-cycle 2+2, Galerkin coarse levels, {args.coarse_sweeps} coarsest sweeps)",
                   "precision": "f32 CG vectors, f64 pressure/residual/reductions",
                   "vcycle_storage": args.storage, "rel_tol": 1e-6,
                   "iterations": last.iterations, "residual": last.residual,
                   "divergence_plain": last.divergence_plain, "degraded": last.degraded,
                   "ms_device_last": round(last.ms_device, 4), "seed": args.seed},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mixed_precision": mixed,
        "learned_fcg": learned,
        "learned_fcg_sweeps": learned_sweeps,
        "batched": batched,
        "surface_step": step,
        "host_pointers": host,
    }
    print(json.dumps(line), flush=True)
    vol.close()


if __name__ == "__main__":
    main()
