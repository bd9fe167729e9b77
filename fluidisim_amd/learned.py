"""The learned preconditioner (include/fsp_learned.h) from Python: the
network, its training pairs made by the plain MGPCG, training, and use
inside the projection (FSP_PRECOND_LEARNED, flexible CG with MGPCG as the
fallback).

Training pairs of a workload: for residuals r_k = b - A p_k of the plain
MGPCG after k iterations (cold start, k = 0, 1, 2, 4), input x = (r_k*sigma,
wet, ghost row, solid) with sigma = 1/rms(r_k), target sigma * (p - p_k) /
kappa on wet cells, p the converged MGPCG pressure, kappa = min(nx, ny,
nz)^2 / 16 (an output scale near 1; flexible CG does not see a constant
factor, so the projection uses z = y / sigma). Inside the projection the
network's output is followed by learned_smooth red-black sweeps.
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from . import _lib
from . import workloads as W
from .volume import _ptr, _raise, _torch_device_first

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "learned_weights.npz")


def _torch():
    import torch
    return torch


def kappa(nx, ny, nz):
    return float(min(nx, ny, nz)) ** 2 / 16.0


class Network:
    """One network on one device (fsp_net)."""

    def __init__(self, device: int = 0):
        _torch_device_first(device)
        h = C.c_void_p()
        _raise(_lib.lib().fsp_net_create(device, C.byref(h)), "fsp_net_create")
        self._h = h
        self.device = device
        self.P = int(_lib.lib().fsp_net_param_count())

    def close(self):
        if self._h:
            _lib.lib().fsp_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, theta):
        t = np.ascontiguousarray(theta, np.float32)
        _raise(_lib.lib().fsp_net_set_params(self._h, t.ctypes.data, t.size), "fsp_net_set_params")

    def params(self):
        t = np.empty(self.P, np.float32)
        _raise(_lib.lib().fsp_net_get_params(self._h, t.ctypes.data, t.size), "fsp_net_get_params")
        return t

    def save(self, path=WEIGHTS, **meta):
        np.savez(path, theta=self.params(), **{k: np.asarray(v) for k, v in meta.items()})

    def load(self, path=WEIGHTS):
        with np.load(path, allow_pickle=False) as f:
            self.set_params(f["theta"])

    def features(self, vol, r, sigma: float = 0.0):
        """(x, sigma): binary16 features of residual r on vol's geometry."""
        torch = _torch()
        vol._upload_geometry()
        r = torch.as_tensor(r, dtype=torch.float32, device=vol.p.device).contiguous()
        x = torch.empty(4 * vol.domain.cells, dtype=torch.int16, device=vol.p.device)
        s = C.c_float()
        vol._sync()
        _raise(_lib.lib().fsp_net_features_device(self._h, vol._h, _ptr(r), _ptr(x), sigma, C.byref(s)),
               "fsp_net_features_device")
        return x, float(s.value)

    def forward(self, x, shape):
        torch = _torch()
        nx, ny, nz = shape
        y = torch.empty(nx * ny * nz, dtype=torch.float32, device=x.device)
        torch.cuda.synchronize(x.device)
        _raise(_lib.lib().fsp_net_forward_device(self._h, nx, ny, nz, _ptr(x), _ptr(y)), "fsp_net_forward_device")
        return y

    def forward_static(self, x, shape, y):
        """y = network(x) through fsp_net_forward_static_device (no device
        memory reserved; y is the caller's). Returns the status."""
        nx, ny, nz = shape
        return _lib.lib().fsp_net_forward_static_device(self._h, nx, ny, nz, _ptr(x), _ptr(y))

    def loss_grad(self, x, target, wet, shape):
        torch = _torch()
        nx, ny, nz = shape
        loss = C.c_double()
        g = np.empty(self.P, np.float32)
        torch.cuda.synchronize(x.device)
        _raise(_lib.lib().fsp_net_loss_grad_device(self._h, nx, ny, nz, _ptr(x), _ptr(target), _ptr(wet),
                                                   C.byref(loss), g.ctypes.data), "fsp_net_loss_grad_device")
        return loss.value, g

    def train_step(self, x, target, wet, shape, lr):
        torch = _torch()
        nx, ny, nz = shape
        loss = C.c_double()
        skipped = C.c_int32()
        torch.cuda.synchronize(x.device)
        _raise(_lib.lib().fsp_net_train_step_device(self._h, nx, ny, nz, _ptr(x), _ptr(target), _ptr(wet), lr,
                                                    C.byref(loss), C.byref(skipped)), "fsp_net_train_step_device")
        return loss.value, bool(skipped.value)

    def attach(self, vol):
        _raise(_lib.lib().fsp_attach_net(vol._h, self._h), "fsp_attach_net")


def init_params(seed: int = 0):
    """He-style initial weights (zero biases), in the library's layer order."""
    layers = [(4, 8, True), (8, 8, True), (8, 16, True), (16, 16, True), (16, 32, True), (32, 32, True),
              (32, 16, True), (16, 16, True), (16, 8, True), (8, 8, True), (8, 1, False)]
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co, relu in layers:
        std = np.sqrt((2.0 if relu else 1.0) / (27 * ci))
        parts.append(rng.standard_normal(co * 27 * ci) * std)
        parts.append(np.zeros(co))
    return np.concatenate(parts).astype(np.float32)


def training_pairs(net: Network, wl: W.Workload, iters=(0, 1, 2, 4), max_iters: int = 2000):
    """Pairs (x, target, wet) of one workload: the residuals r_k = b - A p_k
    of the plain MGPCG after k iterations (k in `iters`, p_0 = 0) and their
    errors e_k = p - p_k, so the network learns A^-1 on the residuals flexible
    CG meets, not on b alone. Returns (pairs, report of the full solve)."""
    torch = _torch()
    vol = W.build_volume(wl)
    rep = vol.project_mobile3(wl.scale, wl.k1, max_iters, warm=False)
    p_star = vol.pressure().clone()
    div = vol.divergence(vol.us, vol.vs, vol.ws)
    rhs, prec = vol.assemble_pressure_problem_for_trials(div, wl.scale)
    d = wl.domain
    kap = kappa(d.nx, d.ny, d.nz)
    pairs = []
    for k in iters:
        if k == 0:
            p_k = torch.zeros_like(p_star)
        else:
            vol.project_mobile3(wl.scale, wl.k1, k, warm=False)
            p_k = vol.pressure().clone()
        r_k = rhs - vol.apply_pressure_operator_for_trials(p_k)
        x, sigma = net.features(vol, r_k)
        wet = (x.view(-1, 4)[:, 1] != 0).to(torch.uint8)
        target = ((p_star - p_k) * (sigma / kap)) * wet.to(torch.float32)
        pairs.append((x, target.contiguous(), wet.contiguous()))
    vol.close()
    return pairs, rep


def tank_shape(X, Y, Z):
    """Reference (nx, ny, nz) of tank(X, Y, Z): y up in the workload, z up in the solver."""
    return X, Z, Y


def train(net: Network, size=(128, 128, 128), samples=64, epochs=4, lr=1e-3, seed0=1000, log=print,
          on_epoch=None, lr_decay=1.0):
    """Trains on `samples` tank workloads of the given size (distinct seeds),
    `epochs` passes, the step size multiplied by lr_decay after each pass;
    returns the per-epoch mean losses. on_epoch(e, history) runs after each
    epoch (e.g. to save the parameters)."""
    X, Y, Z = size
    shape = tank_shape(X, Y, Z)
    t0 = time.time()
    pairs = []
    for s in range(samples):
        wl = W.tank(X, Y, Z, seed=seed0 + s)
        got, rep = training_pairs(net, wl)
        pairs.extend(got)
        if s % 16 == 0:
            log(f"[learned] pair {s}: {rep.iterations} MGPCG iterations ({time.time() - t0:.1f} s)")
    history = []
    rng = np.random.default_rng(seed0)
    for e in range(epochs):
        losses, skipped = [], 0
        for q in rng.permutation(len(pairs)):
            x, t, w = pairs[q]
            loss, sk = net.train_step(x, t, w, shape, lr * lr_decay ** e)
            skipped += int(sk)
            if not sk:
                losses.append(loss)
        history.append(float(np.mean(losses)) if losses else float("nan"))
        log(f"[learned] epoch {e}: loss {history[-1]:.4e} ({skipped} skipped steps, {time.time() - t0:.1f} s)")
        if on_epoch is not None:
            on_epoch(e, history)
    return history


def main():
    import argparse
    ap = argparse.ArgumentParser(description="train the learned preconditioner on MGPCG pairs")
    ap.add_argument("--size", default="128x128x128")
    ap.add_argument("--samples", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-decay", type=float, default=1.0, help="step-size factor per epoch")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=WEIGHTS)
    a = ap.parse_args()
    size = tuple(int(v) for v in a.size.lower().split("x"))
    torch = _torch()
    torch.zeros(1, device="cuda")  # the device's runtime state, before the library's first call
    net = Network()
    net.set_params(init_params(a.seed))
    def save(e, hist):
        net.save(a.out, size=np.array(size), samples=a.samples, pairs=4 * a.samples, epochs=e + 1, lr=a.lr,
                 lr_decay=a.lr_decay, history=np.array(hist))
        print(f"[learned] saved {a.out} after epoch {e}", flush=True)

    train(net, size, a.samples, a.epochs, a.lr, log=lambda m: print(m, flush=True), on_epoch=save,
          lr_decay=a.lr_decay)


if __name__ == "__main__":
    main()
