"""CPU ORACLE — test infrastructure only.

ctypes bindings of ``oracle/liboracle.so`` (the C++ restatement of the
reference projection in ``oracle.hpp``). Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import
this module, as the checker or as the timed CPU baseline; the product path
never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class Geometry(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("dx", C.c_float),
        ("rho", C.c_float), ("g_eff", C.c_float), ("rest", C.c_float),
        ("homogeneous_ghost", C.c_int32), ("closed_lid", C.c_int32),
        ("eta", C.c_void_p), ("eta_roundoff", C.c_void_p), ("solid", C.c_void_p),
        ("open_u", C.c_void_p), ("open_v", C.c_void_p), ("open_w", C.c_void_p),
    ]


class Options(C.Structure):
    _fields_ = [
        ("multigrid", C.c_int32), ("smoother", C.c_int32), ("max_levels", C.c_int32),
        ("pre_sweeps", C.c_int32), ("post_sweeps", C.c_int32), ("coarse_sweeps", C.c_int32),
        ("gauge", C.c_int32), ("coarse_op", C.c_int32), ("storage", C.c_int32), ("flexible", C.c_int32),
        ("level_pre", C.c_int32), ("level_post", C.c_int32), ("rel_tol", C.c_double),
    ]


class Report(C.Structure):
    _fields_ = [
        ("refinements", C.c_uint32), ("iterations", C.c_uint32), ("degraded", C.c_int32),
        ("floor", C.c_int32), ("residual", C.c_double), ("divergence", C.c_double),
        ("backward_error", C.c_double), ("divergence_plain", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, f32, u32, i32, sz = C.c_void_p, C.c_float, C.c_uint32, C.c_int, C.c_size_t
        L.orc_project.argtypes = [i32, vp, vp, vp, vp, vp, vp, f32, f32, u32, vp, vp, vp, vp, vp, vp]
        L.orc_apply.argtypes = [vp, vp, vp]
        L.orc_project_graded.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, vp, f32, f32, u32, vp, vp, vp, vp, vp, vp]
        L.orc_graded_op.argtypes = [i32, vp, vp, i32, vp, vp, C.c_double]
        L.orc_project_coupled.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, f32, f32, u32, i32, u32, vp, vp, vp, vp, vp,
                                          vp]
        L.orc_cut.argtypes = [i32, i32, i32, f32, vp, vp, vp, vp, vp, vp]
        L.orc_apic_project.argtypes = [i32, i32, i32, i32, f32, f32, f32, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_apic_apply.argtypes = [i32, i32, i32, vp, vp, vp, vp]
        L.orc_apic_divergence_metric.argtypes = [i32, i32, i32, vp, vp, vp, vp]
        L.orc_apic_divergence_metric.restype = C.c_double
        L.orc_round_storage.argtypes = [i32, vp, C.c_size_t]
        L.orc_round_storage.restype = None
        L.orc_apply64.argtypes = [vp, vp, vp]
        L.orc_assemble.argtypes = [vp, vp, f32, vp, vp]
        L.orc_divergence.argtypes = [vp, vp, vp, vp, vp]
        L.orc_correct.argtypes = [vp, vp, vp, vp, vp, f32, vp, vp, vp, vp, vp]
        L.orc_correct_coupled.argtypes = [vp, vp, vp, vp, vp, vp, f32, vp, vp, vp, vp, vp]
        L.orc_divergence_metric.argtypes = [vp, vp, vp, vp, vp, vp]
        L.orc_backward_error.argtypes = [i32, vp, vp, vp, vp, vp]
        L.orc_coarse_apply.argtypes = [i32, vp, vp, i32, vp, vp]
        L.orc_vcycle.argtypes = [i32, vp, vp, vp, vp]
        L.orc_classify.argtypes = [vp, vp, vp]
        L.orc_coarse_level.argtypes = [vp, vp, i32, vp, vp]
        L.orc_level_count3.argtypes = [sz, sz, sz]
        L.orc_level_count3.restype = sz
        L.orc_hierarchy_floats3.argtypes = [sz, sz, sz]
        L.orc_hierarchy_floats3.restype = sz
        L.orc_fingerprint.argtypes = [vp, sz]
        L.orc_fingerprint.restype = C.c_uint64
        L.orc_roundoff_3d.restype = C.c_float
        L.orc_fnv_bytes.argtypes = [C.c_char_p, sz]
        L.orc_fnv_bytes.restype = C.c_uint64
        L.orc_step_surface.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, vp, u32, C.c_uint64, u32, vp, vp, vp]
        L.orc_advect.argtypes = [vp, vp, vp, vp, f32, vp, vp, vp]
        L.orc_extrapolate.argtypes = [vp, vp, vp, vp]
        L.orc_transport.argtypes = [vp, vp, vp, f32, vp, vp, vp, vp]
        L.orc_mobile_in_bounds.argtypes = [vp]
        L.orc_perturbation_volume.argtypes = [vp]
        L.orc_perturbation_volume.restype = C.c_double
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


@dataclass
class Problem:
    """Geometry arrays of one projection (reference layout, x fastest, z up)."""
    nx: int
    ny: int
    nz: int
    dx: float
    eta: np.ndarray
    rho: float = 1025.0
    g_eff: float = 9.81
    rest: float = 0.0
    homogeneous_ghost: bool = True
    closed_lid: bool = False
    solid: np.ndarray | None = None
    eta_roundoff: np.ndarray | None = None
    open_u: np.ndarray | None = None
    open_v: np.ndarray | None = None
    open_w: np.ndarray | None = None

    @property
    def cells(self):
        return self.nx * self.ny * self.nz

    @property
    def face_counts(self):
        nx, ny, nz = self.nx, self.ny, self.nz
        return (nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)

    def geometry(self):
        self._keep = [np.ascontiguousarray(self.eta, np.float32)]
        def keep(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dt)
            self._keep.append(a)
            return a.ctypes.data
        g = Geometry()
        g.nx, g.ny, g.nz, g.dx = self.nx, self.ny, self.nz, self.dx
        g.rho, g.g_eff, g.rest = self.rho, self.g_eff, self.rest
        g.homogeneous_ghost = int(self.homogeneous_ghost)
        g.closed_lid = int(self.closed_lid)
        g.eta = self._keep[0].ctypes.data
        g.eta_roundoff = keep(self.eta_roundoff, np.float32)
        g.solid = keep(self.solid, np.uint8)
        g.open_u = keep(self.open_u, np.float32)
        g.open_v = keep(self.open_v, np.float32)
        g.open_w = keep(self.open_w, np.float32)
        return g


def options(multigrid=1, smoother=0, max_levels=-1, pre=2, post=2, coarse=8, gauge=1, coarse_op=0, storage=0,
            flexible=0, level_pre=-1, level_post=-1, rel_tol=1e-6):
    return Options(multigrid, smoother, max_levels, pre, post, coarse, gauge, coarse_op, storage, flexible,
                   level_pre, level_post, rel_tol)


def project(prob: Problem, us, vs, ws, scale, k1, max_iters=20000, opts=None, p_warm=None, precision=32):
    """Run the projection; returns (report dict, p, u, v, w[, p64])."""
    g = prob.geometry()
    o = opts if opts is not None else options()
    nu, nv, nw = prob.face_counts
    us = np.ascontiguousarray(us, np.float32)
    vs = np.ascontiguousarray(vs, np.float32)
    ws = np.ascontiguousarray(ws, np.float32)
    pw = None if p_warm is None else np.ascontiguousarray(p_warm, np.float32)
    p = np.zeros(prob.cells, np.float32)
    u = np.zeros(nu, np.float32)
    v = np.zeros(nv, np.float32)
    w = np.zeros(nw, np.float32)
    p64 = np.zeros(prob.cells, np.float64) if precision == 64 else None
    rep = Report()
    st = lib().orc_project(precision, C.byref(g), C.byref(o), _ptr(us), _ptr(vs), _ptr(ws), _ptr(pw),
                           scale, k1, max_iters, _ptr(p), _ptr(u), _ptr(v), _ptr(w), _ptr(p64), C.byref(rep))
    if st != 0:
        raise RuntimeError(f"oracle project failed: status {st}")
    out = (rep.as_dict(), p, u, v, w)
    return out + (p64,) if precision == 64 else out


class Coupling(C.Structure):
    _fields_ = [("total", C.c_void_p), ("bg_up", C.c_void_p), ("bg_x", C.c_void_p), ("bg_y", C.c_void_p)]


REFINE_NONE, REFINE_STEP, REFINE_ALWAYS = 0, 1, 2


def project_coupled(prob: Problem, us, vs, ws, scale, k1, total, bg_up=None, bg_x=None, bg_y=None, refine=REFINE_NONE,
                    max_iters=20000, opts=None, p_warm=None, precision=32, refine_iters=0):
    """The projection over a coupled surface (total heights and background
    ghost terms; None terms are zeros), with refine_mobile3 as `refine`
    says (bounded by refine_iters, else max_iters); returns (report dict, p,
    u, v, w[, p64])."""
    g = prob.geometry()
    o = opts if opts is not None else options()
    nu, nv, nw = prob.face_counts
    arr = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
    keep = [arr(total), arr(bg_up), arr(bg_x), arr(bg_y)]
    cpl = Coupling(*[_ptr(a) for a in keep])
    us, vs, ws = arr(us), arr(vs), arr(ws)
    pw = arr(p_warm)
    p = np.zeros(prob.cells, np.float32)
    u = np.zeros(nu, np.float32)
    v = np.zeros(nv, np.float32)
    w = np.zeros(nw, np.float32)
    p64 = np.zeros(prob.cells, np.float64) if precision == 64 else None
    rep = Report()
    st = lib().orc_project_coupled(precision, C.byref(g), C.byref(o), C.byref(cpl), _ptr(us), _ptr(vs), _ptr(ws),
                                   _ptr(pw), scale, k1, max_iters, refine, refine_iters, _ptr(p), _ptr(u), _ptr(v),
                                   _ptr(w),
                                   _ptr(p64), C.byref(rep))
    if st != 0:
        raise RuntimeError(f"oracle project_coupled failed: status {st}")
    out = (rep.as_dict(), p, u, v, w)
    return out + (p64,) if precision == 64 else out


class Nodes(C.Structure):
    _fields_ = [("layers", C.c_void_p), ("count", C.c_size_t)]


def _nodes(nodes):
    arr = np.ascontiguousarray(nodes, np.int32)
    return Nodes(arr.ctypes.data, arr.size), arr


def project_graded(prob: Problem, us, vs, ws, scale, k1, nodes, coupling=None, max_iters=20000, opts=None,
                   p_warm=None, precision=32):
    """project_mobile3 over the graded column on `nodes` (fine layers, from 0
    to nz - 1, increasing; delta3d_graded.rs:310-447), coupling (total, bg_up,
    bg_x, bg_y) or None; returns (report dict, p, u, v, w[, p64])."""
    g = prob.geometry()
    o = opts if opts is not None else options()
    nu, nv, nw = prob.face_counts
    arr = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
    keep = [arr(x) for x in (coupling or ())]
    cpl = Coupling(*[_ptr(x) for x in keep]) if coupling is not None else None
    nd, nd_keep = _nodes(nodes)
    us, vs, ws, pw = arr(us), arr(vs), arr(ws), arr(p_warm)
    p = np.zeros(prob.cells, np.float32)
    u, v, w = np.zeros(nu, np.float32), np.zeros(nv, np.float32), np.zeros(nw, np.float32)
    p64 = np.zeros(prob.cells, np.float64) if precision == 64 else None
    rep = Report()
    st = lib().orc_project_graded(precision, C.byref(g), C.byref(o), None if cpl is None else C.byref(cpl),
                                  C.byref(nd), _ptr(us), _ptr(vs), _ptr(ws), _ptr(pw), scale, k1, max_iters, _ptr(p),
                                  _ptr(u), _ptr(v), _ptr(w), _ptr(p64), C.byref(rep))
    if st != 0:
        raise RuntimeError(f"oracle project_graded failed: status {st}")
    out = (rep.as_dict(), p, u, v, w)
    return out + (p64,) if precision == 64 else out


GRADED_EXPAND, GRADED_RESTRICT, GRADED_APPLY, GRADED_WEIGHTS, GRADED_RHS = 0, 1, 2, 3, 4


def graded_op(prob: Problem, nodes, op, x=None, precision=64, scale=1.0):
    """The graded column's P (op 0: graded -> cells), P^T (1: cells ->
    graded), P^T A P (2), node weights (3) and right-hand side P^T b of the
    predicted faces x = (u, v, w) concatenated (4), computed in f64 or f32;
    a non-OK status raises RuntimeError."""
    g = prob.geometry()
    nd, nd_keep = _nodes(nodes)
    n = prob.nx * prob.ny * len(nd_keep)
    size = {0: prob.cells, 1: n, 2: n, 3: len(nd_keep), 4: n}[op]
    x = None if x is None else np.ascontiguousarray(x, np.float64)
    out = np.zeros(size, np.float64)
    st = lib().orc_graded_op(precision, C.byref(g), C.byref(nd), op, _ptr(x), _ptr(out), scale)
    if st != 0:
        raise RuntimeError(f"oracle graded_op failed: status {st}")
    return out


def apply(prob: Problem, x):
    g = prob.geometry()
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(prob.cells, np.float32)
    assert lib().orc_apply(C.byref(g), _ptr(x), _ptr(out)) == 0
    return out


def apply64(prob: Problem, x):
    g = prob.geometry()
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros(prob.cells, np.float64)
    assert lib().orc_apply64(C.byref(g), _ptr(x), _ptr(out)) == 0
    return out


def assemble(prob: Problem, div, scale):
    g = prob.geometry()
    div = np.ascontiguousarray(div, np.float32)
    rhs = np.zeros(prob.cells, np.float32)
    prec = np.zeros(prob.cells, np.float32)
    assert lib().orc_assemble(C.byref(g), _ptr(div), scale, _ptr(rhs), _ptr(prec)) == 0
    return rhs, prec


def divergence(prob: Problem, u, v, w):
    g = prob.geometry()
    u, v, w = (np.ascontiguousarray(a, np.float32) for a in (u, v, w))
    out = np.zeros(prob.cells, np.float32)
    assert lib().orc_divergence(C.byref(g), _ptr(u), _ptr(v), _ptr(w), _ptr(out)) == 0
    return out


def correct(prob: Problem, us, vs, ws, p, k1, coupling=None):
    """correct_mobile3 of p (delta3d_mobile.rs:284-343) and the acceptance
    metric of the result; coupling: (total, bg_up, bg_x, bg_y) or None."""
    g = prob.geometry()
    us, vs, ws, p = (np.ascontiguousarray(a, np.float32) for a in (us, vs, ws, p))
    nu, nv, nw = prob.face_counts
    u, v, w = np.zeros(nu, np.float32), np.zeros(nv, np.float32), np.zeros(nw, np.float32)
    a, b = C.c_double(), C.c_double()
    keep = [None if x is None else np.ascontiguousarray(x, np.float32) for x in (coupling or ())]
    cpl = Coupling(*[_ptr(x) for x in keep]) if coupling is not None else None
    assert lib().orc_correct_coupled(C.byref(g), None if cpl is None else C.byref(cpl), _ptr(us), _ptr(vs), _ptr(ws),
                                     _ptr(p), k1, _ptr(u), _ptr(v), _ptr(w), C.byref(a), C.byref(b)) == 0
    return u, v, w, a.value, b.value


def divergence_metric(prob: Problem, u, v, w):
    """(all, plain): the acceptance metric of delta3d_mobile.rs:398-433 on the given faces."""
    g = prob.geometry()
    u, v, w = (np.ascontiguousarray(a, np.float32) for a in (u, v, w))
    a, b = C.c_double(), C.c_double()
    assert lib().orc_divergence_metric(C.byref(g), _ptr(u), _ptr(v), _ptr(w), C.byref(a), C.byref(b)) == 0
    return a.value, b.value


def backward_error(prob: Problem, p, rhs, res, precision=64):
    """max |r| / (|b| + |A| |p|) over wet rows (delta3d_mobile.rs:243-271)."""
    g = prob.geometry()
    p, rhs, res = (np.ascontiguousarray(a, np.float64) for a in (p, rhs, res))
    out = C.c_double()
    assert lib().orc_backward_error(precision, C.byref(g), _ptr(p), _ptr(rhs), _ptr(res), C.byref(out)) == 0
    return out.value


def coarse_apply(prob: Problem, level, x, opts=None, precision=64):
    """row_sums of coarse level `level` applied to x (delta3d_multigrid.rs:216-255)."""
    g = prob.geometry()
    o = opts if opts is not None else options()
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    st = lib().orc_coarse_apply(precision, C.byref(g), C.byref(o), level, _ptr(x), _ptr(out))
    if st != 0:
        raise RuntimeError(f"no coarse level {level}")
    return out


def vcycle(prob: Problem, r, opts=None, precision=32):
    g = prob.geometry()
    o = opts if opts is not None else options()
    r = np.ascontiguousarray(r, np.float32)
    z = np.zeros(prob.cells, np.float64)
    assert lib().orc_vcycle(precision, C.byref(g), C.byref(o), _ptr(r), _ptr(z)) == 0
    return z


def classify(prob: Problem):
    g = prob.geometry()
    kind = np.zeros(prob.cells, np.int8)
    ghost = np.zeros(prob.cells, np.uint8)
    assert lib().orc_classify(C.byref(g), _ptr(kind), _ptr(ghost)) == 0
    return kind, ghost


def coarse_level(prob: Problem, level, opts=None):
    g = prob.geometry()
    o = opts if opts is not None else options()
    nx, ny, nz = prob.nx, prob.ny, prob.nz
    for _ in range(level + 1):
        nx, ny, nz = nx // 2, ny // 2, nz // 2
    kind = np.zeros(nx * ny * nz, np.float32)
    diag = np.zeros(nx * ny * nz, np.float32)
    st = lib().orc_coarse_level(C.byref(g), C.byref(o), level, _ptr(kind), _ptr(diag))
    if st != 0:
        raise RuntimeError(f"no coarse level {level}")
    return kind, diag


def level_count3(nx, ny, nz):
    return lib().orc_level_count3(nx, ny, nz)


def hierarchy_floats3(nx, ny, nz):
    return lib().orc_hierarchy_floats3(nx, ny, nz)


def round_storage(storage, values):
    """The oracle's rounding of a stored V-cycle value (0 f32, 1 binary16, 2 bfloat16)."""
    a = np.array(values, np.float32, copy=True)
    lib().orc_round_storage(storage, _ptr(a), a.size)
    return a


def fnv_bytes(data: bytes):
    return lib().orc_fnv_bytes(data, len(data))


def fingerprint(values):
    a = np.ascontiguousarray(values, np.float32)
    return lib().orc_fingerprint(_ptr(a), a.size)


# ---------------------------------------------------------------- S385 bench case
_libm = C.CDLL("libm.so.6")
_libm.sinf.argtypes = [C.c_float]
_libm.sinf.restype = C.c_float


def _sinf(x):
    # Rust's f32::sin is the platform libm's sinf.
    return _libm.sinf(x)


def cut_geometry(nx: int, ny: int, nz: int, dx: float, bottom) -> dict:
    """The reference's cut of a bottom height field given at the column
    centres (delta3d_cut.rs:172-258): volume fractions, the three families
    of face apertures and the column floors, f32."""
    b = np.ascontiguousarray(bottom, np.float32)
    assert b.size == nx * ny
    out = {"frac": np.zeros(nx * ny * nz, np.float32), "open_u": np.zeros((nx + 1) * ny * nz, np.float32),
           "open_v": np.zeros(nx * (ny + 1) * nz, np.float32), "open_w": np.zeros(nx * ny * (nz + 1), np.float32),
           "floor": np.zeros(nx * ny, np.float32)}
    st = lib().orc_cut(nx, ny, nz, float(np.float32(dx)), b.ctypes.data, *(out[k].ctypes.data for k in
                                                                             ("frac", "open_u", "open_v", "open_w",
                                                                              "floor")))
    assert st == 0
    return out


_expf = None


def _c_expf(x: float) -> np.float32:
    """The C library's single-precision exp, which Rust's f32::exp calls."""
    global _expf
    if _expf is None:
        m = C.CDLL("libm.so.6")
        m.expf.restype, m.expf.argtypes = C.c_float, [C.c_float]
        _expf = m.expf
    return np.float32(_expf(float(x)))


def s385_bump(nx: int, dx) -> np.ndarray:
    """The bump of S324 off centre in y (examples/delta3d_multigrille.rs:43-47):
    0.4 + 0.6 exp(-(((x - 3)/1.2)^2 + ((y - 1.3)/0.9)^2)) at the column
    centres, f32 (exp the C library's expf)."""
    ny = nx // 2
    f32 = np.float32
    dx = f32(dx)
    out = np.empty(nx * ny, f32)
    for j in range(ny):
        y = (f32(j) + f32(0.5)) * dx
        for i in range(nx):
            x = (f32(i) + f32(0.5)) * dx
            a, b = (x - f32(3.0)) / f32(1.2), (y - f32(1.3)) / f32(0.9)
            out[j * nx + i] = f32(0.4) + f32(0.6) * _c_expf(-(a * a + b * b))
    return out


def s385_problem(nx: int, bump: bool = False) -> tuple[Problem, float, float]:
    """The reference's S385 bench case (examples/delta3d_multigrille.rs:28-36,
    58-79): an 8 x 4 x 6 m tank, surface 4 m plus a 1 cm sine along x,
    rho = 1025, g = 9.81, dt = 2 ms; the bottom flat, or (bump) the S324 bump
    cut into the grid (configure_with_bottom: solid where the volume fraction
    is 0, the cut apertures on every face). Returns (problem, scale, k1) of
    the first mobile step from rest (delta3d_mobile.rs:836-837)."""
    ny, nz = nx // 2, 3 * nx // 4
    f32 = np.float32
    dx = f32(8.0) / f32(nx)
    tau = f32(2.0 * np.pi)
    i = np.arange(nx, dtype=f32)
    # LZ + A * (TAU * (i + 0.5) * dx / LX).sin(), every operation in f32
    arg = ((tau * (i + f32(0.5))) * dx) / f32(8.0)
    s = np.array([_sinf(float(a)) for a in arg], dtype=f32)
    row = (f32(4.0) + f32(0.01) * s).astype(f32)
    eta = np.tile(row, ny).astype(f32)
    prob = Problem(nx, ny, nz, float(dx), eta, rho=1025.0, g_eff=9.81, rest=4.0, homogeneous_ghost=False)
    if bump:
        g = cut_geometry(nx, ny, nz, float(dx), s385_bump(nx, dx))
        prob.solid = (g["frac"] == 0).astype(np.uint8)
        prob.open_u, prob.open_v, prob.open_w = g["open_u"], g["open_v"], g["open_w"]
    dt = 2000 * 1e-6
    scale = float(np.float32(-1025.0 / dt))
    k1 = float(np.float32(dt / 1025.0))
    return prob, scale, k1


# ---------------------------------------------------------------- the surface step
ERR_DOMAIN, ERR_NOT_FINITE, ERR_SHAPE, ERR_CONVERGENCE = 1, 2, 3, 4
BALANCE_FIELDS = ("volume", "delta", "band_in", "perturbation_in", "sponge_out", "residual", "outgoing")


def step_surface(prob: Problem, u, v, w, p, steps, duration_us, max_iters=4000, opts=None, precision=32):
    """`steps` surface steps (delta3d_mobile.rs:819-927) from (u, v, w, p,
    prob.eta, prob.eta_roundoff). Returns a dict: status of the last step,
    steps done, per-step reports and balances, and the final fields."""
    g = prob.geometry()
    o = opts if opts is not None else options()
    f = lambda a: np.array(a, np.float32, copy=True)
    u, v, w, p = f(u), f(v), f(w), f(p)
    eta = f(prob.eta)
    ero = f(prob.eta_roundoff) if prob.eta_roundoff is not None else np.zeros_like(eta)
    reps = (Report * max(1, steps))()
    bal = np.zeros((max(1, steps), 7), np.float64)
    done = C.c_uint32()
    st = lib().orc_step_surface(precision, C.byref(g), C.byref(o), _ptr(u), _ptr(v), _ptr(w), _ptr(p), _ptr(eta),
                                _ptr(ero), steps, duration_us, max_iters, reps, _ptr(bal), C.byref(done))
    n = done.value
    return {"status": st, "done": n, "reports": [reps[q].as_dict() for q in range(min(steps, n + (st != 0)))],
            "balances": [dict(zip(BALANCE_FIELDS, bal[q])) for q in range(n)],
            "u": u, "v": v, "w": w, "p": p, "eta": eta, "eta_roundoff": ero}


def advect(prob: Problem, u, v, w, dt):
    """delta3d_mobile.rs:615-732: (us, vs, ws)."""
    g = prob.geometry()
    u, v, w = (np.ascontiguousarray(a, np.float32) for a in (u, v, w))
    us, vs, ws = np.empty_like(u), np.empty_like(v), np.empty_like(w)
    assert lib().orc_advect(C.byref(g), _ptr(u), _ptr(v), _ptr(w), dt, _ptr(us), _ptr(vs), _ptr(ws)) == 0
    return us, vs, ws


def extrapolate(prob: Problem, u, v, w):
    """delta3d_mobile.rs:345-394 on copies of (u, v, w)."""
    g = prob.geometry()
    u, v, w = (np.array(a, np.float32, copy=True) for a in (u, v, w))
    assert lib().orc_extrapolate(C.byref(g), _ptr(u), _ptr(v), _ptr(w)) == 0
    return u, v, w


def transport(prob: Problem, u, v, factor):
    """delta3d_mobile.rs:734-782: (eta, eta_roundoff, flux_x, flux_y)."""
    g = prob.geometry()
    u, v = (np.ascontiguousarray(a, np.float32) for a in (u, v))
    nx, ny = prob.nx, prob.ny
    eta, ero = np.zeros(nx * ny, np.float32), np.zeros(nx * ny, np.float32)
    fx, fy = np.zeros((nx + 1) * ny, np.float32), np.zeros(nx * (ny + 1), np.float32)
    assert lib().orc_transport(C.byref(g), _ptr(u), _ptr(v), factor, _ptr(eta), _ptr(ero), _ptr(fx), _ptr(fy)) == 0
    return eta, ero, fx, fy


def mobile_in_bounds(prob: Problem):
    return lib().orc_mobile_in_bounds(C.byref(prob.geometry())) == 1


def perturbation_volume(prob: Problem):
    return lib().orc_perturbation_volume(C.byref(prob.geometry()))


# ---------------------------------------------------------------- the APIC-3D projection
APIC_AIR, APIC_WATER, APIC_SOLID = 0, 1, 2


def apic_project(nx, ny, nz, dx, rho, dt, label, phi, u, v, w, precision=32):
    """The APIC-3D pressure projection (apic3d.rs:940-1051): returns (p f64,
    u, v, w corrected, iterations, relative residual, divergence metric)."""
    label = np.ascontiguousarray(label, np.uint8)
    phi = np.ascontiguousarray(phi, np.float32)
    u, v, w = (np.array(a, np.float32, copy=True) for a in (u, v, w))
    p = np.zeros(nx * ny * nz, np.float64)
    it, res, div = C.c_uint32(), C.c_double(), C.c_double()
    st = lib().orc_apic_project(precision, nx, ny, nz, float(np.float32(dx)), float(np.float32(rho)),
                                float(np.float32(dt)), label.ctypes.data, phi.ctypes.data, u.ctypes.data,
                                v.ctypes.data, w.ctypes.data, p.ctypes.data, C.byref(it), C.byref(res), C.byref(div))
    assert st == 0
    return p, u, v, w, it.value, res.value, div.value


def apic_apply(nx, ny, nz, label, phi, x):
    """y = A x of the APIC operator (apic3d.rs:912-935), f64."""
    label = np.ascontiguousarray(label, np.uint8)
    phi = np.ascontiguousarray(phi, np.float32)
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros_like(x)
    assert lib().orc_apic_apply(nx, ny, nz, label.ctypes.data, phi.ctypes.data, x.ctypes.data, y.ctypes.data) == 0
    return y


def apic_divergence_metric(nx, ny, nz, label, u, v, w):
    """The APIC divergence metric (apic3d.rs:1053-1080) of the faces u, v, w."""
    label = np.ascontiguousarray(label, np.uint8)
    u, v, w = (np.ascontiguousarray(a, np.float32) for a in (u, v, w))
    return lib().orc_apic_divergence_metric(nx, ny, nz, label.ctypes.data, u.ctypes.data, v.ctypes.data, w.ctypes.data)
