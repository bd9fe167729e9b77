"""Host-side mirror of the reference's projection interface.

``Volume3`` keeps the names, argument meaning and error behaviour of the
reference's ``Volume3`` for the projection path
(code/water-core/src/delta3d.rs:102-294, delta3d_mobile.rs:9-25, :465-609,
:936-979): ``configure`` reserves everything, ``set_free_surface`` sets the
column heights and resets the pressure, ``set_velocity`` gives the faces,
``project_mobile3(scale, k1, max_iters)`` returns a ``Report`` or raises an
``Error``. Fields live on the GPU as torch tensors; every computation runs in
libfsp.so.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass

import numpy as np

from . import _lib


class Error(enum.IntEnum):
    """delta_projection.rs:133-150 (Domain, NotFinite, Shape, Convergence) plus device errors."""
    Domain = _lib.FSP_DOMAIN
    NotFinite = _lib.FSP_NOT_FINITE
    Shape = _lib.FSP_SHAPE
    Convergence = _lib.FSP_CONVERGENCE
    Device = _lib.FSP_DEVICE
    Invalid = _lib.FSP_INVALID


class ProjectionError(RuntimeError):
    def __init__(self, error: Error, what: str = ""):
        self.error = error
        super().__init__(f"{what}: {error.name}")


def _raise(status: int, what: str):
    if status != _lib.FSP_OK:
        raise ProjectionError(Error(status), what)


@dataclass(frozen=True)
class Domain3:
    """delta3d.rs:44-51."""
    nx: int
    ny: int
    nz: int
    dx: float

    @property
    def cells(self):
        return self.nx * self.ny * self.nz

    @property
    def columns(self):
        return self.nx * self.ny

    @property
    def faces(self):
        nx, ny, nz = self.nx, self.ny, self.nz
        return (nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)


@dataclass
class Report:
    """delta_projection.rs:164-188, plus the call's device and transfer time."""
    refinements: int
    iterations: int
    degraded: bool
    residual: float
    divergence: float
    floor: bool
    backward_error: float
    divergence_plain: float
    ms_device: float = 0.0
    ms_transfer: float = 0.0

    @staticmethod
    def from_c(r: _lib.Report) -> "Report":
        return Report(r.refinements, r.iterations, bool(r.degraded), r.residual, r.divergence, bool(r.floor),
                      r.backward_error, r.divergence_plain, r.ms_device, r.ms_transfer)


def _torch_device_first(device: int):
    """Initialises torch's state on the device before the library's first
    HIP call: torch reports no GPU when another HIP user initialised the
    runtime before it in the same process."""
    import torch
    torch.zeros(1, device=torch.device("cuda", device))


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


class Volume3:
    """The projection state of one δ domain, resident on one GPU."""

    def __init__(self, handle, domain: Domain3, rho: float, g_eff: float, options: _lib.Options):
        import torch
        self._h = handle
        self.domain = domain
        self.rho = rho
        self.g_eff = g_eff
        self.options = options
        self.rest = domain.nz * domain.dx
        self.homogeneous_ghost = False
        self.closed_lid = False
        self._solid = None
        self._open = (None, None, None)
        self._eta = np.full(domain.columns, domain.nz * domain.dx, np.float32)
        self._eta_roundoff = None
        dev = torch.device("cuda", options.device)
        nu, nv, nw = domain.faces
        f32 = torch.float32
        self.p = torch.zeros(domain.cells, dtype=f32, device=dev)
        self.u = torch.zeros(nu, dtype=f32, device=dev)
        self.v = torch.zeros(nv, dtype=f32, device=dev)
        self.w = torch.zeros(nw, dtype=f32, device=dev)
        self.us = torch.zeros(nu, dtype=f32, device=dev)
        self.vs = torch.zeros(nv, dtype=f32, device=dev)
        self.ws = torch.zeros(nw, dtype=f32, device=dev)
        self._dirty = True

    def _sync(self):
        # Inputs written on torch's stream must be complete before the
        # library's own stream reads them.
        import torch
        torch.cuda.synchronize(self.p.device)

    # ------------------------------------------------------------ configuration
    @classmethod
    def configure(cls, domain: Domain3, rho: float = 1025.0, g_eff: float = 9.81, options: _lib.Options | None = None,
                  **option_kw) -> "Volume3":
        """Volume3::configure + enable_multigrid (delta3d.rs:210-294, delta3d_multigrid.rs:339-349)."""
        if not (np.isfinite(rho) and rho > 0 and np.isfinite(g_eff)):
            raise ProjectionError(Error.NotFinite, "configure")
        o = options if options is not None else _lib.default_options()
        for k, v in option_kw.items():
            setattr(o, k, v)
        _torch_device_first(o.device)
        d = _lib.Domain(domain.nx, domain.ny, domain.nz, domain.dx)
        h = C.c_void_p()
        _raise(_lib.lib().fsp_create(C.byref(d), C.byref(o), C.byref(h)), "configure")
        return cls(h, domain, rho, g_eff, o)

    def close(self):
        if self._h:
            _lib.lib().fsp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def multigrid_levels(self) -> int | None:
        """delta3d_multigrid.rs:352-354."""
        if self.options.preconditioner != _lib.PRECOND_MULTIGRID:
            return None
        return _lib.lib().fsp_levels(self._h)

    # ------------------------------------------------------------ geometry
    def set_free_surface(self, eta, rest: float, eta_roundoff=None):
        """delta3d_mobile.rs:9-25: heights per column; the pressure restarts from 0."""
        eta = np.ascontiguousarray(eta, np.float32)
        if eta.size != self.domain.columns:
            raise ProjectionError(Error.Shape, "set_free_surface")
        if not np.isfinite(rest) or not np.all(np.isfinite(eta)):
            raise ProjectionError(Error.NotFinite, "set_free_surface")
        self._eta = eta
        self._eta_roundoff = None if eta_roundoff is None else np.ascontiguousarray(eta_roundoff, np.float32)
        self.rest = float(rest)
        self.p.zero_()
        self._dirty = True

    def set_solid(self, solid=None, open_u=None, open_v=None, open_w=None):
        """Solid cells (the reference's cut geometry with frac == 0) and optional face apertures in [0, 1]."""
        self._solid = None if solid is None else np.ascontiguousarray(solid, np.uint8)
        self._open = tuple(None if a is None else np.ascontiguousarray(a, np.float32) for a in (open_u, open_v, open_w))
        self._dirty = True

    def set_level_set(self, labels, phi, theta_min: float = 0.01):
        """The APIC-3D projection's geometry (apic3d.rs:940-1051): cell labels
        (LABEL_AIR / LABEL_WATER / LABEL_SOLID) and the level set phi at the
        cell centres; replaces the column geometry until set_solid or
        set_free_surface is called again."""
        labels = np.ascontiguousarray(labels, np.uint8)
        phi = np.ascontiguousarray(phi, np.float32)
        _raise(_lib.lib().fsp_set_level_set(self._h, _ptr(labels), labels.size, _ptr(phi), phi.size,
                                            float(theta_min)), "set_level_set")
        self._dirty = False

    def set_coupling(self, total=None, bg_up=None, bg_x=None, bg_y=None):
        """The coupled surface of the perturbation step over a background wave
        (delta3d_coupling.rs:97-189 as the projection reads it,
        delta3d_mobile.rs:92-147): total heights per column (wetness and
        ghost fractions), and the background's terms added to the up ghosts
        (per column) and to the side ghosts (per u face, per v face); None
        terms are zeros. total=None switches the coupling off."""
        self._upload_geometry()
        if total is None:
            _raise(_lib.lib().fsp_set_coupling(self._h, None, 0, None, 0, None, 0, None, 0), "set_coupling")
            return
        arr = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        total, bg_up, bg_x, bg_y = arr(total), arr(bg_up), arr(bg_x), arr(bg_y)
        sz = lambda a: 0 if a is None else a.size
        _raise(_lib.lib().fsp_set_coupling(self._h, _ptr(total), total.size, _ptr(bg_up), sz(bg_up), _ptr(bg_x),
                                           sz(bg_x), _ptr(bg_y), sz(bg_y)), "set_coupling")

    def enable_graded(self, nodes=None):
        """Volume3::enable_graded (delta3d_graded.rs:64-100): the pressure at
        the node layers only (from 0 to nz - 1, increasing), linear between
        them; None returns to the fine scheme."""
        self._upload_geometry()
        if nodes is None:
            _raise(_lib.lib().fsp_enable_graded(self._h, None, 0), "enable_graded")
            return
        arr = np.ascontiguousarray(nodes, np.int32)
        self._graded_nodes = arr
        _raise(_lib.lib().fsp_enable_graded(self._h, _ptr(arr), arr.size), "enable_graded")

    def pressure_unknowns_per_column(self) -> int:
        """delta3d_graded.rs:102-105."""
        return _lib.lib().fsp_pressure_unknowns_per_column(self._h)

    def set_surface_mode(self, homogeneous_ghost: bool, closed_lid: bool = False):
        """homogeneous_ghost: p = 0 at the surface (delta3d_coupling.rs:226-249); closed_lid: Neumann top."""
        self.homogeneous_ghost = bool(homogeneous_ghost)
        self.closed_lid = bool(closed_lid)
        self._dirty = True

    def _upload_geometry(self):
        if not self._dirty:
            return
        s = _lib.Surface(self.rho, self.g_eff, self.rest, int(self.homogeneous_ghost), int(self.closed_lid))
        ou, ov, ow = self._open
        sz = lambda a: 0 if a is None else a.size
        _raise(_lib.lib().fsp_set_geometry(
            self._h, C.byref(s), _ptr(self._eta), self._eta.size, _ptr(self._eta_roundoff),
            _ptr(self._solid), sz(self._solid), _ptr(ou), sz(ou), _ptr(ov), sz(ov), _ptr(ow), sz(ow)),
            "set_geometry")
        self._dirty = False

    # ------------------------------------------------------------ fields
    def set_velocity(self, u, v, w):
        """The predicted faces of the projection (us, vs, ws of delta3d_mobile.rs:477)."""
        import torch
        nu, nv, nw = self.domain.faces
        for src, dst, n in ((u, self.us, nu), (v, self.vs, nv), (w, self.ws, nw)):
            t = torch.as_tensor(src, dtype=torch.float32)
            if t.numel() != n:
                raise ProjectionError(Error.Shape, "set_velocity")
            dst.copy_(t.reshape(-1).to(dst.device))

    def set_step_velocity(self, u, v, w):
        """The velocity the surface step advects (u, v, w of the reference's
        Volume3, which its set_velocity sets; delta3d.rs)."""
        import torch
        nu, nv, nw = self.domain.faces
        for src, dst, n in ((u, self.u, nu), (v, self.v, nv), (w, self.w, nw)):
            t = torch.as_tensor(src, dtype=torch.float32)
            if t.numel() != n:
                raise ProjectionError(Error.Shape, "set_step_velocity")
            if not bool(torch.isfinite(t).all()):
                raise ProjectionError(Error.NotFinite, "set_step_velocity")
            dst.copy_(t.reshape(-1).to(dst.device))

    def pressure(self):
        return self.p

    def velocity_u(self):
        return self.u

    def velocity_v(self):
        return self.v

    def velocity_w(self):
        return self.w

    # ------------------------------------------------------------ the projection
    def project_mobile3(self, scale: float, k1: float, max_iters: int, rel_tol: float = 1e-6, warm: bool = True) -> Report:
        """delta3d_mobile.rs:465-609: solves, corrects u, v, w and returns the Report."""
        self._upload_geometry()
        self._sync()
        rep = _lib.Report()
        st = _lib.lib().fsp_project_device(self._h, _ptr(self.us), _ptr(self.vs), _ptr(self.ws), _ptr(self.p),
                                           int(warm), _ptr(self.u), _ptr(self.v), _ptr(self.w), scale, k1, rel_tol,
                                           max_iters, C.byref(rep))
        _raise(st, "project_mobile3")
        return Report.from_c(rep)

    def refine_mobile3(self, scale: float, k1: float, max_iters: int, rel_tol: float = 1e-6) -> Report:
        """delta3d_coupling.rs:226-249: the corrected u, v, w become the
        predicted faces, the pressure is solved again from 0 with homogeneous
        ghosts and added to the current p; refinements = 1."""
        self._upload_geometry()
        self._sync()
        rep = _lib.Report()
        st = _lib.lib().fsp_refine_device(self._h, _ptr(self.p), _ptr(self.u), _ptr(self.v), _ptr(self.w), scale, k1,
                                          rel_tol, max_iters, C.byref(rep))
        _raise(st, "refine_mobile3")
        return Report.from_c(rep)

    def project_perturbation(self, scale: float, k1: float, max_iters: int, rel_tol: float = 1e-6,
                             warm: bool = True) -> Report:
        """The projection path of step_perturbation_mobile
        (delta3d_coupling.rs:835-843): project; after a degraded result that
        stopped on the floor, refine, with the first solve's iterations added;
        a result still degraded raises Error.Convergence."""
        report = self.project_mobile3(scale, k1, max_iters, rel_tol, warm)
        if report.degraded and report.floor:
            first = report.iterations
            report = self.refine_mobile3(scale, k1, max_iters, rel_tol)
            report.iterations = min(report.iterations + first, 0xFFFFFFFF)
        if report.degraded:
            raise ProjectionError(Error.Convergence, "step_perturbation_mobile")
        return report

    # ------------------------------------------------------------ the surface step
    def step_surface_mobile(self, duration_us: int, max_iters: int) -> Report:
        """delta3d_mobile.rs:819-927: advect u, v, w, project (warm start
        from p), extrapolate, move the heights. Returns the projection's
        Report; the step's Balance3 is in self.balance. A refusal raises
        ProjectionError and leaves u, v, w, p and the surface as they were."""
        self._upload_geometry()
        self._sync()
        rep, bal = _lib.Report(), _lib.Balance()
        st = _lib.lib().fsp_step_surface_mobile_device(self._h, _ptr(self.u), _ptr(self.v), _ptr(self.w),
                                                       _ptr(self.p), int(duration_us), int(max_iters),
                                                       C.byref(rep), C.byref(bal))
        self.last_report = Report.from_c(rep)
        _raise(st, "step_surface_mobile")
        self.balance = bal.as_dict()
        self._eta, self._eta_roundoff = self.surface()
        return self.last_report

    def surface(self):
        """(eta, eta_roundoff) of the device state, host arrays."""
        self._upload_geometry()
        eta = np.empty(self.domain.columns, np.float32)
        ero = np.empty(self.domain.columns, np.float32)
        _raise(_lib.lib().fsp_get_surface(self._h, _ptr(eta), _ptr(ero)), "surface")
        return eta, ero

    def surface_scalars(self):
        """(perturbation_volume, mobile_in_bounds) of the current surface
        (delta3d_balance.rs:105-113, delta3d_mobile.rs:801-813)."""
        self._upload_geometry()
        vol, ok = C.c_double(), C.c_int32()
        _raise(_lib.lib().fsp_surface_scalars(self._h, C.byref(vol), C.byref(ok)), "surface_scalars")
        return vol.value, bool(ok.value)

    def advect_for_trials(self, u, v, w, dt: float):
        """advect_mobile3 (delta3d_mobile.rs:615-732) alone: (us, vs, ws)."""
        import torch
        self._upload_geometry()
        dev = self.p.device
        u, v, w = (torch.as_tensor(a, dtype=torch.float32, device=dev).contiguous() for a in (u, v, w))
        us, vs, ws = torch.empty_like(u), torch.empty_like(v), torch.empty_like(w)
        self._sync()
        _raise(_lib.lib().fsp_advect_device(self._h, _ptr(u), _ptr(v), _ptr(w), dt, _ptr(us), _ptr(vs), _ptr(ws)),
               "advect")
        return us, vs, ws

    def extrapolate_for_trials(self, u, v, w):
        """extrapolate_mobile3 (delta3d_mobile.rs:345-394) alone, on copies."""
        import torch
        self._upload_geometry()
        dev = self.p.device
        u, v, w = (torch.as_tensor(a, dtype=torch.float32, device=dev).clone().contiguous() for a in (u, v, w))
        self._sync()
        _raise(_lib.lib().fsp_extrapolate_device(self._h, _ptr(u), _ptr(v), _ptr(w)), "extrapolate")
        return u, v, w

    def transport_for_trials(self, u, v, factor: float):
        """transport_mobile3 (delta3d_mobile.rs:734-782) alone: moves the
        surface; returns the discharges (flux_x, flux_y)."""
        import torch
        self._upload_geometry()
        dev = self.p.device
        nx, ny = self.domain.nx, self.domain.ny
        u, v = (torch.as_tensor(a, dtype=torch.float32, device=dev).contiguous() for a in (u, v))
        fx = torch.empty((nx + 1) * ny, dtype=torch.float32, device=dev)
        fy = torch.empty(nx * (ny + 1), dtype=torch.float32, device=dev)
        self._sync()
        _raise(_lib.lib().fsp_transport_device(self._h, _ptr(u), _ptr(v), factor, _ptr(fx), _ptr(fy)), "transport")
        self._eta, self._eta_roundoff = self.surface()
        return fx, fy

    def apply_pressure_operator_for_trials(self, p):
        """delta3d_mobile.rs:936-949."""
        import torch
        self._upload_geometry()
        p = torch.as_tensor(p, dtype=torch.float32, device=self.p.device).contiguous()
        if p.numel() != self.domain.cells:
            raise ProjectionError(Error.Domain, "apply_pressure_operator_for_trials")
        if not bool(torch.isfinite(p).all()):
            raise ProjectionError(Error.NotFinite, "apply_pressure_operator_for_trials")
        out = torch.empty_like(p)
        self._sync()
        _raise(_lib.lib().fsp_apply_operator_device(self._h, _ptr(p), _ptr(out)), "apply_operator")
        return out

    def assemble_pressure_problem_for_trials(self, divergence, scale: float):
        """delta3d_mobile.rs:960-979: (rhs, prec)."""
        import torch
        self._upload_geometry()
        d = torch.as_tensor(divergence, dtype=torch.float32, device=self.p.device).contiguous()
        if d.numel() != self.domain.cells:
            raise ProjectionError(Error.Domain, "assemble_pressure_problem_for_trials")
        if not np.isfinite(scale) or not bool(torch.isfinite(d).all()):
            raise ProjectionError(Error.NotFinite, "assemble_pressure_problem_for_trials")
        rhs, prec = torch.empty_like(d), torch.empty_like(d)
        self._sync()
        _raise(_lib.lib().fsp_assemble_device(self._h, _ptr(d), scale, _ptr(rhs), _ptr(prec)), "assemble")
        return rhs, prec

    def precondition(self, r):
        """z = M^-1 r with the solver's preconditioner (v_cycle3, delta3d_multigrid.rs:393-429)."""
        import torch
        self._upload_geometry()
        r = torch.as_tensor(r, dtype=torch.float32, device=self.p.device).contiguous()
        z = torch.empty_like(r)
        self._sync()
        _raise(_lib.lib().fsp_precondition_device(self._h, _ptr(r), _ptr(z)), "precondition")
        return z

    def autotune(self):
        """Picks the fastest fine-level workgroup tile for this grid
        (fsp_autotune); returns (tile width in quads, microseconds per pass)."""
        self._upload_geometry()
        self._sync()
        tx, us = C.c_int32(), C.c_double()
        _raise(_lib.lib().fsp_autotune(self._h, C.byref(tx), C.byref(us)), "autotune")
        return tx.value, us.value

    def divergence(self, u, v, w):
        """delta3d.rs:747-763, :806-823."""
        import torch
        self._upload_geometry()
        out = torch.empty(self.domain.cells, dtype=torch.float32, device=self.p.device)
        self._sync()
        _raise(_lib.lib().fsp_divergence_device(self._h, _ptr(u.contiguous()), _ptr(v.contiguous()),
                                                _ptr(w.contiguous()), _ptr(out)), "divergence")
        return out


def project_batch(volumes, scales, k1s, max_iters: int, rel_tol: float = 1e-6, concurrency: int = 0,
                  warm: bool = False):
    """project_mobile3 of independent patches, `concurrency` at a time on their
    own streams (fsp_project_batch); returns the reports."""
    n = len(volumes)
    io = (_lib.PatchIO * max(n, 1))()
    handles = (C.c_void_p * max(n, 1))()
    for i, vol in enumerate(volumes):
        vol._upload_geometry()
        handles[i] = vol._h
        io[i] = _lib.PatchIO(_ptr(vol.us), _ptr(vol.vs), _ptr(vol.ws), _ptr(vol.p), int(warm), _ptr(vol.u),
                             _ptr(vol.v), _ptr(vol.w), float(scales[i]), float(k1s[i]))
    if n:
        volumes[0]._sync()
    reps = (_lib.Report * max(n, 1))()
    st = _lib.lib().fsp_project_batch(handles, io, n, rel_tol, max_iters, concurrency, reps)
    _raise(st, "project_batch")
    return [Report.from_c(reps[i]) for i in range(n)]


def generate(cfg: int, nx: int, ny: int, nz: int, seed: int):
    """Seeded synthetic inputs (host): (eta, solid, us, vs, ws) as numpy arrays."""
    d = Domain3(nx, ny, nz, 1.0)
    nu, nv, nw = d.faces
    eta = np.zeros(d.columns, np.float32)
    solid = np.zeros(d.cells, np.uint8)
    us, vs, ws = np.zeros(nu, np.float32), np.zeros(nv, np.float32), np.zeros(nw, np.float32)
    st = _lib.lib().fsp_generate(cfg, nx, ny, nz, seed, _ptr(eta), _ptr(solid), _ptr(us), _ptr(vs), _ptr(ws))
    _raise(st, "generate")
    return eta, solid, us, vs, ws
