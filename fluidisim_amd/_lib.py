"""ctypes bindings of libfsp.so, the C ABI declared in include/fsp.h and
include/fsp_learned.h.

The library is the product: loading fails loudly when it has not been built
(there is no fallback implementation).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# FSP_LIBRARY selects another build of the same library (an alternative
# compile of the same sources, e.g. for tuning); the default is the in-tree one.
LIB_PATH = os.environ.get("FSP_LIBRARY") or os.path.join(HERE, "libfsp.so")

FSP_OK, FSP_DOMAIN, FSP_NOT_FINITE, FSP_SHAPE, FSP_CONVERGENCE, FSP_DEVICE, FSP_INVALID = range(7)
PRECOND_JACOBI, PRECOND_MULTIGRID, PRECOND_LEARNED = 0, 1, 2
SMOOTH_JACOBI, SMOOTH_RBGS = 0, 1
STORE_F32, STORE_F16, STORE_BF16 = 0, 1, 2
COARSE_REDISCRETISED, COARSE_GALERKIN = 0, 1
SWEEP_FUSED, SWEEP_SINGLE = 0, 1
LABEL_AIR, LABEL_WATER, LABEL_SOLID = 0, 1, 2

# Every entry point include/fsp.h declares.
EXPORTS = [
    "fsp_abi_version", "fsp_status_string", "fsp_default_options", "fsp_create", "fsp_destroy", "fsp_levels",
    "fsp_set_geometry", "fsp_set_level_set", "fsp_set_coupling", "fsp_set_coupling_device", "fsp_refine_device",
    "fsp_enable_graded", "fsp_pressure_unknowns_per_column",
    "fsp_project", "fsp_project_device", "fsp_apply_operator_device", "fsp_assemble_device",
    "fsp_precondition_device", "fsp_divergence_device", "fsp_profile_begin", "fsp_profile_end", "fsp_generate",
    "fsp_project_batch", "fsp_autotune", "fsp_step_surface_mobile_device", "fsp_get_surface", "fsp_advect_device",
    "fsp_extrapolate_device", "fsp_transport_device", "fsp_surface_scalars",
    # include/fsp_learned.h
    "fsp_net_param_count", "fsp_net_create", "fsp_net_destroy", "fsp_net_set_params", "fsp_net_get_params",
    "fsp_net_reserve", "fsp_net_features_device", "fsp_net_forward_device", "fsp_net_forward_static_device",
    "fsp_net_loss_grad_device",
    "fsp_net_train_step_device", "fsp_attach_net", "fsp_net_mfma_probe",
]


class Domain(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("dx", C.c_float)]


class Options(C.Structure):
    _fields_ = [
        ("preconditioner", C.c_int32), ("smoother", C.c_int32), ("max_levels", C.c_int32),
        ("pre_sweeps", C.c_int32), ("post_sweeps", C.c_int32), ("coarse_sweeps", C.c_int32),
        ("gauge", C.c_int32), ("vcycle_storage", C.c_int32), ("replace_every", C.c_int32),
        ("refuse_degraded", C.c_int32), ("device", C.c_int32), ("coarse_operator", C.c_int32),
        ("graph", C.c_int32), ("learned_budget", C.c_int32), ("learned_smooth", C.c_int32),
        ("learned_vcycle", C.c_int32), ("fine_sweeps", C.c_int32),
    ]


class PatchIO(C.Structure):
    _fields_ = [("us", C.c_void_p), ("vs", C.c_void_p), ("ws", C.c_void_p), ("p", C.c_void_p),
                ("p_warm", C.c_int32), ("u", C.c_void_p), ("v", C.c_void_p), ("w", C.c_void_p),
                ("scale", C.c_float), ("k1", C.c_float)]


class Surface(C.Structure):
    _fields_ = [("rho", C.c_float), ("g_eff", C.c_float), ("rest", C.c_float),
                ("homogeneous_ghost", C.c_int32), ("closed_lid", C.c_int32)]


class Report(C.Structure):
    _fields_ = [
        ("refinements", C.c_uint32), ("iterations", C.c_uint32), ("degraded", C.c_int32), ("floor", C.c_int32),
        ("residual", C.c_double), ("divergence", C.c_double), ("backward_error", C.c_double),
        ("divergence_plain", C.c_double), ("ms_device", C.c_double), ("ms_transfer", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Balance(C.Structure):
    _fields_ = [(f, C.c_double) for f in ("volume", "delta", "band_in", "perturbation_in", "sponge_out", "residual",
                                          "outgoing")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class FspError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        super().__init__(f"{what}: {status_string(status)} ({status})")


_lib = None


def lib():
    """Load libfsp.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    # torch first: libfsp.so then binds to the HIP runtime torch already
    # loaded. Loaded the other way round, the process holds two HIP
    # runtimes, and the GPU is visible to only one of them.
    import torch  # noqa: F401
    L = C.CDLL(LIB_PATH)
    vp, sz, f32, f64, i32, u32 = C.c_void_p, C.c_size_t, C.c_float, C.c_double, C.c_int32, C.c_uint32
    L.fsp_abi_version.restype = i32
    L.fsp_status_string.argtypes = [i32]
    L.fsp_status_string.restype = C.c_char_p
    L.fsp_default_options.argtypes = [vp]
    L.fsp_create.argtypes = [vp, vp, vp]
    L.fsp_create.restype = i32
    L.fsp_destroy.argtypes = [vp]
    L.fsp_levels.argtypes = [vp]
    L.fsp_levels.restype = i32
    L.fsp_set_geometry.argtypes = [vp, vp, vp, sz, vp, vp, sz, vp, sz, vp, sz, vp, sz]
    L.fsp_set_geometry.restype = i32
    L.fsp_set_level_set.argtypes = [vp, vp, sz, vp, sz, f32]
    L.fsp_set_level_set.restype = i32
    L.fsp_set_coupling.argtypes = [vp, vp, sz, vp, sz, vp, sz, vp, sz]
    L.fsp_set_coupling.restype = i32
    L.fsp_set_coupling_device.argtypes = [vp, vp, vp, vp, vp]
    L.fsp_set_coupling_device.restype = i32
    L.fsp_refine_device.argtypes = [vp, vp, vp, vp, vp, f32, f32, f64, u32, vp]
    L.fsp_refine_device.restype = i32
    L.fsp_enable_graded.argtypes = [vp, vp, sz]
    L.fsp_enable_graded.restype = i32
    L.fsp_pressure_unknowns_per_column.argtypes = [vp]
    L.fsp_pressure_unknowns_per_column.restype = i32
    L.fsp_project.argtypes = [vp, vp, sz, vp, sz, vp, sz, vp, sz, i32, vp, sz, vp, sz, vp, sz, f32, f32, f64, u32, vp]
    L.fsp_project.restype = i32
    L.fsp_project_device.argtypes = [vp, vp, vp, vp, vp, i32, vp, vp, vp, f32, f32, f64, u32, vp]
    L.fsp_project_device.restype = i32
    L.fsp_apply_operator_device.argtypes = [vp, vp, vp]
    L.fsp_apply_operator_device.restype = i32
    L.fsp_assemble_device.argtypes = [vp, vp, f32, vp, vp]
    L.fsp_assemble_device.restype = i32
    L.fsp_precondition_device.argtypes = [vp, vp, vp]
    L.fsp_precondition_device.restype = i32
    L.fsp_divergence_device.argtypes = [vp, vp, vp, vp, vp]
    L.fsp_divergence_device.restype = i32
    L.fsp_profile_begin.argtypes = [vp, C.c_char_p, u32]
    L.fsp_profile_begin.restype = i32
    L.fsp_profile_end.argtypes = [vp, vp, vp, vp, vp, u32]
    L.fsp_profile_end.restype = i32
    L.fsp_generate.argtypes = [i32, i32, i32, i32, C.c_uint64, vp, vp, vp, vp, vp]
    L.fsp_generate.restype = i32
    L.fsp_project_batch.argtypes = [vp, vp, i32, f64, u32, i32, vp]
    L.fsp_project_batch.restype = i32
    L.fsp_autotune.argtypes = [vp, vp, vp]
    L.fsp_autotune.restype = i32
    L.fsp_step_surface_mobile_device.argtypes = [vp, vp, vp, vp, vp, C.c_uint64, u32, vp, vp]
    L.fsp_step_surface_mobile_device.restype = i32
    L.fsp_get_surface.argtypes = [vp, vp, vp]
    L.fsp_get_surface.restype = i32
    L.fsp_advect_device.argtypes = [vp, vp, vp, vp, f32, vp, vp, vp]
    L.fsp_advect_device.restype = i32
    L.fsp_extrapolate_device.argtypes = [vp, vp, vp, vp]
    L.fsp_extrapolate_device.restype = i32
    L.fsp_transport_device.argtypes = [vp, vp, vp, f32, vp, vp]
    L.fsp_transport_device.restype = i32
    L.fsp_surface_scalars.argtypes = [vp, vp, vp]
    L.fsp_surface_scalars.restype = i32
    L.fsp_net_param_count.restype = u32
    L.fsp_net_create.argtypes = [i32, vp]
    L.fsp_net_create.restype = i32
    L.fsp_net_destroy.argtypes = [vp]
    L.fsp_net_destroy.restype = None
    L.fsp_net_set_params.argtypes = [vp, vp, sz]
    L.fsp_net_set_params.restype = i32
    L.fsp_net_get_params.argtypes = [vp, vp, sz]
    L.fsp_net_get_params.restype = i32
    L.fsp_net_reserve.argtypes = [vp, i32, i32, i32]
    L.fsp_net_reserve.restype = i32
    L.fsp_net_features_device.argtypes = [vp, vp, vp, vp, f32, vp]
    L.fsp_net_features_device.restype = i32
    L.fsp_net_forward_device.argtypes = [vp, i32, i32, i32, vp, vp]
    L.fsp_net_forward_device.restype = i32
    L.fsp_net_forward_static_device.argtypes = [vp, i32, i32, i32, vp, vp]
    L.fsp_net_forward_static_device.restype = i32
    L.fsp_net_loss_grad_device.argtypes = [vp, i32, i32, i32, vp, vp, vp, vp, vp]
    L.fsp_net_loss_grad_device.restype = i32
    L.fsp_net_train_step_device.argtypes = [vp, i32, i32, i32, vp, vp, vp, f32, vp, vp]
    L.fsp_net_train_step_device.restype = i32
    L.fsp_attach_net.argtypes = [vp, vp]
    L.fsp_attach_net.restype = i32
    L.fsp_net_mfma_probe.argtypes = [vp, vp, vp]
    L.fsp_net_mfma_probe.restype = i32
    _lib = L
    return L


def status_string(status: int) -> str:
    return lib().fsp_status_string(status).decode()


def default_options(**kw) -> Options:
    o = Options()
    lib().fsp_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def check(status: int, what: str):
    if status != FSP_OK:
        raise FspError(status, what)
