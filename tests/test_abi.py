"""The C ABI library loads and exports every entry point include/fsp.h and
include/fsp_learned.h declare (CPU only: no compute call needs a device here)."""
import ctypes as C
import os
import re

from fluidisim_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for header in ("fsp.h", "fsp_learned.h"):
        text = open(os.path.join(ROOT, "include", header)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(fsp_\w+)\s*\(", text))
    return sorted(names)


def test_header_and_binding_list_agree():
    assert declared_functions() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in declared_functions():
        assert hasattr(L, name), name


def test_abi_version_and_status_strings():
    L = _lib.lib()
    assert L.fsp_abi_version() == 1
    assert _lib.status_string(_lib.FSP_SHAPE) == "shape"
    assert _lib.status_string(_lib.FSP_CONVERGENCE) == "convergence"


def test_default_options_are_the_reference_multigrid():
    """delta3d_multigrid.rs:28-34: damped Jacobi, 2 pre, 2 post, 8 coarsest sweeps."""
    o = _lib.default_options()
    assert (o.preconditioner, o.smoother, o.max_levels) == (_lib.PRECOND_MULTIGRID, _lib.SMOOTH_JACOBI, -1)
    assert (o.pre_sweeps, o.post_sweeps, o.coarse_sweeps) == (2, 2, 8)


def test_create_refuses_bad_domains_without_a_device():
    """delta3d.rs:210-214: an empty or non-finite domain is a Domain error
    (checked before any device call)."""
    L = _lib.lib()
    h = C.c_void_p()
    for d in [(0, 2, 2, 1.0), (2, 0, 2, 1.0), (2, 2, 0, 1.0), (2, 2, 2, 0.0), (2, 2, 2, float("nan"))]:
        dom = _lib.Domain(*d)
        assert L.fsp_create(C.byref(dom), None, C.byref(h)) == _lib.FSP_DOMAIN


def test_create_refuses_inconsistent_learned_options_without_a_device():
    """The network-then-V-cycle preconditioner needs f32 V-cycle storage and
    red-black smoothing (the network writes f32 into the V-cycle's iterate);
    the options are checked before any device call."""
    L = _lib.lib()
    h = C.c_void_p()
    dom = _lib.Domain(64, 8, 8, 1.0)
    base = dict(preconditioner=_lib.PRECOND_LEARNED, smoother=_lib.SMOOTH_RBGS, learned_vcycle=1)
    for bad in [dict(vcycle_storage=_lib.STORE_F16), dict(vcycle_storage=_lib.STORE_BF16),
                dict(smoother=_lib.SMOOTH_JACOBI), dict(pre_sweeps=0), dict(post_sweeps=0), dict(learned_vcycle=2),
                dict(learned_vcycle=-1)]:
        o = _lib.default_options(**{**base, **bad})
        assert L.fsp_create(C.byref(dom), C.byref(o), C.byref(h)) == _lib.FSP_INVALID, bad
    assert _lib.default_options().learned_vcycle == 0
