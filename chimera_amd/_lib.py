"""ctypes binding of libchimera_hip.so (the C ABI declared in include/chimera_hip.h).

There is no CPU fallback: if the library is missing or a call fails the error is
raised, mapped to the exception type the reference raises for the same condition
(ValueError for bad arguments, AssertionError for the reference's asserts,
RuntimeError otherwise).
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libchimera_hip.so")

CHM_E_ARG, CHM_E_HIP, CHM_E_STATE, CHM_E_ASSERT = -1, -2, -3, -4
PREC_F64, PREC_F32 = 0, 1

_lib = None

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
vp = C.c_void_p


class Geom(C.Structure):
    _fields_ = [("natoms", C.c_int32), ("nsp", C.c_int32), ("ncoord", C.c_int32),
                ("nshell", C.c_int32), ("nblk", C.c_int32), ("nstencil", C.c_int32),
                ("nabsrec", C.c_int32), ("nkw", C.c_int32), ("nwin", C.c_int32),
                ("box", C.c_int32 * 3), ("ncell", C.c_int32 * 3), ("nbasis", C.c_int32),
                ("basis_of_parity", C.c_int32 * 8), ("has_const", C.c_int32)]


_SIGS = {
    "chm_last_error": (C.c_char_p, []),
    "chm_version": (C.c_int, []),
    "chm_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "chm_ctx_destroy": (C.c_int, [vp]),
    "chm_ctx_sync": (C.c_int, [vp]),
    "chm_fm_set_tables": (C.c_int, [vp, C.c_int, C.c_int, f64p, f64p, f64p, f64p, C.c_double]),
    "chm_fm_set_topology": (C.c_int, [vp, C.c_int, C.c_int, i32p, i32p, f64p, i32p, f64p]),
    "chm_fm_set_shift": (C.c_int, [vp, f64p]),
    "chm_fm_set_sig2": (C.c_int, [vp, f64p]),
    "chm_fm_set_cumulants": (C.c_int, [vp, C.c_int, f64p, f64p, f64p]),
    "chm_fm_distances": (C.c_int, [vp, C.c_int, f64p, f64p, vp]),
    "chm_fm_ptypes": (C.c_int, [vp, C.c_int, i32p, i32p]),
    "chm_fm_chi": (C.c_int, [vp, C.c_int, f64p, i32p, f64p, C.c_int]),
    "chm_fm_chi_noalloc": (C.c_int, [vp, C.c_int, vp, vp, vp, C.c_int, vp]),
    "chm_fm_term_rows": (C.c_int, [vp, C.c_int, f64p, f64p, f64p, i32p, i32p, i32p, C.c_double,
                                   f64p]),
    "chm_ft_create": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, i32p, i32p, f64p, f64p,
                                f64p, f64p, f64p, C.c_double, C.POINTER(vp)]),
    "chm_ft_destroy": (C.c_int, [vp]),
    "chm_ft_apply": (C.c_int, [vp, C.c_int, f64p, vp, f64p]),
    "chm_ft_apply_edge": (C.c_int, [vp, C.c_int, C.c_int, f64p, f64p]),
    "chm_ft_resample": (C.c_int, [vp, C.c_int, C.c_int, f64p, f64p]),
    "chm_sampler_create": (C.c_int, [vp, C.POINTER(Geom), C.c_int, C.c_int, i32p, i32p, i32p, vp,
                                     f64p, f64p, f64p, f64p, C.c_int, C.POINTER(vp)]),
    "chm_sampler_destroy": (C.c_int, [vp]),
    "chm_sampler_init": (C.c_int, [vp, i32p, f64p, C.c_uint64]),
    "chm_sampler_run": (C.c_int, [vp, C.c_int, C.c_int]),
    "chm_sampler_set_recording": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
    "chm_sampler_read": (C.c_int, [vp, C.c_int, vp, C.c_int64]),
    "chm_sampler_counts": (C.c_int, [vp, C.c_int, i32p, i32p]),
    "chm_sampler_chi2": (C.c_int, [vp, C.c_int, f64p, f64p]),
    "chm_fm_tab_rows": (C.c_int, [vp, C.c_int, f64p, f64p, f64p, f64p, f64p, C.c_double, f64p]),
    "chm_fm_grad_pos": (C.c_int, [vp, C.c_int, f64p, i32p, f64p, f64p]),
    "chm_fm_chi_soft": (C.c_int, [vp, C.c_int, f64p, f64p, f64p]),
    "chm_fm_grad_soft": (C.c_int, [vp, C.c_int, f64p, f64p, f64p, f64p, f64p]),
    "chm_fm_profile": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, f64p]),
    "chm_sampler_exchange": (C.c_int, [vp]),
    "chm_sampler_set_beta": (C.c_int, [vp, f64p]),
    "chm_sampler_chik": (C.c_int, [vp, C.c_int, i32p, f64p]),
    "chm_sampler_profile": (C.c_int, [vp, C.c_int, f64p]),
    "chm_sampler_set_precision": (C.c_int, [vp, C.c_int, C.c_int]),
    "chm_sampler_set_tuning": (C.c_int, [vp, C.c_int]),
    "chm_sampler_get_tuning": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "chm_sampler_set_pairs": (C.c_int, [vp, C.c_int]),
    "chm_sampler_reserve_marks": (C.c_int, [vp, C.c_int]),
    "chm_sampler_mark": (C.c_int, [vp, C.c_int]),
    "chm_sampler_read_marks": (C.c_int, [vp, C.c_int, f64p]),
    "chm_sampler_state_size": (C.c_int, [vp, C.POINTER(C.c_int64)]),
    "chm_sampler_export": (C.c_int, [vp, vp, C.c_int64]),
    "chm_sampler_import": (C.c_int, [vp, vp, C.c_int64]),
    "chm_lp_info": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                              C.POINTER(C.c_int)]),
    "chm_lp_set_params": (C.c_int, [vp, vp]),
    "chm_lp_get_params": (C.c_int, [vp, vp]),
    "chm_lp_configure": (C.c_int, [vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_int]),
    "chm_lp_set_tilt_weights": (C.c_int, [vp, vp, C.c_int]),
    "chm_lp_set_records": (C.c_int, [vp, C.c_int]),
    "chm_lp_read_records": (C.c_int, [vp, C.POINTER(C.c_int64), vp, vp, vp, C.c_int64]),
    "chm_lp_eval": (C.c_int, [vp, C.c_int, vp, vp]),
    "chm_lp_grad": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, vp]),
    "chm_lp_train": (C.c_int, [vp, C.c_int, C.c_int, C.c_float, C.c_int, C.c_float, C.c_uint64,
                               vp]),
    "chm_lp_profile": (C.c_int, [vp, C.c_int, vp]),
    "chm_psampler_create": (C.c_int, [vp, C.c_int, f64p, i32p, C.c_int, C.c_int, f64p, vp, f64p,
                                      f64p, C.c_int, C.c_int, C.POINTER(vp)]),
    "chm_psampler_destroy": (C.c_int, [vp]),
    "chm_psampler_set_moves": (C.c_int, [vp, C.c_double, C.c_double, C.c_double]),
    "chm_psampler_set_prior": (C.c_int, [vp, vp, C.c_double]),
    "chm_psampler_set_ms": (C.c_int, [vp, vp, i32p, i8p, f64p]),
    "chm_psampler_init": (C.c_int, [vp, i32p, f64p, vp, f64p, C.c_uint64]),
    "chm_psampler_run": (C.c_int, [vp, C.c_int, C.c_int]),
    "chm_psampler_exchange": (C.c_int, [vp]),
    "chm_psampler_set_beta": (C.c_int, [vp, f64p]),
    "chm_psampler_set_recording": (C.c_int, [vp, C.c_int]),
    "chm_psampler_read": (C.c_int, [vp, C.c_int, vp, C.c_int64]),
    "chm_psampler_chi2": (C.c_int, [vp, C.c_int, f64p, f64p]),
    "chm_psampler_reserve_marks": (C.c_int, [vp, C.c_int]),
    "chm_psampler_mark": (C.c_int, [vp, C.c_int]),
    "chm_psampler_read_marks": (C.c_int, [vp, C.c_int, f64p]),
    "chm_batch_create": (C.c_int, [vp, C.c_int, vp, C.POINTER(vp)]),
    "chm_batch_run": (C.c_int, [vp, C.c_int, C.c_int]),
    "chm_batch_destroy": (C.c_int, [vp]),
    "chm_batch_profile": (C.c_int, [vp, C.c_int, vp]),
    "chm_batch_set_transform": (C.c_int, [vp, C.c_int]),
    "chm_batch_get_shape": (C.c_int, [vp, C.POINTER(C.c_int)]),
}

EXPORTS = tuple(_SIGS)


def load(path=LIB_PATH):
    """Load the library once (no device work).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build() "
                           "(make -C chimera_amd/csrc); there is no CPU fallback")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc == 0:
        return
    msg = load().chm_last_error().decode(errors="replace")
    if rc == CHM_E_ARG:
        raise ValueError(msg)
    if rc == CHM_E_ASSERT:
        raise AssertionError(msg)
    raise RuntimeError(f"libchimera_hip: {msg} (code {rc})")


class Context:
    """One device + one stream (owns every device buffer of the objects on it)."""

    _default = None

    def __init__(self, device=0):
        lib = load()
        h = vp()
        check(lib.chm_ctx_create(int(device), C.byref(h)))
        self.h = h
        self.device = device

    def sync(self):
        check(load().chm_ctx_sync(self.h))

    def __del__(self, _finalizing=sys.is_finalizing):
        # bound at definition: module globals may already be gone at interpreter exit
        if _finalizing():
            return
        if getattr(self, "h", None) and _lib is not None:
            _lib.chm_ctx_destroy(self.h)
            self.h = None

    @classmethod
    def default(cls):
        if cls._default is None:
            cls._default = Context(int(os.environ.get("CHM_DEVICE", "0")))
        return cls._default


def f64(a):
    return np.ascontiguousarray(a, np.float64)


def i32(a):
    return np.ascontiguousarray(a, np.int32)
