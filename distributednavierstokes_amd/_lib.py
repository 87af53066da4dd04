"""ctypes binding of libdnsrocm.so (the C ABI in include/dns_rocm.h).

The library is built in-tree (``distributednavierstokes_amd/libdnsrocm.so``)
by ``__graft_entry__.build()``; there is no fallback: if it is missing or
cannot be loaded, every call raises.
"""
from __future__ import annotations

import ctypes as ct
import os

from .problem import DNSError

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libdnsrocm.so")

DNS_OK = 0
DNS_FP64, DNS_MIXED, DNS_PLAIN_FP64 = 0, 1, 2
LAYOUT_INTERIOR, LAYOUT_GHOSTED = 0, 1
FIELD = dict(ux=0, uy=1, uz=2, p=3, qx=4, qy=5, qz=6, delta=7)

_dp = ct.POINTER(ct.c_double)
_fp = ct.POINTER(ct.c_float)


class FVParams(ct.Structure):
    _fields_ = [
        ("n", ct.c_int32 * 3),
        ("lims", ct.c_double * 6),
        ("ywall", ct.c_int32),
        ("yfaces", _dp),
        ("visc", ct.c_double),
        ("bodyforce", ct.c_double * 3),
        ("ydiffusion", ct.c_int32),
        ("precision", ct.c_int32),
    ]


class SpecParams(ct.Structure):
    _fields_ = [
        ("n", ct.c_int32),
        ("kcut", ct.c_int32),
        ("l", ct.c_double),
        ("visc", ct.c_double),
        ("precision", ct.c_int32),
    ]


# name -> (restype, argtypes)
_SIGS = {
    "dns_last_error": (ct.c_char_p, []),
    "dns_version": (ct.c_char_p, []),
    "dns_fv_create": (ct.c_int, [ct.POINTER(FVParams), ct.POINTER(ct.c_void_p)]),
    "dns_fv_destroy": (ct.c_int, [ct.c_void_p]),
    "dns_fv_set_field": (ct.c_int, [ct.c_void_p, ct.c_int32, _dp, ct.c_int32]),
    "dns_fv_get_field": (ct.c_int, [ct.c_void_p, ct.c_int32, _dp, ct.c_int32]),
    "dns_fv_momentum": (ct.c_int, [ct.c_void_p, ct.c_double, ct.c_double]),
    "dns_fv_project": (ct.c_int, [ct.c_void_p, ct.c_int32]),
    "dns_fv_step": (ct.c_int, [ct.c_void_p, ct.c_double]),
    "dns_fv_cfl_timestep": (ct.c_int, [ct.c_void_p, ct.c_double, _dp]),
    "dns_fv_advance_cfl_noalloc": (ct.c_int, [ct.c_void_p, ct.c_double, ct.c_int32, ct.c_double, _dp, _dp,
                                              ct.POINTER(ct.c_int32)]),
    "dns_fv_describe": (ct.c_int, [ct.c_void_p, ct.c_char_p, ct.c_int32]),
    "dns_fv_advance_cfl": (ct.c_int, [ct.c_void_p, ct.c_double, ct.c_int32, ct.c_double, _dp, _dp,
                                      ct.POINTER(ct.c_int32)]),
    "dns_fv_maxdiv": (ct.c_int, [ct.c_void_p, _dp]),
    "dns_fv_kinetic_energy": (ct.c_int, [ct.c_void_p, _dp]),
    "dns_fv_channelstats_sample": (ct.c_int, [ct.c_void_p, _dp]),
    "dns_fv_spectrum_sample": (ct.c_int, [ct.c_void_p, _dp, ct.c_int32]),
    "dns_fv_sync": (ct.c_int, [ct.c_void_p]),
    "dns_fv_timing": (ct.c_int, [ct.c_void_p, ct.c_int32]),
    "dns_fv_timing_read": (ct.c_int, [ct.c_void_p, ct.c_int32, _dp, ct.POINTER(ct.c_int64)]),
    "dns_spec_create": (ct.c_int, [ct.POINTER(SpecParams), ct.POINTER(ct.c_void_p)]),
    "dns_spec_destroy": (ct.c_int, [ct.c_void_p]),
    "dns_spec_set_state": (ct.c_int, [ct.c_void_p, ct.c_int32, _dp]),
    "dns_spec_get_state": (ct.c_int, [ct.c_void_p, ct.c_int32, _dp]),
    "dns_spec_set_physical": (ct.c_int, [ct.c_void_p, _dp]),
    "dns_spec_get_physical": (ct.c_int, [ct.c_void_p, _dp]),
    "dns_spec_spec_to_phys": (ct.c_int, [ct.c_void_p]),
    "dns_spec_phys_to_spec": (ct.c_int, [ct.c_void_p]),
    "dns_spec_project": (ct.c_int, [ct.c_void_p]),
    "dns_spec_rhs": (ct.c_int, [ct.c_void_p]),
    "dns_spec_step": (ct.c_int, [ct.c_void_p, ct.c_double, _dp]),
    "dns_spec_vsummax": (ct.c_int, [ct.c_void_p, _dp]),
    "dns_spec_shell_energies": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.c_int32, _dp]),
    "dns_spec_rescale_shells": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.c_int32, _dp]),
    "dns_spec_noise": (ct.c_int, [ct.c_void_p, ct.c_uint64]),
    "dns_spec_energy": (ct.c_int, [ct.c_void_p, _dp]),
    "dns_spec_dissipation": (ct.c_int, [ct.c_void_p, _dp]),
    "dns_spec_maxdiv": (ct.c_int, [ct.c_void_p, _dp]),
    "dns_spec_save": (ct.c_int, [ct.c_void_p, ct.c_char_p, ct.c_double, ct.c_int64, ct.c_char_p]),
    "dns_spec_load": (ct.c_int, [ct.c_void_p, ct.c_char_p, _dp, ct.POINTER(ct.c_int64)]),
    "dns_spec_sync": (ct.c_int, [ct.c_void_p]),
    "dns_spec_products": (ct.c_int, [ct.c_void_p]),
    "dns_spec_get_products": (ct.c_int, [ct.c_void_p, _dp]),
    "dns_sfs_sample": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_double, ct.c_int32]),
    "dns_sfs_get": (ct.c_int, [ct.c_void_p, _dp, _dp]),
    "dns_sfs_physical": (ct.c_int, [ct.c_void_p, _dp, _dp]),
    "dns_cnn_create": (ct.c_int, [ct.c_int32, ct.c_int32, ct.c_uint64, ct.POINTER(ct.c_void_p)]),
    "dns_cnn_destroy": (ct.c_int, [ct.c_void_p]),
    "dns_cnn_num_params": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int64)]),
    "dns_cnn_set_params": (ct.c_int, [ct.c_void_p, _fp]),
    "dns_cnn_get_params": (ct.c_int, [ct.c_void_p, _fp]),
    "dns_cnn_get_grads": (ct.c_int, [ct.c_void_p, _fp]),
    "dns_cnn_forward": (ct.c_int, [ct.c_void_p, _fp, _fp]),
    "dns_cnn_loss_grad": (ct.c_int, [ct.c_void_p, _fp, _fp, _dp]),
    "dns_cnn_adam_step": (ct.c_int, [ct.c_void_p, ct.c_float, ct.c_float, ct.c_float, ct.c_float]),
    "dns_cnn_sync": (ct.c_int, [ct.c_void_p]),
    "dns_cnn_add_sample": (ct.c_int, [ct.c_void_p, _fp, _fp]),
    "dns_cnn_dataset_size": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int32)]),
    "dns_cnn_fit_scales": (ct.c_int, [ct.c_void_p, _dp, _dp]),
    "dns_cnn_set_scales": (ct.c_int, [ct.c_void_p, ct.c_double, ct.c_double]),
    "dns_cnn_get_scales": (ct.c_int, [ct.c_void_p, _dp, _dp]),
    "dns_cnn_train": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.c_float, ct.c_float, ct.c_float, ct.c_float,
                                 ct.c_uint64, ct.c_int32, _dp]),
    "dns_cnn_eval": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.c_int32, _dp]),
    "dns_cnn_predict": (ct.c_int, [ct.c_void_p, ct.c_int32, _fp]),
    "dns_cnn_timing": (ct.c_int, [ct.c_void_p, ct.c_int32]),
    "dns_cnn_timing_read": (ct.c_int, [ct.c_void_p, ct.c_int32, _dp, ct.POINTER(ct.c_int64)]),
    "dns_sfs_append": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "dns_les_closure_step": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_double, _dp]),
    "dns_spec_timing": (ct.c_int, [ct.c_void_p, ct.c_int32]),
    "dns_spec_timing_read": (ct.c_int, [ct.c_void_p, ct.c_int32, _dp, ct.POINTER(ct.c_int64)]),
}

_LIB = None


def lib():
    """Loads libdnsrocm.so once; raises DNSError if it is absent."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise DNSError(f"HIP library not built: {LIBPATH} is missing (run __graft_entry__.build())")
        L = ct.CDLL(LIBPATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def exported_symbols():
    return list(_SIGS)


def check(rc):
    if rc != DNS_OK:
        msg = lib().dns_last_error()
        raise DNSError(msg.decode() if msg else f"libdnsrocm error {rc}")
    return rc


def dptr(a):
    return a.ctypes.data_as(_dp)


def fptr(a):
    assert a.dtype == __import__("numpy").float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)
