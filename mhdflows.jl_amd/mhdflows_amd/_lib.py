"""ctypes binding of libmhdflows_amd.so (the C ABI in include/mhdflows_amd.h).

Loading fails loudly when the shared library is missing: there is no CPU
fallback for the hot path.
"""

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libmhdflows_amd.so")

MHD_F32, MHD_F64 = 0, 1
MHD_PHYS_HD, MHD_PHYS_MHD = 0, 1
MHD_RK4, MHD_LSRK54 = 0, 1

STATUS = {0: "ok", -1: "invalid argument", -2: "unsupported size", -3: "over HBM budget",
          -4: "HIP error", -5: "NaN detected"}


class MHDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} [{STATUS.get(code, code)}]")
        self.code = code


class Config(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
                ("nu", C.c_double), ("eta", C.c_double),
                ("n_nu", C.c_int32), ("n_eta", C.c_int32),
                ("precision", C.c_int32), ("physics", C.c_int32), ("stepper", C.c_int32),
                ("device", C.c_int32), ("stream", C.c_void_p), ("hbm_budget", C.c_uint64),
                ("dt", C.c_double)]


class Diag(C.Structure):
    _fields_ = [("E_kin", C.c_double), ("E_mag", C.c_double), ("H_kin", C.c_double),
                ("H_mag", C.c_double), ("H_cross", C.c_double), ("max_div_b", C.c_double),
                ("max_div_u", C.c_double), ("has_nan", C.c_int32), ("pad", C.c_int32)]


class KTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_int64), ("total_ms", C.c_double),
                ("bytes", C.c_double)]


# name -> (restype, argtypes)
_P = C.c_void_p
_I = C.c_int32
SIGNATURES = {
    "mhd_abi_version": (_I, []),
    "mhd_last_error": (C.c_char_p, []),
    "mhd_supported_length": (_I, [_I]),
    "mhd_query_bytes": (_I, [C.POINTER(Config), C.POINTER(C.c_uint64)]),
    "mhd_create": (_I, [C.POINTER(Config), C.POINTER(_P)]),
    "mhd_destroy": (_I, [_P]),
    "mhd_set_ic_real": (_I, [_P, C.POINTER(_P)]),
    "mhd_get_sol": (_I, [_P, _P]),
    "mhd_set_sol": (_I, [_P, _P]),
    "mhd_set_sol_modes": (_I, [_P, C.c_int64, _P, _P]),
    "mhd_batch_set_sol_modes": (_I, [_P, _I, C.c_int64, _P, _P]),
    "mhd_get_sol_modes": (_I, [_P, C.c_int64, _P, _P]),
    "mhd_set_sol_modes_noalloc": (_I, [_P, C.c_int64, _P, _P]),
    "mhd_batch_get_sol_modes": (_I, [_P, _I, C.c_int64, _P, _P]),
    "mhd_get_vars": (_I, [_P, _I, _P]),
    "mhd_device_ptrs": (_I, [_P, C.POINTER(_P), C.POINTER(_P)]),
    "mhd_get_clock": (_I, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "mhd_set_clock": (_I, [_P, C.c_double, C.c_double, C.c_int64]),
    "mhd_calcN": (_I, [_P, _P, _P, C.c_double]),
    "mhd_advection": (_I, [_P, _P, _P, C.c_double]),
    "mhd_u_update": (_I, [_P, _P, _I]),
    "mhd_b_update": (_I, [_P, _P, _I]),
    "mhd_step": (_I, [_P, _I]),
    "mhd_cfl_dt": (_I, [_P, C.c_double, C.POINTER(C.c_double)]),
    "mhd_integrate": (_I, [_P, C.c_double, C.c_int64, C.c_double, C.c_double, C.POINTER(C.c_int64)]),
    "mhd_use_graph": (_I, [_P, _I]),
    "mhd_diagnostics": (_I, [_P, C.POINTER(Diag)]),
    "mhd_refresh_vars": (_I, [_P]),
    "mhd_set_forcing": (_I, [_P, C.c_uint64, C.c_double, C.c_double]),
    "mhd_set_forcing_field": (_I, [_P, C.POINTER(_P)]),
    "mhd_set_forcing_nd": (_I, [_P, C.c_double, C.POINTER(_P)]),
    "mhd_set_forcing_a99": (_I, [_P, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]),
    "mhd_time_steps": (_I, [_P, _I, C.POINTER(C.c_double)]),
    "mhd_spectra": (_I, [_P, _P, _I, C.POINTER(_I)]),
    "mhd_scale_decomposition": (_I, [_P, _I, _I, C.c_double, C.c_double, _P]),
    "mhd_structure_function": (_I, [_P, _I, _P, _I, _P, _P]),
    "mhd_div_correction": (_I, [_P, _I]),
    "mhd_set_vp": (_I, [_P, _I, _P, _P, _P]),
    "mhd_field_op": (_I, [_P, _I, _P, _P]),
    "mhd_savefile": (_I, [_P, C.c_char_p, _I]),
    "mhd_restart": (_I, [_P, C.c_char_p, _I]),
    "mhd_autotune": (_I, [_P, _I, C.c_char_p, _I]),
    "mhd_set_variants": (_I, [_P, _P, _P]),
    "mhd_set_step_pruning": (_I, [_P, _I]),
    "mhd_batch_set_step_pruning": (_I, [_P, _I]),
    "mhd_batch_autotune": (_I, [_P, _I]),
    "mhd_profile_steps": (_I, [_P, _I, C.POINTER(KTime), _I, C.POINTER(_I)]),
    "mhd_step_bytes_model": (C.c_double, [C.POINTER(Config)]),
    "mhd_bench_copy": (_I, [C.c_uint64, _I, C.POINTER(C.c_double)]),
    "mhd_batch_create": (_I, [C.POINTER(Config), _I, C.POINTER(_P)]),
    "mhd_batch_destroy": (_I, [_P]),
    "mhd_batch_set_ic_real": (_I, [_P, _I, C.POINTER(_P)]),
    "mhd_batch_set_sol": (_I, [_P, _I, _P]),
    "mhd_batch_get_sol": (_I, [_P, _I, _P]),
    "mhd_batch_get_clock": (_I, [_P, _I, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "mhd_batch_set_cfl": (_I, [_P, _I, C.c_double]),
    "mhd_batch_step": (_I, [_P, _I]),
    "mhd_batch_use_graph": (_I, [_P, _I]),
    "mhd_batch_multi": (_I, [_P, _I, C.POINTER(_I)]),
    "mhd_set_pipeline": (_I, [_P, _I]),
    "mhd_batch_multi_classes": (_I, [_P, _I, _I, C.POINTER(_I)]),
    "mhd_batch_time_steps": (_I, [_P, _I, C.POINTER(C.c_double)]),
    "mhd_batch_set_params": (_I, [_P, _I, C.c_double, C.c_double, _I]),
    "mhd_batch_start_run": (_I, [_P, _I, C.c_double, C.c_double, C.c_int64]),
    "mhd_batch_launch": (_I, [_P, _I]),
    "mhd_batch_query": (_I, [_P, C.POINTER(_I)]),
    "mhd_batch_sync": (_I, [_P]),
}

_lib = None


def lib():
    """Load the library (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} not found: build it with __graft_entry__.build() "
                              "(the MI355X path has no CPU fallback)")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().mhd_last_error()
        raise MHDError(rc, msg.decode() if msg else "error")
    return rc
