"""ctypes mirror of include/ocpx.h and the loader of the HIP library.

The structures here are byte-for-byte the C structures of the C ABI; the
library itself (libocpx.so, HIP kernels + C++ host code) is built by
opencaeporox_amd/csrc/build.py.  There is no CPU fallback: loading fails
loudly when the library is missing.
"""
import ctypes as C
import os

import numpy as np

MAX_NC = 16
MAX_TABROWS = 64

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libocpx.so")


class Fluid(C.Structure):
    _fields_ = [
        ("nc", C.c_int32), ("lbc_exact", C.c_int32),
        ("Tc", C.c_double * MAX_NC), ("Pc", C.c_double * MAX_NC),
        ("Vc", C.c_double * MAX_NC), ("Vcvis", C.c_double * MAX_NC),
        ("MW", C.c_double * MAX_NC), ("acf", C.c_double * MAX_NC),
        ("OmegaA", C.c_double * MAX_NC), ("OmegaB", C.c_double * MAX_NC),
        ("Vshift", C.c_double * MAX_NC),
        ("BIC", C.c_double * (MAX_NC * MAX_NC)),
        ("LBCcoef", C.c_double * 5), ("PVTW", C.c_double * 5),
        ("rhoW_std", C.c_double),
        ("RR_maxIt", C.c_int32), ("SSMsta_maxIt", C.c_int32), ("NRsta_maxIt", C.c_int32),
        ("SSMsp_maxIt", C.c_int32), ("NRsp_maxIt", C.c_int32), ("skipPSA", C.c_int32),
        ("RR_tol", C.c_double), ("SSMsta_tol", C.c_double), ("SSMsta_eYt", C.c_double),
        ("NRsta_tol", C.c_double), ("SSMsp_tol", C.c_double), ("NRsp_tol", C.c_double),
    ]


class Sat(C.Structure):
    _fields_ = [("nswof", C.c_int32), ("nsgof", C.c_int32),
                ("swof", C.c_double * (4 * MAX_TABROWS)),
                ("sgof", C.c_double * (4 * MAX_TABROWS))]


class Rock(C.Structure):
    _fields_ = [("Pref", C.c_double), ("cp1", C.c_double), ("cp2", C.c_double)]


class Grid(C.Structure):
    _fields_ = [("nb", C.c_int64), ("nconn", C.c_int64),
                ("v", C.c_void_p), ("depth", C.c_void_p), ("poroInit", C.c_void_p),
                ("conn_b", C.c_void_p), ("conn_e", C.c_void_p), ("trans", C.c_void_p)]


class Well(C.Structure):
    _fields_ = [("type", C.c_int32), ("mode", C.c_int32), ("perf0", C.c_int32), ("nperf", C.c_int32),
                ("bhp", C.c_double), ("target", C.c_double),
                ("zi", C.c_double * (MAX_NC + 1)), ("factor", C.c_double * (MAX_NC + 1))]


class ResNorms(C.Structure):
    _fields_ = [("maxRelRes_V", C.c_double), ("maxId_V", C.c_int64),
                ("maxRelRes_N", C.c_double), ("maxId_N", C.c_int64)]


# enum ocpx_field, in order
FIELDS = ["P", "T", "NI", "LNI", "PHASEEXIST", "S", "XIJ", "RHO", "XI", "MU", "RHOP",
          "XIP", "MUP", "RHOX", "XIX", "MUX", "NT", "VF", "VFP", "VFI", "DSEC_DPRI", "KR",
          "PC", "PJ", "DKRDS", "DPCDS", "PORO", "POROP", "ROCKVP", "FTYPE_IN", "FTYPE_OUT",
          "SKIP_FLAG", "SKIP_MINEIG", "SKIP_P", "SKIP_T", "SKIP_ZI", "ITERS"]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
INT_FIELDS = {"PHASEEXIST": np.uint32, "FTYPE_IN": np.int32, "FTYPE_OUT": np.int32,
              "SKIP_FLAG": np.uint32, "ITERS": np.int32}


def field_len(name, Nc):
    """Entries per cell of a field (mirrors ocpx_field_len)."""
    nc, np_ = Nc + 1, 3
    table = {"NI": nc, "LNI": nc, "PHASEEXIST": np_, "S": np_, "XIJ": np_ * nc,
             "RHO": np_, "XI": np_, "MU": np_, "RHOP": np_, "XIP": np_, "MUP": np_,
             "RHOX": np_ * nc, "XIX": np_ * nc, "MUX": np_ * nc, "VFI": nc,
             "DSEC_DPRI": np_ * (nc + 1) * (nc + 1), "KR": np_, "PC": np_, "PJ": np_,
             "DKRDS": np_ * np_, "DPCDS": np_ * np_, "SKIP_ZI": Nc, "ITERS": 8}
    return table.get(name, 1)


_lib = None


def load(path=None):
    """Load libocpx.so; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libocpx.so not built ({p}); run __graft_entry__.build()")
    lib = C.CDLL(p)
    vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    sig = {
        "ocpx_field_len": (i64, [C.c_int, i32]),
        "ocpx_ctx_create": (C.c_int, [C.POINTER(Fluid), C.POINTER(Sat), C.POINTER(Rock),
                                      C.POINTER(Grid), vp, C.POINTER(vp)]),
        "ocpx_ctx_create_multi": (C.c_int, [i32, vp, vp, vp, vp, vp, C.POINTER(vp)]),
        "ocpx_ctx_destroy": (None, [vp]),
        "ocpx_last_error": (C.c_char_p, [vp]),
        "ocpx_sync": (C.c_int, [vp]),
        "ocpx_set_field": (C.c_int, [vp, C.c_int, vp]),
        "ocpx_get_field": (C.c_int, [vp, C.c_int, vp]),
        "ocpx_field_device_ptr": (C.c_int, [vp, C.c_int, C.POINTER(vp)]),
        "ocpx_flash_der": (C.c_int, [vp, C.c_int]),
        "ocpx_phase_equilibrium": (C.c_int, [vp, C.c_int]),
        "ocpx_get_pe_result": (C.c_int, [vp, vp, vp, vp]),
        "ocpx_krpc_der": (C.c_int, [vp]),
        "ocpx_cal_rock": (C.c_int, [vp]),
        "ocpx_cal_res": (C.c_int, [vp, dbl, vp, C.POINTER(ResNorms)]),
        "ocpx_assemble": (C.c_int, [vp, dbl]),
        "ocpx_get_bsr": (C.c_int, [vp, C.POINTER(i64), vp, vp, vp, vp]),
        "ocpx_bsr_device_ptrs": (C.c_int, [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp),
                                           C.POINTER(vp)]),
        "ocpx_get_solution": (C.c_int, [vp, vp, dbl]),
        "ocpx_newton_pass": (C.c_int, [vp, dbl, C.c_int, C.c_int]),
        "ocpx_update_last_step": (C.c_int, [vp]),
        "ocpx_set_flash_mapping": (C.c_int, [vp, C.c_int]),
        "ocpx_set_geometry": (C.c_int, [vp, vp, vp]),
        "ocpx_get_flash_mapping": (C.c_int, [vp, C.POINTER(i32)]),
        "ocpx_autotune_flash": (C.c_int, [vp, vp]),
        "ocpx_reset_to_last_step": (C.c_int, [vp]),
        "ocpx_jt_lambda": (C.c_int, [vp, vp, vp]),
        "ocpx_dR_dparams": (C.c_int, [vp, dbl, vp, vp, vp]),
        "ocpx_jt_lambda_dev": (C.c_int, [vp, vp, vp]),
        "ocpx_dR_dparams_dev": (C.c_int, [vp, dbl, vp, vp, vp]),
        "ocpx_linear_solve": (C.c_int, [vp, dbl, i32, vp, C.POINTER(i32), C.POINTER(dbl)]),
        "ocpx_set_flash_queue_order": (C.c_int, [vp, C.c_int]),
        "ocpx_get_flash_queue_order": (C.c_int, [vp, C.POINTER(i32)]),
        "ocpx_set_flash_stages": (C.c_int, [vp, C.c_int]),
        "ocpx_set_linear_solver": (C.c_int, [vp, C.c_int]),
        "ocpx_set_flash_long": (C.c_int, [vp, i32]),
        "ocpx_get_flash_long": (C.c_int, [vp, C.POINTER(i32)]),
        "ocpx_batch_create": (C.c_int, [C.POINTER(vp), i32, vp, C.POINTER(vp)]),
        "ocpx_batch_newton_pass": (C.c_int, [vp, dbl]),
        "ocpx_batch_destroy": (None, [vp]),
        "ocpx_get_linear_solver": (C.c_int, [vp, C.POINTER(i32)]),
        "ocpx_phase_equilibrium_noalloc": (C.c_int, [C.POINTER(Fluid), i64, vp, vp, vp, vp, vp, vp, vp, vp]),
        "ocpx_get_flash_stages": (C.c_int, [vp, C.POINTER(i32)]),
        "ocpx_timing_enable": (C.c_int, [vp, C.c_int]),
        "ocpx_timing_get": (C.c_int, [vp, C.c_int, C.POINTER(dbl), C.POINTER(i64)]),
        "ocpx_set_wells": (C.c_int, [vp, i32, vp, vp, vp, vp, vp]),
        "ocpx_well_terms": (C.c_int, [vp, dbl, i32, C.POINTER(ResNorms), vp]),
        "ocpx_get_well_blocks": (C.c_int, [vp, vp, vp, vp]),
        "ocpx_get_wells": (C.c_int, [vp, vp, vp]),
        "ocpx_set_well_bhp": (C.c_int, [vp, vp]),
        "ocpx_peaceman_wi": (dbl, [i32, dbl, dbl, dbl, dbl, dbl, dbl, dbl, dbl, dbl]),
        "ocpx_xi_phase": (dbl, [C.POINTER(Fluid), dbl, dbl, vp, i32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib if path is None else _lib
    return lib


def declared_symbols():
    """Names of the entry points declared in include/ocpx.h."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "ocpx.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"\b(ocpx_[a-z_A-Z0-9]+)\s*\(", txt)))
