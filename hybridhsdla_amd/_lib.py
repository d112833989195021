"""ctypes binding of the hsgen C ABI (include/hsgen.h).

The library is the in-tree hybridhsdla_amd/lib/libhsgen.so built for gfx950 by
``python -m hybridhsdla_amd.build``.  There is no fallback: if it is missing
or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libhsgen.so"

HSGEN_OK = 0
HSGEN_PREC_FP64 = 0
HSGEN_PREC_OZAKI_INT8 = 1
HSGEN_PREC_OZAKI_BF16 = 2
HSGEN_FILL_UPPER = 1
HSGEN_SKIP_H = 2
HSGEN_SKIP_S = 4

_p = C.c_void_p


class HsgenSystem(C.Structure):
    _fields_ = [
        ("natoms", C.c_int32),
        ("max_lmax", C.c_int32),
        ("lmax", _p),
        ("rmt", _p),
        ("pos", _p),
        ("rotmat", _p),
        ("match", _p),
        ("omega", C.c_double),
        ("bmat", _p),
        ("t_aa", _p),
        ("t_ab", _p),
        ("t_bb", _p),
    ]


class HsgenKpoint(C.Structure):
    _fields_ = [("k", C.c_double * 3), ("ng", C.c_int32), ("g", _p)]


class HsgenOpts(C.Structure):
    _fields_ = [
        ("precision", C.c_int32),
        ("slices", C.c_int32),
        ("target_rel_err", C.c_double),
        ("flags", C.c_int32),
        ("tile", C.c_int32),
    ]


class HsgenReport(C.Structure):
    _fields_ = [
        ("ms_total", C.c_double),
        ("ms_abgen", C.c_double),
        ("ms_tmul", C.c_double),
        ("ms_slice", C.c_double),
        ("ms_contract", C.c_double),
        ("flops_alg", C.c_double),
        ("bytes_abgen", C.c_double),
        ("ng", C.c_int32),
        ("kdim", C.c_int32),
        ("slices_used", C.c_int32),
        ("tile_used", C.c_int32),
        ("err_est_h", C.c_double),
        ("err_est_s", C.c_double),
        ("ms_igemm", C.c_double),
        ("err_smp_h", C.c_double),
        ("err_smp_s", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


# (name, restype, argtypes) of every symbol include/hsgen.h declares
SIGNATURES = [
    ("hsgen_init", C.c_int, [C.POINTER(_p), C.c_int32]),
    ("hsgen_finalize", C.c_int, [_p]),
    ("hsgen_last_error", C.c_char_p, [_p]),
    ("hsgen_status_string", C.c_char_p, [C.c_int]),
    ("hsgen_release_workspace", C.c_int, [_p]),
    ("hsgen_system_create", C.c_int, [_p, C.POINTER(HsgenSystem), C.POINTER(_p)]),
    ("hsgen_system_destroy", C.c_int, [_p, _p]),
    ("hsgen_construct_ab", C.c_int, [_p, _p, C.POINTER(HsgenKpoint), _p, _p]),
    ("hsgen_create_hs", C.c_int,
     [_p, _p, C.POINTER(HsgenKpoint), C.POINTER(HsgenOpts), _p, _p, C.c_int64,
      C.POINTER(HsgenReport)]),
    ("hsgen_create_hs_noalloc", C.c_int,
     [_p, _p, C.POINTER(HsgenKpoint), C.POINTER(HsgenOpts), _p, _p, C.c_int64,
      C.POINTER(HsgenReport)]),
    ("hsgen_create_hs_batched", C.c_int,
     [_p, _p, C.c_int32, C.POINTER(HsgenKpoint), C.POINTER(HsgenOpts), C.POINTER(_p),
      C.POINTER(_p), C.POINTER(C.c_int64), C.POINTER(HsgenReport)]),
    ("hsgen_contract_hs", C.c_int,
     [_p, C.c_int32, C.c_int32, C.c_int32, _p, _p, _p, C.POINTER(_p), C.POINTER(_p),
      C.POINTER(_p), _p, C.c_int32, _p, C.POINTER(HsgenOpts), _p, _p, C.c_int64,
      C.POINTER(HsgenReport)]),
    ("hsgen_autotune", C.c_int,
     [_p, _p, C.POINTER(HsgenKpoint), C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    ("hsgen_eigensolve", C.c_int,
     [_p, C.c_int32, _p, _p, C.c_int64, _p, _p, C.c_int64, C.POINTER(C.c_double)]),
    ("hsgen_int8_mfma_ceiling", C.c_int, [_p, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("hsgen_radial_solve", C.c_int,
     [C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int32, _p, _p, _p, _p]),
    ("hsgen_quadrature_weights", C.c_int, [C.c_int32, _p]),
    ("hsgen_tmat_assemble", C.c_int, [C.c_int32, C.c_int32, _p, _p, _p, _p, _p, _p]),
]

_lib = None


class HsgenError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"hsgen status {status}: {msg}")
        self.status = status


def load() -> C.CDLL:
    """Load libhsgen.so (raises if it is absent: there is no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m hybridhsdla_amd.build` "
            "(hipcc, gfx950); hsgen has no CPU fallback")
    lib = C.CDLL(str(LIB_PATH))
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int, ctx=None) -> None:
    if status != HSGEN_OK:
        lib = load()
        msg = lib.hsgen_last_error(ctx).decode() if ctx else lib.hsgen_status_string(status).decode()
        raise HsgenError(status, msg)
