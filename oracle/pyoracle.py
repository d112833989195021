"""ctypes binding of the CPU oracle (oracle/hs_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to compare results with or as the timed CPU
baseline.  The product path (hybridhsdla_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import glob
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liborc.so"

_p = C.c_void_p
_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        # the oracle is plain C and cheap to build
        import subprocess
        subprocess.run(["gcc", "-O2", "-std=c99", "-fPIC", "-shared", "-o", str(LIB_PATH),
                        str(HERE / "hs_oracle.c"), "-ldl", "-lm"], check=True)
    lib = C.CDLL(str(LIB_PATH))
    lib.orc_sphbes.restype = C.c_int
    lib.orc_sphbes.argtypes = [C.c_int, C.c_double, _p]
    lib.orc_dsphbs.restype = None
    lib.orc_dsphbs.argtypes = [C.c_int, C.c_double, _p, _p]
    lib.orc_ylm4.restype = None
    lib.orc_ylm4.argtypes = [C.c_int, _p, _p]
    lib.orc_kmesh_rk.restype = None
    lib.orc_kmesh_rk.argtypes = [_p, _p, C.c_int, _p, _p]
    lib.orc_construct_ab.restype = C.c_int
    lib.orc_construct_ab.argtypes = [C.c_int, C.c_int, _p, _p, _p, _p, _p, C.c_double, _p, _p,
                                     C.c_int, _p, _p, _p]
    lib.orc_blas_load.restype = C.c_int
    lib.orc_blas_load.argtypes = [C.c_char_p, C.c_char_p]
    lib.orc_blas_set_threads.restype = C.c_int
    lib.orc_blas_set_threads.argtypes = [C.c_int]
    lib.orc_create_hs2_v2.restype = C.c_int
    lib.orc_create_hs2_v2.argtypes = [C.c_int, C.c_int, C.c_int, _p, _p, _p, _p, _p, C.c_int, _p,
                                      _p, _p, _p, _p]
    lib.orc_naive_hs.restype = None
    lib.orc_naive_hs.argtypes = [C.c_int, C.c_int, C.c_int, _p, _p, C.c_int, _p, _p, _p, _p, _p,
                                 _p]
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data


def sphbes(lmax: int, x: float) -> np.ndarray:
    fj = np.zeros(lmax + 1)
    rc = load().orc_sphbes(lmax, x, _ptr(fj))
    if rc:
        raise ValueError(f"sphbes: lmax={lmax}, x={x} rejected ({rc})")
    return fj


def dsphbs(lmax: int, x: float, fj: np.ndarray) -> np.ndarray:
    dfj = np.zeros(lmax + 1)
    fj = np.ascontiguousarray(fj, dtype=np.float64)
    load().orc_dsphbs(lmax, x, _ptr(fj), _ptr(dfj))
    return dfj


def ylm4(lmax: int, v) -> np.ndarray:
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros((lmax + 1) ** 2, dtype=np.complex128)
    load().orc_ylm4(lmax, _ptr(v), _ptr(out))
    return out


def construct_ab(system, kp):
    """Reference-layout A, B: (L, natoms, ng) complex, Fortran order."""
    s = system
    L = s.L
    A = np.zeros((L, s.natoms, kp.ng), dtype=np.complex128, order="F")
    B = np.zeros_like(A)
    lmax = np.ascontiguousarray(s.lmax, dtype=np.int32)
    rmt = np.ascontiguousarray(s.rmt, dtype=np.float64)
    pos = np.ascontiguousarray(s.pos, dtype=np.float64)
    rot = np.ascontiguousarray(s.rotmat, dtype=np.float64)
    match = np.ascontiguousarray(s.match, dtype=np.float64)
    bmat = np.ascontiguousarray(s.bmat, dtype=np.float64)
    k = np.ascontiguousarray(kp.k, dtype=np.float64)
    g = np.ascontiguousarray(kp.g, dtype=np.int32)
    rc = load().orc_construct_ab(s.natoms, s.max_lmax, _ptr(lmax), _ptr(rmt), _ptr(pos), _ptr(rot),
                                 _ptr(match), float(s.omega), _ptr(bmat), _ptr(k), kp.ng, _ptr(g),
                                 _ptr(A), _ptr(B))
    if rc:
        raise ValueError(f"construct_ab failed ({rc})")
    return A, B


def _tmat_pack(taa, tab, tbb):
    blocks = []
    sizes = []
    offs = []
    off = 0
    for aa, ab, bb in zip(taa, tab, tbb):
        n = aa.shape[0]
        sizes.append(n)
        for blk in (aa, ab, bb):
            offs.append(off)
            blocks.append(np.asfortranarray(blk, dtype=np.complex128).reshape(-1, order="F"))
            off += n * n
    return (np.ascontiguousarray(np.concatenate(blocks)), np.array(sizes, dtype=np.int32),
            np.array(offs, dtype=np.int64))


def find_openblas():
    """(path, prefix) of a host OpenBLAS: the scipy wheel's copy first."""
    try:
        import scipy
        cands = sorted(glob.glob(os.path.join(os.path.dirname(scipy.__file__) + ".libs",
                                              "libscipy_openblas*.so")))
        if cands:
            return cands[0], "scipy_"
    except ImportError:
        pass
    for pat in ("/usr/lib/x86_64-linux-gnu/libopenblas.so*", "/usr/lib/libopenblas.so*",
                "/usr/local/lib/libopenblas.so*"):
        c = sorted(glob.glob(pat))
        if c:
            return c[0], ""
    return None, None


def blas_ready(threads=None) -> bool:
    path, prefix = find_openblas()
    if path is None:
        return False
    lib = load()
    if lib.orc_blas_load(path.encode(), prefix.encode()) != 0:
        return False
    if threads:
        lib.orc_blas_set_threads(int(threads))
    return True


def create_hs2_v2(L, a, G, A, B, taa, tab, tbb, l_maxs, max_lmax, sqrt_u_norms, threads=None):
    """The reference's BLAS sequence (h_and_s_stripped_v2.cpp:100-382).  A and
    B are copied (the reference destroys its inputs).  Returns lower H, S."""
    if not blas_ready(threads):
        raise RuntimeError("no host OpenBLAS available for the oracle contraction")
    A = np.array(A.reshape(L * a, G, order="F"), dtype=np.complex128, order="F", copy=True)
    B = np.array(B.reshape(L * a, G, order="F"), dtype=np.complex128, order="F", copy=True)
    W = np.zeros_like(A)
    H = np.zeros((G, G), dtype=np.complex128, order="F")
    S = np.zeros((G, G), dtype=np.complex128, order="F")
    T_all, tsz, toff = _tmat_pack(taa, tab, tbb)
    lm = np.ascontiguousarray(l_maxs, dtype=np.int32)
    un = np.ascontiguousarray(sqrt_u_norms, dtype=np.float64).reshape(-1)
    rc = load().orc_create_hs2_v2(L, a, G, _ptr(A), _ptr(B), _ptr(H), _ptr(S), _ptr(lm), int(max_lmax),
                                  _ptr(un), _ptr(tsz), _ptr(T_all), _ptr(toff), _ptr(W))
    if rc:
        raise RuntimeError(f"oracle contraction failed ({rc})")
    return H, S


def naive_hs(L, a, G, A, B, taa, tab, tbb, max_lmax, dE_u_norm):
    """Element-wise sums (h_and_s.cpp:84-162); full H, S."""
    A = np.asfortranarray(A.reshape(L * a, G, order="F"), dtype=np.complex128)
    B = np.asfortranarray(B.reshape(L * a, G, order="F"), dtype=np.complex128)
    H = np.zeros((G, G), dtype=np.complex128, order="F")
    S = np.zeros((G, G), dtype=np.complex128, order="F")
    T_all, tsz, toff = _tmat_pack(taa, tab, tbb)
    nrm = np.ascontiguousarray(dE_u_norm, dtype=np.float64).reshape(-1)
    load().orc_naive_hs(L, a, G, _ptr(A), _ptr(B), int(max_lmax), _ptr(nrm), _ptr(tsz), _ptr(T_all),
                        _ptr(toff), _ptr(H), _ptr(S))
    return H, S


def reference_hs(system, kp, threads=None):
    """construct_AB + the v2 contraction: the reference CPU path for one k."""
    A, B = construct_ab(system, kp)
    s = system
    return create_hs2_v2(s.L, s.natoms, kp.ng, A, B, s.taa, s.tab, s.tbb, s.lmax, s.max_lmax,
                         s.sqrt_u_norms(), threads=threads)
