"""Python host interface over the hsgen C ABI.

Mirrors the reference's entry points for the FLAPW H/S path:

  construct_AB            flapw_opt/src/tensor_AB_match.cpp:63-129
  create_HS2              flapw_opt/src/h_and_s.cpp:591-598 (STRIPPED branch)
  create_HS2_stripped_v2  flapw_opt/src/h_and_s_stripped_v2.cpp:100-108

with the same argument meaning.  Every call runs on the GPU through
libhsgen.so; there is no CPU path.  Errors raise HsgenError (the reference
calls exit()/abort() instead).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import (HSGEN_FILL_UPPER, HSGEN_PREC_FP64, HSGEN_PREC_OZAKI_BF16, HSGEN_PREC_OZAKI_INT8, HSGEN_SKIP_H,
                   HSGEN_SKIP_S, HsgenError, HsgenKpoint, HsgenOpts, HsgenReport, HsgenSystem)
from .system import KPoint, System

__all__ = ["Context", "SystemHandle", "HsgenError", "HSGEN_FILL_UPPER", "HSGEN_SKIP_H",
           "HSGEN_SKIP_S", "HSGEN_PREC_FP64", "HSGEN_PREC_OZAKI_INT8", "HSGEN_PREC_OZAKI_BF16"]


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _kpoint(kp: KPoint) -> Tuple[HsgenKpoint, np.ndarray]:
    g = np.ascontiguousarray(kp.g, dtype=np.int32)
    s = HsgenKpoint()
    for i in range(3):
        s.k[i] = float(kp.k[i])
    s.ng = int(g.shape[0])
    s.g = g.ctypes.data if g.size else None
    return s, g


def _opts(precision=HSGEN_PREC_FP64, slices=0, target_rel_err=0.0, flags=0, tile=0) -> HsgenOpts:
    o = HsgenOpts()
    o.precision = precision
    o.slices = slices
    o.target_rel_err = target_rel_err
    o.flags = flags
    o.tile = tile
    return o


class SystemHandle:
    """Device-resident copy of a System (hsgen_system_create)."""

    def __init__(self, ctx: "Context", system: System):
        self.ctx = ctx
        self.system = system
        lib = _lib.load()
        self._keep = []
        s = HsgenSystem()
        s.natoms = system.natoms
        s.max_lmax = system.max_lmax
        arrs = dict(
            lmax=np.ascontiguousarray(system.lmax, dtype=np.int32),
            rmt=np.ascontiguousarray(system.rmt, dtype=np.float64),
            pos=np.ascontiguousarray(system.pos, dtype=np.float64),
            rotmat=np.ascontiguousarray(system.rotmat, dtype=np.float64),
            match=np.ascontiguousarray(system.match, dtype=np.float64),
            bmat=np.ascontiguousarray(system.bmat, dtype=np.float64),
        )
        self._keep.extend(arrs.values())
        for k, v in arrs.items():
            setattr(s, k, v.ctypes.data)
        s.omega = float(system.omega)
        ptr_arrays = []
        for blocks in (system.taa, system.tab, system.tbb):
            fort = [np.asfortranarray(b, dtype=np.complex128) for b in blocks]
            self._keep.extend(fort)
            pa = (C.c_void_p * len(fort))(*[f.ctypes.data for f in fort])
            ptr_arrays.append(pa)
        self._keep.extend(ptr_arrays)
        s.t_aa = C.cast(ptr_arrays[0], C.c_void_p)
        s.t_ab = C.cast(ptr_arrays[1], C.c_void_p)
        s.t_bb = C.cast(ptr_arrays[2], C.c_void_p)
        h = C.c_void_p()
        _lib.check(lib.hsgen_system_create(ctx.handle, C.byref(s), C.byref(h)), ctx.handle)
        self.handle = h
        self._keep = []  # the library copied everything

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().hsgen_system_destroy(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One GPU and one HIP stream (hsgen_init)."""

    def __init__(self, device: int = 0):
        lib = _lib.load()
        h = C.c_void_p()
        st = lib.hsgen_init(C.byref(h), int(device))
        if st != 0:
            raise HsgenError(st, lib.hsgen_status_string(st).decode())
        self.handle = h
        self.lib = lib

    def close(self):
        if getattr(self, "handle", None):
            self.lib.hsgen_finalize(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def release_workspace(self):
        _lib.check(self.lib.hsgen_release_workspace(self.handle), self.handle)

    def upload(self, system: System) -> SystemHandle:
        return SystemHandle(self, system)

    # -- construct_AB -----------------------------------------------------
    def construct_AB(self, sysh: SystemHandle, kp: KPoint):
        """A, B as (L, natoms, ng) complex128 arrays in the reference's
        arma::cx_cube layout (lm fastest)."""
        s = sysh.system
        L = s.L
        A = np.zeros((L, s.natoms, kp.ng), dtype=np.complex128, order="F")
        B = np.zeros_like(A)
        ks, g = _kpoint(kp)
        _lib.check(self.lib.hsgen_construct_ab(self.handle, sysh.handle, C.byref(ks), _ptr(A), _ptr(B)),
                   self.handle)
        return A, B

    # -- create_HS2 -------------------------------------------------------
    def create_HS2(self, sysh: SystemHandle, kp: KPoint, precision=HSGEN_PREC_FP64, slices=0,
                   target_rel_err=0.0, flags=0, tile=0, H=None, S=None, report=False):
        """Lower-triangle H and S (ng x ng, column-major).  With H/S given as
        preallocated arrays (numpy, or ints = device pointers with ld = ng)
        the library writes into them and leaves their upper triangles alone."""
        ng = kp.ng
        if H is None:
            H = np.zeros((ng, ng), dtype=np.complex128, order="F")
        if S is None:
            S = np.zeros((ng, ng), dtype=np.complex128, order="F")
        hp = H if isinstance(H, int) else _ptr(H)
        sp = S if isinstance(S, int) else _ptr(S)
        ks, g = _kpoint(kp)
        o = _opts(precision, slices, target_rel_err, flags, tile)
        rep = HsgenReport()
        _lib.check(self.lib.hsgen_create_hs(self.handle, sysh.handle, C.byref(ks), C.byref(o), hp, sp,
                                            max(ng, 1), C.byref(rep)), self.handle)
        if report:
            return H, S, rep.as_dict()
        return H, S

    def create_hs_device(self, sysh: SystemHandle, kp: KPoint, h_ptr: int, s_ptr: int, ld: int,
                         precision=HSGEN_PREC_FP64, slices=0, target_rel_err=0.0, flags=0, tile=0):
        ks, g = _kpoint(kp)
        o = _opts(precision, slices, target_rel_err, flags, tile)
        rep = HsgenReport()
        _lib.check(self.lib.hsgen_create_hs(self.handle, sysh.handle, C.byref(ks), C.byref(o), h_ptr,
                                            s_ptr, int(ld), C.byref(rep)), self.handle)
        return rep.as_dict()

    def create_hs_noalloc(self, sysh: SystemHandle, kp: KPoint, h_ptr: int, s_ptr: int, ld: int, flags=0,
                          tile=0):
        """hsgen_create_hs_noalloc: FP64 H and S into device pointers with no
        device allocation (the library's fixed per-GPU scratch array)."""
        ks, g = _kpoint(kp)
        o = _opts(HSGEN_PREC_FP64, 0, 0.0, flags, tile)
        rep = HsgenReport()
        _lib.check(self.lib.hsgen_create_hs_noalloc(self.handle, sysh.handle, C.byref(ks), C.byref(o), h_ptr,
                                                    s_ptr, int(ld), C.byref(rep)), self.handle)
        return rep.as_dict()

    def autotune(self, sysh: SystemHandle, kp: KPoint, precision=HSGEN_PREC_FP64):
        """Time the tile configurations on kp, remember the fastest for this
        problem size; returns (config id, contraction ms)."""
        ks, g = _kpoint(kp)
        best = C.c_int32()
        ms = C.c_double()
        _lib.check(self.lib.hsgen_autotune(self.handle, sysh.handle, C.byref(ks), int(precision), C.byref(best),
                                           C.byref(ms)), self.handle)
        return best.value, ms.value

    def eigensolve(self, H, S, vectors: bool = False, n: int = None, ld: int = None):
        """Generalized eigenproblem H x = lambda S x from lower-triangle H, S
        (numpy column-major arrays, or device pointers with n and ld given).
        Returns (w, V or None, gpu_ms); V is column-major."""
        if isinstance(H, int):
            if n is None or ld is None:
                raise ValueError("device pointers need n and ld")
            hp, sp = H, S
        else:
            H = np.asfortranarray(H, dtype=np.complex128)
            S = np.asfortranarray(S, dtype=np.complex128)
            n, ld = H.shape[0], H.shape[0]
            hp, sp = _ptr(H), _ptr(S)
        w = np.zeros(n)
        V = np.zeros((n, n), dtype=np.complex128, order="F") if vectors else None
        ms = C.c_double()
        _lib.check(self.lib.hsgen_eigensolve(self.handle, n, hp, sp, ld, _ptr(w), _ptr(V) if vectors else None,
                                             n, C.byref(ms)), self.handle)
        return w, V, ms.value

    def int8_mfma_ceiling(self, iters: int = 0):
        """Diagnostic: sustained INT8 TOP/s of the split-precision GEMMs' MFMA
        instruction with random operands and nothing else running (the
        board's power-limited ceiling for them); returns (tops, gpu_ms)."""
        tops, ms = C.c_double(), C.c_double()
        _lib.check(self.lib.hsgen_int8_mfma_ceiling(self.handle, int(iters), C.byref(tops), C.byref(ms)),
                   self.handle)
        return tops.value, ms.value

    def create_HS2_batched(self, sysh: SystemHandle, kps: Sequence[KPoint], precision=HSGEN_PREC_FP64,
                           slices=0, target_rel_err=0.0, flags=0, tile=0, outputs=None):
        """H, S for many k-points; outputs = list of (H_ptr, S_ptr, ld) device
        pointers, or None for host numpy results."""
        nk = len(kps)
        kst = (HsgenKpoint * nk)()
        keep = []
        for i, kp in enumerate(kps):
            s, g = _kpoint(kp)
            kst[i] = s
            keep.append(g)
        res = []
        if outputs is None:
            for kp in kps:
                res.append((np.zeros((kp.ng, kp.ng), np.complex128, order="F"),
                            np.zeros((kp.ng, kp.ng), np.complex128, order="F")))
            hp = (C.c_void_p * nk)(*[_ptr(h) for h, _ in res])
            sp = (C.c_void_p * nk)(*[_ptr(s) for _, s in res])
            lds = (C.c_int64 * nk)(*[max(kp.ng, 1) for kp in kps])
        else:
            hp = (C.c_void_p * nk)(*[o[0] for o in outputs])
            sp = (C.c_void_p * nk)(*[o[1] for o in outputs])
            lds = (C.c_int64 * nk)(*[o[2] for o in outputs])
        o = _opts(precision, slices, target_rel_err, flags, tile)
        reps = (HsgenReport * max(nk, 1))()
        _lib.check(self.lib.hsgen_create_hs_batched(self.handle, sysh.handle, nk, kst, C.byref(o), hp, sp,
                                                    lds, reps), self.handle)
        return res, [reps[i].as_dict() for i in range(nk)]

    # -- create_HS2_stripped_v2 --------------------------------------------
    def create_HS2_stripped_v2(self, L: int, a: int, G: int, A: np.ndarray, B: np.ndarray,
                               taa: Sequence[np.ndarray], tab: Sequence[np.ndarray],
                               tbb: Sequence[np.ndarray], l_maxs: np.ndarray, max_lmax: int,
                               u_norms: np.ndarray, precision=HSGEN_PREC_FP64, slices=0,
                               target_rel_err=0.0, flags=0, tile=0, H=None, S=None):
        """Contraction only.  A, B: (L*a, G) complex (or (L, a, G) cube),
        column-major; u_norms are sqrt(||u_dot||^2) per (atom, l) as in the
        reference.  Returns lower-triangle H, S."""
        A = np.asfortranarray(A.reshape(L * a, G, order="F"), dtype=np.complex128)
        B = np.asfortranarray(B.reshape(L * a, G, order="F"), dtype=np.complex128)
        fa = [np.asfortranarray(t, dtype=np.complex128) for t in taa]
        fb = [np.asfortranarray(t, dtype=np.complex128) for t in tab]
        fc = [np.asfortranarray(t, dtype=np.complex128) for t in tbb]
        tsz = np.array([t.shape[0] for t in fa], dtype=np.int32)
        pa = (C.c_void_p * a)(*[t.ctypes.data for t in fa])
        pb = (C.c_void_p * a)(*[t.ctypes.data for t in fb])
        pc = (C.c_void_p * a)(*[t.ctypes.data for t in fc])
        lm = np.ascontiguousarray(l_maxs, dtype=np.int32)
        un = np.ascontiguousarray(u_norms, dtype=np.float64).reshape(-1)
        if H is None:
            H = np.zeros((G, G), dtype=np.complex128, order="F")
        if S is None:
            S = np.zeros((G, G), dtype=np.complex128, order="F")
        o = _opts(precision, slices, target_rel_err, flags, tile)
        rep = HsgenReport()
        _lib.check(self.lib.hsgen_contract_hs(self.handle, L, a, G, _ptr(A), _ptr(B), _ptr(tsz), pa, pb,
                                              pc, _ptr(lm), int(max_lmax), _ptr(un), C.byref(o), _ptr(H),
                                              _ptr(S), max(G, 1), C.byref(rep)), self.handle)
        return H, S
