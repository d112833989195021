"""Fleur dump reader and H/S writer (ctypes binding of include/hsgen_io.h,
libhsgen_io.so).  Host only; no GPU needed.

Python mirrors of the reference's
  FleurData           FleurData.cpp:175-458 (reader; populate_atomdata :38-68)
  HDF5Store           HDF5Store.cpp:6-47 (H/S output file)
  pack_to_full        flapw_verify.py:68-82 (Fleur's packed lower triangle)
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import HsgenKpoint, HsgenSystem
from .system import KPoint, System

IO_PATH = Path(__file__).resolve().parent / "lib" / "libhsgen_io.so"
_p = C.c_void_p


class HsgenDumpInfo(C.Structure):
    _fields_ = [
        ("num_atoms", C.c_int32), ("num_types", C.c_int32), ("max_lmax", C.c_int32),
        ("max_mesh_pts", C.c_int32), ("num_kpts", C.c_int32), ("max_gpts", C.c_int32),
        ("real_basis", C.c_double * 9), ("rec_basis", C.c_double * 9), ("omega", C.c_double),
        ("num_hs", C.c_int32), ("has_final", C.c_int32),
    ]


IO_SIGNATURES = [
    ("hsgen_io_last_error", C.c_char_p, []),
    ("hsgen_dump_open", C.c_int, [C.c_char_p, C.POINTER(_p)]),
    ("hsgen_dump_close", C.c_int, [_p]),
    ("hsgen_dump_get_info", C.c_int, [_p, C.POINTER(HsgenDumpInfo)]),
    ("hsgen_dump_get_type", C.c_int, [_p, C.c_int32, _p, _p, _p, _p, _p, _p, _p, _p]),
    ("hsgen_dump_get_atoms", C.c_int, [_p, _p, _p, _p]),
    ("hsgen_dump_get_kpoint", C.c_int, [_p, C.c_int32, _p, _p, _p]),
    ("hsgen_dump_get_system", C.c_int, [_p, C.POINTER(HsgenSystem)]),
    ("hsgen_dump_get_hs_packed", C.c_int, [_p, C.c_int32, C.POINTER(C.c_int64), _p, _p]),
    ("hsgen_unpack_lower", C.c_int, [C.c_int64, _p, _p, C.c_int64]),
    ("hsgen_dump_write", C.c_int, [C.c_char_p, C.POINTER(HsgenSystem), _p, _p, C.c_int32,
                                   C.POINTER(HsgenKpoint)]),
    ("hsgen_dump_write_hs_packed", C.c_int, [C.c_char_p, C.c_int32, C.c_int32, _p, _p, C.c_int64]),
    ("hsgen_store_open", C.c_int, [C.c_char_p, C.POINTER(_p)]),
    ("hsgen_store_add_hs", C.c_int, [_p, C.c_int32, C.c_int32, _p, _p, C.c_int64]),
    ("hsgen_store_close", C.c_int, [_p]),
]

_io = None


def load_io() -> C.CDLL:
    """Load libhsgen_io.so (after libhsgen.so, whose C++ runtime it shares)."""
    global _io
    if _io is not None:
        return _io
    _lib.load()
    if not IO_PATH.exists():
        raise RuntimeError(f"{IO_PATH} is missing: build it with `python -m hybridhsdla_amd.build` "
                           "(needs the HDF5 headers and library)")
    lib = C.CDLL(str(IO_PATH))
    for name, res, args in IO_SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _io = lib
    return lib


def _check(st: int) -> None:
    if st != 0:
        raise _lib.HsgenError(st, load_io().hsgen_io_last_error().decode())


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _complex_block(addr: int, n: int) -> np.ndarray:
    buf = (C.c_double * (2 * n * n)).from_address(addr)
    return np.frombuffer(buf, dtype=np.complex128).reshape((n, n), order="F").copy()


class FleurDump:
    """A Fleur dump (fleur_dump.h5 layout) opened for reading."""

    def __init__(self, path):
        self.path = str(path)
        lib = load_io()
        h = _p()
        _check(lib.hsgen_dump_open(self.path.encode(), C.byref(h)))
        self.handle = h
        info = HsgenDumpInfo()
        _check(lib.hsgen_dump_get_info(h, C.byref(info)))
        self.info = {name: (list(getattr(info, name)) if name.endswith("basis") else getattr(info, name))
                     for name, _ in HsgenDumpInfo._fields_}

    def close(self):
        if getattr(self, "handle", None):
            load_io().hsgen_dump_close(self.handle)
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

    # ------------------------------------------------------------------
    def type_data(self, t: int) -> dict:
        """Cutoffs, mesh, energies, potential and raw T blocks of type t (0-based)."""
        lib = load_io()
        lmax, lns, nm = C.c_int32(), C.c_int32(), C.c_int32()
        dx, rmt = C.c_double(), C.c_double()
        _check(lib.hsgen_dump_get_type(self.handle, t, C.byref(lmax), C.byref(lns), C.byref(nm), C.byref(dx),
                                       C.byref(rmt), None, None, None))
        n = (lmax.value + 1) ** 2
        E = np.zeros(lmax.value + 1)
        vr = np.zeros(nm.value)
        t4 = np.zeros(8 * n * n)
        _check(lib.hsgen_dump_get_type(self.handle, t, None, None, None, None, None, _ptr(E), _ptr(vr), _ptr(t4)))
        t4c = t4.view(np.complex128).reshape(4, n * n)
        blocks = np.stack([t4c[s].reshape((n, n), order="F") for s in range(4)])
        return dict(lmax=lmax.value, lnonsph=lns.value, mesh_n=nm.value, dx=dx.value, rmt=rmt.value,
                    E=E, vr=vr, t4=blocks)

    def atoms(self):
        """(types 0-based, positions (natoms, 3), rotations (natoms, 9) as read)."""
        na = self.info["num_atoms"]
        ty = np.zeros(na, np.int32)
        pos = np.zeros((na, 3))
        rot = np.zeros((na, 9))
        _check(load_io().hsgen_dump_get_atoms(self.handle, _ptr(ty), _ptr(pos), _ptr(rot)))
        return ty, pos, rot

    def kpoint(self, k: int) -> KPoint:
        lib = load_io()
        kp = np.zeros(3)
        ng = C.c_int32()
        _check(lib.hsgen_dump_get_kpoint(self.handle, k, _ptr(kp), C.byref(ng), None))
        g = np.zeros((ng.value, 3), np.int32)
        _check(lib.hsgen_dump_get_kpoint(self.handle, k, None, None, _ptr(g)))
        return KPoint(k=kp, g=g)

    def system(self) -> System:
        """The system description (radial solutions and final T computed by
        the library, or the stored final data)."""
        lib = load_io()
        d = HsgenSystem()
        _check(lib.hsgen_dump_get_system(self.handle, C.byref(d)))
        na, nl = d.natoms, d.max_lmax + 1

        def arr(addr, n, dt=np.float64):
            ct = C.c_int32 if dt == np.int32 else C.c_double
            return np.frombuffer((ct * n).from_address(addr), dtype=dt).copy()

        lmax = arr(d.lmax, na, np.int32)
        taa, tab, tbb = [], [], []
        for blocks, ptrs in ((taa, d.t_aa), (tab, d.t_ab), (tbb, d.t_bb)):
            pa = (C.c_void_p * na).from_address(ptrs)
            for a in range(na):
                blocks.append(_complex_block(pa[a], int((lmax[a] + 1) ** 2)))
        ty, _, _ = self.atoms()
        return System(lmax=lmax, rmt=arr(d.rmt, na), pos=arr(d.pos, 3 * na).reshape(na, 3),
                      rotmat=arr(d.rotmat, 9 * na).reshape(na, 9), match=arr(d.match, 5 * nl * na).reshape(na, nl, 5),
                      omega=float(d.omega), bmat=arr(d.bmat, 9), taa=taa, tab=tab, tbb=tbb,
                      amat=np.array(self.info["real_basis"]).reshape(3, 3, order="F"), types=ty,
                      label=Path(self.path).name)

    def fleur_hs(self, k: int) -> Tuple[np.ndarray, np.ndarray]:
        """Fleur's H and S of k-point k as column-major matrices holding the
        lower triangle (upper zero)."""
        lib = load_io()
        n = C.c_int64()
        _check(lib.hsgen_dump_get_hs_packed(self.handle, k, C.byref(n), None, None))
        m = n.value * (n.value + 1) // 2
        hp = np.zeros(m, np.complex128)
        sp = np.zeros(m, np.complex128)
        _check(lib.hsgen_dump_get_hs_packed(self.handle, k, C.byref(n), _ptr(hp), _ptr(sp)))
        return unpack_lower(hp), unpack_lower(sp)


def unpack_lower(packed: np.ndarray) -> np.ndarray:
    """Row-wise packed lower triangle -> column-major n x n (upper zero)."""
    packed = np.ascontiguousarray(packed, dtype=np.complex128)
    m = packed.size
    n = int(round((np.sqrt(8 * m + 1) - 1) / 2))
    M = np.zeros((n, n), np.complex128, order="F")
    _check(load_io().hsgen_unpack_lower(n, _ptr(packed), _ptr(M), n))
    return M


def _system_struct(system: System):
    keep = []
    s = HsgenSystem()
    s.natoms = system.natoms
    s.max_lmax = system.max_lmax
    for name, dt in (("lmax", np.int32), ("rmt", np.float64), ("pos", np.float64), ("rotmat", np.float64),
                     ("match", np.float64), ("bmat", np.float64)):
        a = np.ascontiguousarray(getattr(system, name), dtype=dt)
        keep.append(a)
        setattr(s, name, a.ctypes.data)
    s.omega = float(system.omega)
    for name, blocks in (("t_aa", system.taa), ("t_ab", system.tab), ("t_bb", system.tbb)):
        fort = [np.asfortranarray(b, dtype=np.complex128) for b in blocks]
        pa = (C.c_void_p * len(fort))(*[f.ctypes.data for f in fort])
        keep.extend(fort)
        keep.append(pa)
        setattr(s, name, C.cast(pa, C.c_void_p))
    return s, keep


def write_dump(path, system: System, kps: Sequence[KPoint]) -> None:
    """Write a system and its k-points in the dump layout (+ final data)."""
    s, keep = _system_struct(system)
    types = np.ascontiguousarray(system.types if system.types is not None else np.arange(system.natoms),
                                 dtype=np.int32)
    amat = np.asfortranarray(system.amat if system.amat is not None else
                             np.linalg.inv(system.bmat33()).T * 2 * np.pi).reshape(-1, order="F").copy()
    kst = (HsgenKpoint * max(len(kps), 1))()
    for i, kp in enumerate(kps):
        g = np.ascontiguousarray(kp.g, dtype=np.int32)
        keep.append(g)
        kst[i].k[:] = list(np.asarray(kp.k, dtype=np.float64))
        kst[i].ng = kp.ng
        kst[i].g = g.ctypes.data
    _check(load_io().hsgen_dump_write(str(path).encode(), C.byref(s), _ptr(types), _ptr(amat), len(kps), kst))


def write_hs_packed(path, k: int, H: np.ndarray, S: np.ndarray) -> None:
    """Add reference H and S of k-point k to a dump written by write_dump, in
    Fleur's packed lower-triangle layout (what FleurDump.fleur_hs reads)."""
    H = np.asfortranarray(H, dtype=np.complex128)
    S = np.asfortranarray(S, dtype=np.complex128)
    n = H.shape[0]
    if H.shape != (n, n) or S.shape != (n, n):
        raise ValueError("H and S must be square and of one size")
    _check(load_io().hsgen_dump_write_hs_packed(str(path).encode(), k, n, _ptr(H), _ptr(S), max(n, 1)))


class HSStore:
    """H/S output file (/matrix_hamilton/<k+1>, /matrix_overlap/<k+1>)."""

    def __init__(self, path):
        h = _p()
        _check(load_io().hsgen_store_open(str(path).encode(), C.byref(h)))
        self.handle = h

    def add(self, k: int, H: np.ndarray, S: np.ndarray) -> None:
        H = np.asfortranarray(H, dtype=np.complex128)
        S = np.asfortranarray(S, dtype=np.complex128)
        n = H.shape[0]
        _check(load_io().hsgen_store_add_hs(self.handle, k, n, _ptr(H), _ptr(S), n))

    def close(self) -> None:
        if getattr(self, "handle", None):
            load_io().hsgen_store_close(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
