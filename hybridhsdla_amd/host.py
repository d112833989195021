"""Host-side system set-up through the C ABI (no GPU needed): the radial
solver and the T-matrix assembly that turn a Fleur dump into an
hsgen_system.

Python mirrors of the reference's
  radial_solve       special_funcs.cpp:21-36 (-> Fleur radfun)
  TMat + augment     types_helper.cpp:41-88, AtomData.cpp:95-145
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def radial_solve(l: int, E: float, dx: float, rmt: float, vr, with_functions: bool = False):
    """{u, u', u_dot, u_dot', <u_dot|u_dot>} at R = rmt for the potential
    vr = r V(r) on the mesh r_i = rmt exp(dx (i + 1 - n)); with_functions also
    returns the (n, 2) large/small components of r u and r u_dot."""
    lib = _lib.load()
    vr = np.ascontiguousarray(vr, dtype=np.float64)
    n = vr.size
    match = np.zeros(5)
    f = np.zeros((2, n)) if with_functions else None
    g = np.zeros((2, n)) if with_functions else None
    _lib.check(lib.hsgen_radial_solve(int(l), float(E), float(dx), float(rmt), n, _ptr(vr), _ptr(match),
                                      _ptr(f) if f is not None else None,
                                      _ptr(g) if g is not None else None))
    if with_functions:
        return match, f.T, g.T
    return match


def quadrature_weights(which: int) -> np.ndarray:
    """Integer weights derived by the library (0: 5x7 first-interval Lagrange
    over 60480, 1: Newton-Cotes 7 over 140, 2: Adams-Bashforth 6 over 1440,
    3: Adams-Moulton 6 over 1440)."""
    lib = _lib.load()
    n = {0: 35, 1: 7, 2: 6, 3: 6}[which]
    w = np.zeros(n)
    _lib.check(lib.hsgen_quadrature_weights(int(which), _ptr(w)))
    return w.reshape(5, 7) if which == 0 else w


def tmat_assemble(lmax: int, lnonsph: int, t4, E, ddn):
    """Final (T_aa, T_ab, T_bb) of one atom from the dumped blocks t_uu, t_dd,
    t_ud, t_du (array (4, n, n) complex; only lower triangles are read)."""
    lib = _lib.load()
    n = (lmax + 1) ** 2
    t4 = np.asarray(t4, dtype=np.complex128)
    if t4.shape != (4, n, n):
        raise ValueError(f"t4 must be (4, {n}, {n}), got {t4.shape}")
    buf = np.ascontiguousarray(np.stack([np.asfortranarray(t4[s]).ravel(order="F") for s in range(4)]))
    E = np.ascontiguousarray(E[: lmax + 1], dtype=np.float64)
    ddn = np.ascontiguousarray(ddn[: lmax + 1], dtype=np.float64)
    out = [np.zeros((n, n), np.complex128, order="F") for _ in range(3)]
    _lib.check(lib.hsgen_tmat_assemble(int(lmax), int(lnonsph), _ptr(buf), _ptr(E), _ptr(ddn),
                                       *[_ptr(o) for o in out]))
    return tuple(out)
