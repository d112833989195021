"""System description (the reference's AtomData after generate_matchdata() and
augment_tmats(), flapw_opt/src/include/AtomData.h:16-52) and the seeded
synthetic systems the benchmarks and tests run on.

Layouts follow the C ABI (include/hsgen.h): per-atom arrays, column-major 3x3
matrices, complex128 T blocks in column-major (Fortran) order.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

TWO_PI = 2.0 * np.pi


@dataclass
class System:
    """Per-atom system description (everything construct_AB and the
    contraction read)."""
    lmax: np.ndarray            # (natoms,) int32
    rmt: np.ndarray             # (natoms,) float64
    pos: np.ndarray             # (natoms, 3) fractional positions
    rotmat: np.ndarray          # (natoms, 9) column-major R_atom * bmat
    match: np.ndarray           # (natoms, max_lmax+1, 5) u, u_dr, dE_u, dE_u_dr, dE_u_norm
    omega: float                # det(amat)
    bmat: np.ndarray            # (9,) column-major reciprocal basis (incl. 2 pi)
    taa: List[np.ndarray]       # per atom (n, n) complex128, Fortran order
    tab: List[np.ndarray]
    tbb: List[np.ndarray]
    amat: Optional[np.ndarray] = None   # (3, 3) real basis, columns = lattice vectors
    types: Optional[np.ndarray] = None  # (natoms,) type index
    label: str = ""

    @property
    def natoms(self) -> int:
        return int(self.lmax.shape[0])

    @property
    def max_lmax(self) -> int:
        return int(self.lmax.max())

    @property
    def L(self) -> int:
        return (self.max_lmax + 1) ** 2

    @property
    def nlm(self) -> np.ndarray:
        return (self.lmax.astype(np.int64) + 1) ** 2

    @property
    def kdim(self) -> int:
        return int(self.nlm.sum())

    def bmat33(self) -> np.ndarray:
        return self.bmat.reshape(3, 3, order="F")

    def sqrt_u_norms(self) -> np.ndarray:
        """(natoms, max_lmax+1) sqrt(||u_dot_l||^2), zero for l > lmax
        (h_and_s.cpp:705-711)."""
        nl = self.max_lmax + 1
        out = np.zeros((self.natoms, nl))
        for a in range(self.natoms):
            out[a, : self.lmax[a] + 1] = np.sqrt(self.match[a, : self.lmax[a] + 1, 4])
        return out

    def flops_alg(self, ng: int) -> float:
        """16 G^2 K + 32 G sum_a L_a^2 (complex MAC = 8 flops, lower triangle)."""
        nlm = self.nlm.astype(np.float64)
        return 16.0 * ng * ng * nlm.sum() + 32.0 * ng * float((nlm ** 2).sum())

    def flops_ref(self, ng: int) -> float:
        """Flops of the reference's v2 BLAS sequence, 20 G^2 K + 24 G sum L_a^2."""
        nlm = self.nlm.astype(np.float64)
        return 20.0 * ng * ng * nlm.sum() + 24.0 * ng * float((nlm ** 2).sum())


@dataclass
class KPoint:
    k: np.ndarray                 # (3,) reduced coordinates
    g: np.ndarray                 # (ng, 3) int32
    kmax: float = 0.0             # cutoff the set was built with (informational)

    @property
    def ng(self) -> int:
        return int(self.g.shape[0])


# ---------------------------------------------------------------------------
# G-vector sets
# ---------------------------------------------------------------------------
def gvectors(kpt, bmat33: np.ndarray, kmax: float) -> np.ndarray:
    """All integer g with |bmat (k+g)| < kmax, lexicographic (g1, g2, g3) order."""
    kpt = np.asarray(kpt, dtype=np.float64)
    # bound |g_i| via the smallest singular value of bmat
    smin = np.linalg.svd(bmat33, compute_uv=False).min()
    n = int(np.ceil(kmax / smin + np.abs(kpt).max())) + 1
    r = np.arange(-n, n + 1)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), axis=-1).reshape(-1, 3)
    K = (g + kpt) @ bmat33.T
    keep = np.einsum("ij,ij->i", K, K) < kmax * kmax
    return np.ascontiguousarray(g[keep].astype(np.int32))


def kmax_for_ng(kpt, bmat33: np.ndarray, target: int) -> float:
    """Smallest cutoff whose G set has at least `target` vectors (bisection on
    the sorted |k+g| list)."""
    kpt = np.asarray(kpt, dtype=np.float64)
    smin = np.linalg.svd(bmat33, compute_uv=False).min()
    vol = abs(np.linalg.det(bmat33))
    kguess = (target * vol * 3.0 / (4.0 * np.pi)) ** (1.0 / 3.0)
    n = int(np.ceil(1.3 * kguess / smin + np.abs(kpt).max())) + 2
    r = np.arange(-n, n + 1)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), axis=-1).reshape(-1, 3)
    K = (g + kpt) @ bmat33.T
    kn = np.sort(np.sqrt(np.einsum("ij,ij->i", K, K)))
    # cutoff halfway between the target-th and the next distinct radius
    k_in = kn[target - 1]
    nxt = kn[kn > k_in]
    return float(0.5 * (k_in + nxt[0])) if nxt.size else float(k_in * 1.000001)


# ---------------------------------------------------------------------------
# Synthetic systems (seeded, deterministic)
# ---------------------------------------------------------------------------
def haar_unitary(n: int, rng: np.random.Generator) -> np.ndarray:
    """Haar-distributed unitary: QR of a complex Ginibre matrix with the
    phases of diag(R) moved into Q."""
    z = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    return q * (d / np.abs(d))[None, :]


def _positions(natoms: int, a: float, dmin: float, rng: np.random.Generator) -> np.ndarray:
    pos = np.zeros((natoms, 3))
    placed = 0
    tries = 0
    while placed < natoms:
        tries += 1
        if tries > 200000:
            raise RuntimeError("could not place atoms with the requested separation")
        p = rng.random(3)
        if placed:
            d = pos[:placed] - p
            d -= np.round(d)                      # periodic minimum image (cubic)
            if (np.sqrt((d * d).sum(1)) * a).min() < dmin:
                continue
        pos[placed] = p
        placed += 1
    return pos


def _match_values(ntypes: int, nl: int, rng: np.random.Generator) -> np.ndarray:
    """Per (type, l): u~U(0.5,1.5), u'~U(-1,1), u_dot~U(-0.5,0.5),
    u_dot'~U(-1,1), ||u_dot||^2~U(0.01,0.1); draws with |W_l| < 1e-3 redrawn."""
    out = np.zeros((ntypes, nl, 5))
    for t in range(ntypes):
        for l in range(nl):
            while True:
                u = rng.uniform(0.5, 1.5)
                udr = rng.uniform(-1.0, 1.0)
                ue = rng.uniform(-0.5, 0.5)
                uedr = rng.uniform(-1.0, 1.0)
                nrm = rng.uniform(0.01, 0.1)
                w = ue * udr - u * uedr
                if abs(w) >= 1e-3:
                    break
            out[t, l] = (u, udr, ue, uedr, nrm)
    return out


def synthetic_tmat(nlm: int, rng: np.random.Generator, lam=(-2.0, 5.0)):
    """2n x 2n Hermitian Q diag(lambda) Q^H (Q Haar, lambda ~ U(lam)) split
    into the T_aa, T_ab, T_bb blocks (T_ba = T_ab^H)."""
    n2 = 2 * nlm
    q = haar_unitary(n2, rng)
    lmb = rng.uniform(lam[0], lam[1], n2)
    t = (q * lmb[None, :]) @ q.conj().T
    t = 0.5 * (t + t.conj().T)
    aa = np.asfortranarray(t[:nlm, :nlm])
    ab = np.asfortranarray(t[:nlm, nlm:])
    bb = np.asfortranarray(t[nlm:, nlm:])
    return aa, ab, bb


def make_system(a_cell: float, natoms: int, ntypes: int, type_lmax, rmt: float, seed_base: int,
                dmin: float = 3.0, label: str = "", share_t_by_type: bool = False) -> System:
    """Cubic cell a_cell, natoms atoms of ntypes round-robin types with per-type
    lmax, identity rotations, random u-values per (type, l) and per-atom T."""
    type_lmax = np.broadcast_to(np.asarray(type_lmax, dtype=np.int32), (ntypes,))
    rng_pos = np.random.default_rng(seed_base + 0)
    rng_u = np.random.default_rng(seed_base + 1)
    rng_t = np.random.default_rng(seed_base + 2)
    amat = a_cell * np.eye(3)
    bmat33 = TWO_PI / a_cell * np.eye(3)
    types = np.arange(natoms) % ntypes
    lmax = type_lmax[types].astype(np.int32)
    max_lmax = int(lmax.max())
    nl = max_lmax + 1
    pos = _positions(natoms, a_cell, dmin, rng_pos)
    mt = _match_values(ntypes, nl, rng_u)
    match = np.zeros((natoms, nl, 5))
    for a in range(natoms):
        match[a, : lmax[a] + 1] = mt[types[a], : lmax[a] + 1]
    rot = np.tile(np.asfortranarray(bmat33).reshape(-1, order="F"), (natoms, 1))
    taa, tab, tbb = [], [], []
    cache = {}
    for a in range(natoms):
        key = types[a] if share_t_by_type else a
        if key not in cache:
            cache[key] = synthetic_tmat((int(lmax[a]) + 1) ** 2, rng_t)
        aa, ab, bb = cache[key]
        taa.append(aa)
        tab.append(ab)
        tbb.append(bb)
    return System(lmax=lmax, rmt=np.full(natoms, float(rmt)), pos=pos,
                  rotmat=np.ascontiguousarray(rot), match=match,
                  omega=float(np.linalg.det(amat)),
                  bmat=np.asfortranarray(bmat33).reshape(-1, order="F").copy(),
                  taa=taa, tab=tab, tbb=tbb, amat=amat, types=types, label=label)


# Workloads (sizes of the reference's benchmark configurations).  N_G is
# honoured; the cutoff that realises it is recorded in KPoint.kmax.
CONFIGS = {
    0: dict(a_cell=10.0, natoms=4, ntypes=2, type_lmax=8, rmt=2.2, ng=600),
    1: dict(a_cell=20.0, natoms=32, ntypes=4, type_lmax=10, rmt=2.0, ng=6000),
    2: dict(a_cell=40.0, natoms=256, ntypes=8, type_lmax=12, rmt=2.2, ng=25000),
    4: dict(a_cell=24.0, natoms=48, ntypes=3, type_lmax=(8, 10, 12), rmt=2.2, ng=None),
}

KPT0 = np.array([0.1, 0.2, 0.3])


def make_config(cfg: int, ng: Optional[int] = None, natoms: Optional[int] = None):
    """Return (System, [KPoint]) for workload `cfg` (0, 1, 2 or 4; 3 is the
    cfg-2 system in split precision).  `ng`/`natoms` shrink a workload for
    sampling."""
    if cfg == 3:
        cfg = 2
    c = dict(CONFIGS[cfg])
    if natoms is not None:
        c["natoms"] = natoms
    sysm = make_system(c["a_cell"], c["natoms"], c["ntypes"], c["type_lmax"], c["rmt"],
                       seed_base=cfg * 1000, label=f"cfg{cfg}")
    b33 = sysm.bmat33()
    if cfg == 4:
        rng = np.random.default_rng(cfg * 1000 + 3)
        r = (2.0 * np.arange(1, 5) - 5.0) / 8.0
        kpts = []
        for k1 in r:
            for k2 in r:
                for k3 in r:
                    kmax = rng.uniform(2.6, 3.3)
                    k = np.array([k1, k2, k3])
                    kpts.append(KPoint(k=k, g=gvectors(k, b33, kmax), kmax=kmax))
        return sysm, kpts
    target = ng if ng is not None else c["ng"]
    kmax = kmax_for_ng(KPT0, b33, target)
    g = gvectors(KPT0, b33, kmax)
    return sysm, [KPoint(k=KPT0.copy(), g=g, kmax=kmax)]


def tiny_system(natoms: int = 2, lmax=(2, 3), a_cell: float = 6.0, rmt: float = 1.8,
                seed: int = 7, ng_target: int = 20, kpt=(0.1, 0.2, 0.3), rotate: bool = False):
    """Small system for brute-force checks; optional random rotations."""
    lmax = np.asarray(lmax, dtype=np.int32)
    ntypes = len(lmax)
    s = make_system(a_cell, natoms, ntypes, lmax, rmt, seed_base=seed, dmin=1.0)
    if rotate:
        rng = np.random.default_rng(seed + 11)
        b33 = s.bmat33()
        for a in range(natoms):
            q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
            s.rotmat[a] = (q @ b33).reshape(-1, order="F")
    k = np.asarray(kpt, dtype=np.float64)
    kmax = kmax_for_ng(k, s.bmat33(), ng_target)
    return s, KPoint(k=k, g=gvectors(k, s.bmat33(), kmax), kmax=kmax)
