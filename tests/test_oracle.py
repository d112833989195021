"""Pins of the CPU oracle (oracle/hs_oracle.c) against library routines,
closed forms, invariants and brute force - never against a retyped copy of
its own formulas.

  sphbes/dsphbs  scipy.special.spherical_jn (+derivative), incl. the series
                 branches x < 1e-3 and x < 1e-4; closed forms j0, j1; the
                 cross-product identity j_l y_l' - j_l' y_l = 1/x^2
  ylm4           scipy.special.sph_harm_y; sum rule; Y_{l,-m} = (-1)^m Y*_lm;
                 addition theorem; z-axis (theta = 0) special case
  construct_ab   Legendre form of the spherical overlap (the form of Fleur's
                 hssphn.F:677-681), which uses neither Y_lm nor the A/B
                 layout; phase shift under translation
  contraction    element-wise sums (h_and_s.cpp:84-162 restated separately),
                 T = identity / T = 0 special cases, atom permutation
                 invariance, Hermiticity
"""
import numpy as np
import pytest
import scipy.special as sp

from hybridhsdla_amd import tiny_system
from oracle import pyoracle as po
from tests.helpers import rel_fro

EPS = np.finfo(float).eps
XS = np.concatenate([[0.0, 1e-6, 5e-5, 9.9e-5, 1.5e-4, 5e-4, 9.99e-4, 1.001e-3, 2e-3],
                     np.linspace(0.01, 10.0, 157), [12.0, 13.5, 15.0, 20.0]])


@pytest.mark.parametrize("lmax", [1, 2, 5, 8, 10, 12, 14])
def test_sphbes_vs_scipy(lmax):
    for x in XS:
        fj = po.sphbes(lmax, x)
        ref = sp.spherical_jn(np.arange(lmax + 1), x)
        err = np.abs(fj - ref)
        # Fleur's closed form j1 = (j0 - cos x)/x cancels for small x: its
        # absolute error is ~eps/x (and feeds the Miller normalisation when
        # j1 is the larger of j0, j1) - the reference's own accuracy, which
        # the restatement keeps.
        tol = 5e-13 * np.abs(ref) + 1e-15 + (4 * EPS / x if x > 0 else 0.0) * np.abs(ref).max()
        assert (err <= tol).all(), (lmax, x, fj, ref)


@pytest.mark.parametrize("lmax", [1, 3, 8, 12])
def test_dsphbs_vs_scipy(lmax):
    for x in XS:
        fj = po.sphbes(lmax, x)
        dfj = po.dsphbs(lmax, x, fj)
        ref = sp.spherical_jn(np.arange(lmax + 1), x, derivative=True)
        err = np.abs(dfj - ref)
        # dj_l = j_{l-1} - (l+1) j_l / x divides the j1 cancellation error by x
        tol = 5e-12 * np.abs(ref) + 2e-15 + (8 * (lmax + 1) * EPS / x ** 2 if x > 0 else 0.0)
        assert (err <= tol).all(), (lmax, x, dfj, ref)


def test_sphbes_closed_forms_and_cross_product():
    for x in np.linspace(0.05, 9.0, 40):
        fj = po.sphbes(8, x)
        assert abs(fj[0] - np.sin(x) / x) <= 1e-15
        assert abs(fj[1] - (np.sin(x) / x ** 2 - np.cos(x) / x)) <= 4 * EPS / x
        dfj = po.dsphbs(8, x, fj)
        y = sp.spherical_yn(np.arange(9), x)
        dy = sp.spherical_yn(np.arange(9), x, derivative=True)
        cross = fj * dy - dfj * y
        assert np.allclose(cross, 1.0 / x ** 2, rtol=1e-10, atol=0)


def test_sphbes_rejects_bad_input():
    with pytest.raises(ValueError):
        po.sphbes(0, 1.0)          # sphbes.f:29 'sphbes1'
    with pytest.raises(ValueError):
        po.sphbes(4, -0.5)         # sphbes.f:30 'sphbes2'


def _dirs(n, rng):
    v = rng.standard_normal((n, 3))
    return v


@pytest.mark.parametrize("lmax", [1, 4, 8, 12])
def test_ylm4_vs_scipy(lmax):
    rng = np.random.default_rng(3)
    for v in list(_dirs(25, rng)) + [np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0]),
                                     np.array([0.0, -1.0, 0.0]), np.array([0.3, 0.4, -2.0])]:
        y = po.ylm4(lmax, v)
        r = np.linalg.norm(v)
        theta = np.arccos(np.clip(v[2] / r, -1, 1))
        phi = np.arctan2(v[1], v[0])
        for l in range(lmax + 1):
            for m in range(-l, l + 1):
                ref = sp.sph_harm_y(l, m, theta, phi)
                assert abs(y[l * (l + 1) + m] - ref) <= 1e-13 * max(1.0, abs(ref)), (l, m, v)


def test_ylm4_invariants():
    rng = np.random.default_rng(5)
    lmax = 12
    for _ in range(20):
        a, b = rng.standard_normal(3), rng.standard_normal(3)
        ya, yb = po.ylm4(lmax, a), po.ylm4(lmax, b)
        cosg = a @ b / np.linalg.norm(a) / np.linalg.norm(b)
        for l in range(lmax + 1):
            blk = slice(l * l, (l + 1) ** 2)
            # sum rule
            assert abs(np.sum(np.abs(ya[blk]) ** 2) - (2 * l + 1) / (4 * np.pi)) <= 1e-13
            # addition theorem
            add = np.sum(np.conj(ya[blk]) * yb[blk])
            assert abs(add - (2 * l + 1) / (4 * np.pi) * sp.eval_legendre(l, cosg)) <= 1e-13
            # Y_{l,-m} = (-1)^m conj(Y_lm)
            for m in range(1, l + 1):
                assert abs(ya[l * (l + 1) - m] - (-1) ** m * np.conj(ya[l * (l + 1) + m])) <= 1e-15
    # z axis: only m = 0 survives, Y_l0 = sqrt((2l+1)/4pi)
    yz = po.ylm4(lmax, np.array([0.0, 0.0, 2.5]))
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            ref = np.sqrt((2 * l + 1) / (4 * np.pi)) if m == 0 else 0.0
            assert abs(yz[l * (l + 1) + m] - ref) <= 1e-14
    # r < 1e-12 takes the theta = 0 branch
    y0 = po.ylm4(lmax, np.zeros(3))
    assert np.allclose(y0, yz, atol=1e-15)


def _legendre_overlap(s, kp):
    """S_GG' = sum_a e^{i 2pi (K'-K).tau} (4pi)^2/Omega sum_l (2l+1)/4pi
    P_l(Kh.Kh') (a_l a_l' + N_l b_l b_l') / W_l^2, with j_l from scipy."""
    b33 = s.bmat33()
    K = kp.g + kp.k
    kn = np.linalg.norm(K @ b33.T, axis=1)
    S = np.zeros((kp.ng, kp.ng), dtype=complex)
    for a in range(s.natoms):
        R = s.rotmat[a].reshape(3, 3, order="F")
        kb = K @ R                      # rows: R^T K
        nb = np.linalg.norm(kb, axis=1)
        kh = kb / np.where(nb > 0, nb, 1)[:, None]
        cosg = np.clip(kh @ kh.T, -1, 1)
        ph = np.exp(2j * np.pi * (K @ s.pos[a]))
        x = s.rmt[a] * kn
        acc = np.zeros((kp.ng, kp.ng))
        for l in range(s.lmax[a] + 1):
            u, udr, ue, uedr, nrm = s.match[a, l]
            w = ue * udr - u * uedr
            j = sp.spherical_jn(l, x)
            dj = sp.spherical_jn(l, x, derivative=True)
            al = (ue * kn * dj - uedr * j) / w
            bl = (-u * kn * dj + udr * j) / w
            acc += (2 * l + 1) / (4 * np.pi) * sp.eval_legendre(l, cosg) * (
                np.outer(al, al) + nrm * np.outer(bl, bl))
        S += (4 * np.pi) ** 2 / s.omega * np.conj(ph)[:, None] * ph[None, :] * acc
    return S


@pytest.mark.parametrize("rotate", [False, True])
def test_construct_ab_legendre_overlap(rotate):
    s, kp = tiny_system(natoms=3, lmax=(3, 2, 4), ng_target=60, rotate=rotate)
    A, B = po.construct_ab(s, kp)
    L, na, G = A.shape
    Am = A.reshape(L * na, G, order="F")
    Bm = B.reshape(L * na, G, order="F")
    nrm = np.zeros((na, L))
    for a in range(na):
        for l in range(s.lmax[a] + 1):
            nrm[a, l * l:(l + 1) ** 2] = s.match[a, l, 4]
    S = Am.conj().T @ Am + Bm.conj().T @ (nrm.reshape(-1)[:, None] * Bm)
    Sref = _legendre_overlap(s, kp)
    assert rel_fro(S, Sref) <= 1e-12


def test_construct_ab_translation_phase():
    """Shifting every atom by t multiplies each column g by e^{i 2pi (k+g).t}."""
    s, kp = tiny_system(natoms=2, lmax=(3, 3), ng_target=30)
    A, B = po.construct_ab(s, kp)
    t = np.array([0.13, -0.07, 0.21])
    s.pos = s.pos + t
    A2, B2 = po.construct_ab(s, kp)
    ph = np.exp(2j * np.pi * ((kp.g + kp.k) @ t))
    assert np.allclose(A2, A * ph[None, None, :], rtol=0, atol=1e-13 * np.abs(A).max())
    assert np.allclose(B2, B * ph[None, None, :], rtol=0, atol=1e-13 * np.abs(B).max())


@pytest.fixture(scope="module")
def blas():
    if not po.blas_ready():
        pytest.skip("no host OpenBLAS")


def test_contraction_vs_naive(blas):
    s, kp = tiny_system(natoms=3, lmax=(2, 3, 1), ng_target=25)
    A, B = po.construct_ab(s, kp)
    H, S = po.create_hs2_v2(s.L, s.natoms, kp.ng, A, B, s.taa, s.tab, s.tbb, s.lmax, s.max_lmax,
                            s.sqrt_u_norms())
    Hn, Sn = po.naive_hs(s.L, s.natoms, kp.ng, A, B, s.taa, s.tab, s.tbb, s.max_lmax, s.match[:, :, 4])
    assert rel_fro(np.tril(H), np.tril(Hn)) <= 1e-13
    assert rel_fro(np.tril(S), np.tril(Sn)) <= 1e-13
    assert rel_fro(Hn, Hn.conj().T) <= 1e-14


def test_contraction_identity_t(blas):
    s, kp = tiny_system(natoms=2, lmax=(3, 2), ng_target=25)
    A, B = po.construct_ab(s, kp)
    eye = [np.eye(t.shape[0], dtype=complex) for t in s.taa]
    zero = [np.zeros_like(t) for t in s.tab]
    H, _ = po.create_hs2_v2(s.L, s.natoms, kp.ng, A, B, eye, zero, eye, s.lmax, s.max_lmax,
                            s.sqrt_u_norms())
    Am = A.reshape(-1, kp.ng, order="F")
    Bm = B.reshape(-1, kp.ng, order="F")
    ref = Am.conj().T @ Am + Bm.conj().T @ Bm
    assert rel_fro(np.tril(H), np.tril(ref)) <= 1e-13
    H0, _ = po.create_hs2_v2(s.L, s.natoms, kp.ng, A, B, [0 * t for t in s.taa], zero,
                             [0 * t for t in s.tbb], s.lmax, s.max_lmax, s.sqrt_u_norms())
    assert np.abs(np.tril(H0)).max() == 0.0


def test_contraction_atom_permutation(blas):
    s, kp = tiny_system(natoms=3, lmax=(3, 3, 3), ng_target=30)
    H, S = po.reference_hs(s, kp)
    perm = [2, 0, 1]
    s.pos = s.pos[perm]
    s.rotmat = s.rotmat[perm]
    s.match = s.match[perm]
    s.rmt = s.rmt[perm]
    s.lmax = s.lmax[perm]
    s.taa = [s.taa[i] for i in perm]
    s.tab = [s.tab[i] for i in perm]
    s.tbb = [s.tbb[i] for i in perm]
    H2, S2 = po.reference_hs(s, kp)
    assert rel_fro(np.tril(H2), np.tril(H)) <= 1e-13
    assert rel_fro(np.tril(S2), np.tril(S)) <= 1e-13
