"""The generalized eigenproblem H x = lambda S x after H/S generation
(rocSOLVER zhegvd behind hsgen_eigensolve) against scipy's eigh."""
import numpy as np
import pytest
import scipy.linalg as sl

from tests.helpers import open_dump_or_fail, rel_fro_lower

pytestmark = pytest.mark.gpu


def _full(M):
    return np.tril(M) + np.tril(M, -1).conj().T


def test_well_conditioned_problem(ctx):
    rng = np.random.default_rng(4)
    n = 300
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    Hf = (X + X.conj().T) / 2
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    Sf = Q @ np.diag(rng.uniform(1.0, 2.0, n)) @ Q.conj().T
    H = np.asfortranarray(np.tril(Hf) + 7.0 * np.triu(np.ones((n, n)), 1))   # junk above: lower is read
    S = np.asfortranarray(np.tril(Sf) + 7.0 * np.triu(np.ones((n, n)), 1))
    w, V, ms = ctx.eigensolve(H, S, vectors=True)
    wr = sl.eigh(Hf, Sf, eigvals_only=True)
    print(f"n={n}: max |dw| {np.abs(w - wr).max():.2e}, {ms:.2f} ms", flush=True)
    assert np.abs(w - wr).max() <= 1e-12 * np.abs(wr).max()
    R = Hf @ V - Sf @ V * w[None, :]
    assert np.linalg.norm(R) <= 1e-12 * np.linalg.norm(Hf) * np.linalg.norm(V)
    assert np.abs(V.conj().T @ Sf @ V - np.eye(n)).max() <= 1e-12


def test_fleur_two_atoms_spectrum(ctx):
    """Eigenvalues of the generated H and S against scipy on Fleur's own H
    and S (S has condition ~1e8 here, which sets the tolerance)."""
    d = open_dump_or_fail("tests/golden/two_atoms/fleur_dump.h5")
    s, kp = d.system(), d.kpoint(0)
    Hf, Sf = d.fleur_hs(0)
    d.close()
    h = ctx.upload(s)
    H, S = ctx.create_HS2(h, kp)
    w, _, ms = ctx.eigensolve(H, S)
    wr = sl.eigh(_full(Hf), _full(Sf), eigvals_only=True)
    err = np.abs(w - wr).max() / np.abs(wr).max()
    print(f"two_atoms: n={kp.ng} max rel |dw| {err:.2e}, {ms:.2f} ms", flush=True)
    assert err <= 1e-7


def test_indefinite_overlap_is_an_error(ctx):
    from hybridhsdla_amd._lib import HsgenError
    n = 16
    H = np.asfortranarray(np.eye(n, dtype=np.complex128))
    S = np.asfortranarray(-np.eye(n, dtype=np.complex128))
    with pytest.raises(HsgenError, match="positive definite"):
        ctx.eigensolve(H, S)
