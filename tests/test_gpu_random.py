"""Randomised parity on the GPU: small systems drawn from a seeded generator,
each with a general (non-orthogonal, non-symmetric) reciprocal lattice, so
that every off-diagonal element of bmat and of the per-atom R bmat takes part
(the synthetic workloads are cubic), random atom counts, l cutoffs per type,
k-points, plane-wave counts, positions and rotations.

Per case, against the oracle on the same inputs:
- A and B (elementwise-normwise <= 1e-13);
- H and S in FP64 (relative Frobenius error of the lower triangle <= 1e-13);
- H and S in split precision at a requested 1e-11 (<= 1e-11);
and three k-points of the case in one batched FP64 call equal the single
calls bit for bit.
"""
import numpy as np
import pytest

from hybridhsdla_amd.system import KPoint, gvectors, kmax_for_ng, make_system
from oracle import pyoracle as po
from tests.helpers import rel_fro_lower

pytestmark = pytest.mark.gpu


def _random_case(seed):
    rng = np.random.default_rng(1000 + seed)
    ntypes = int(rng.integers(1, 4))
    type_lmax = rng.integers(1, 11, ntypes)
    natoms = int(rng.integers(1, 7))
    s = make_system(8.0, natoms, ntypes, type_lmax, 1.6, seed_base=2000 + 10 * seed, dmin=1.0)
    # a general lattice: cubic reciprocal vectors sheared and stretched
    b = (2.0 * np.pi / 8.0) * (np.eye(3) + 0.35 * rng.uniform(-1.0, 1.0, (3, 3)))
    assert abs(np.linalg.det(b)) > 0.2 * (2.0 * np.pi / 8.0) ** 3
    s.bmat = np.asfortranarray(b).reshape(-1, order="F").copy()
    s.omega = float((2.0 * np.pi) ** 3 / abs(np.linalg.det(b)))
    for a in range(natoms):
        if rng.random() < 0.5:
            q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
            s.rotmat[a] = (q @ b).reshape(-1, order="F")
        else:
            s.rotmat[a] = s.bmat
    kps = []
    for _ in range(3):
        k = rng.uniform(-0.5, 0.5, 3)
        ng = int(rng.integers(1, 400))
        kmax = kmax_for_ng(k, b, ng)
        kps.append(KPoint(k=k, g=gvectors(k, b, kmax), kmax=kmax))
    return s, kps


@pytest.mark.parametrize("seed", range(6))
def test_random_general_lattice(ctx, seed):
    from hybridhsdla_amd.api import HSGEN_PREC_OZAKI_INT8
    s, kps = _random_case(seed)
    print(f"seed {seed}: natoms {s.natoms}, lmax {s.lmax.tolist()}, ng {[k.ng for k in kps]}", flush=True)
    h = ctx.upload(s)
    kp = kps[0]
    A, B = ctx.construct_AB(h, kp)
    Ar, Br = po.construct_ab(s, kp)
    for X, Xr in ((A, Ar), (B, Br)):
        assert np.linalg.norm((X - Xr).ravel()) <= 1e-13 * np.linalg.norm(Xr.ravel())
    singles = []
    for k in kps:
        H, S = ctx.create_HS2(h, k)
        Hr, Sr = po.reference_hs(s, k)
        eh, es = rel_fro_lower(H, Hr), rel_fro_lower(S, Sr)
        Hs, Ss = ctx.create_HS2(h, k, precision=HSGEN_PREC_OZAKI_INT8, target_rel_err=1e-11)
        fh, fs = rel_fro_lower(Hs, Hr), rel_fro_lower(Ss, Sr)
        print(f"  ng {k.ng}: FP64 H {eh:.2e} S {es:.2e}; split H {fh:.2e} S {fs:.2e}", flush=True)
        assert eh <= 1e-13 and es <= 1e-13
        assert fh <= 1e-11 and fs <= 1e-11
        singles.append((H, S))
    res, _ = ctx.create_HS2_batched(h, kps)
    for (Hb, Sb), (H, S) in zip(res, singles):
        assert np.array_equal(Hb, H) and np.array_equal(Sb, S)
