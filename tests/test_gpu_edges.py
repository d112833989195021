"""Edge cases on the GPU, against the oracle and through the error paths.

- no G-vectors, one G-vector;
- sizes on either side of every tile and padding boundary (128, 256, 768),
  in FP64 and split precision;
- generator atom classes: many distinct rotations / muffin-tin radii;
- an l cutoff of 20 next to 7;
- atoms with lmax = 0 (the reference allows any lmax_a <= max_lmax; Fleur's
  sphbes stops for a maximum lmax < 1, `sphbes.f:29`, which the library
  reports as an argument error);
- repeated G-vectors (equal rows and columns of H and S);
- batched k-points of ragged sizes, including one with a single G-vector and
  one with none, in FP64 and split precision;
- a contraction depth at which the split-precision sample splits its depth;
- argument errors come back as status codes with a message, never a crash.

Tolerances as in test_gpu_parity.py: relative Frobenius error of the lower
triangle <= 1e-13 (FP64) or the requested target (split precision).
"""
import numpy as np
import pytest

from hybridhsdla_amd import tiny_system
from hybridhsdla_amd.system import KPoint
from oracle import pyoracle as po
from tests.helpers import rel_fro_lower

pytestmark = pytest.mark.gpu

TOL = 1e-13


def _subset(kp, idx):
    return KPoint(k=kp.k, g=np.ascontiguousarray(kp.g[np.asarray(idx, dtype=np.int64)]), kmax=kp.kmax)


@pytest.fixture(scope="module")
def big_list():
    # ~800 G-vectors to cut subsets from
    return tiny_system(natoms=3, lmax=(3, 2, 4), ng_target=800, seed=5)


def test_no_g_vectors(ctx, big_list):
    s, kp = big_list
    h = ctx.upload(s)
    H, S = ctx.create_HS2(h, _subset(kp, []))
    assert H.shape == (0, 0) and S.shape == (0, 0)
    A, B = ctx.construct_AB(h, _subset(kp, []))
    assert A.shape == (s.L, s.natoms, 0)


@pytest.mark.parametrize("ng", [1, 2, 127, 128, 129, 255, 257, 767, 769])
def test_ragged_sizes_fp64(ctx, big_list, ng):
    s, kp = big_list
    assert kp.ng >= 769
    sub = _subset(kp, np.arange(ng) * (kp.ng // ng) if ng > 1 else [kp.ng // 2])
    print(f"ng={sub.ng}", flush=True)
    h = ctx.upload(s)
    Hr, Sr = po.reference_hs(s, sub)
    # the default tile (64 x 64 at these sizes) and the 128 x 128 tile
    H, S = ctx.create_HS2(h, sub)
    H11, S11 = ctx.create_HS2(h, sub, tile=11)
    assert np.array_equal(H, H11) and np.array_equal(S, S11)
    assert rel_fro_lower(H, Hr) <= TOL and rel_fro_lower(S, Sr) <= TOL
    assert not np.any(np.triu(H, 1)) and not np.any(np.triu(S, 1))
    assert np.all(np.diag(H).imag == 0) and np.all(np.diag(S).imag == 0)


@pytest.mark.parametrize("ng", [1, 191, 193, 255, 257, 767, 769])
def test_ragged_sizes_split_precision(ctx, big_list, ng):
    from hybridhsdla_amd.api import HSGEN_PREC_OZAKI_INT8
    s, kp = big_list
    sub = _subset(kp, np.arange(ng) * (kp.ng // ng) if ng > 1 else [3])
    print(f"ng={sub.ng}", flush=True)
    h = ctx.upload(s)
    H, S, rep = ctx.create_HS2(h, sub, precision=HSGEN_PREC_OZAKI_INT8, target_rel_err=1e-12, report=True)
    # the 256 x 192 INT8 tile over the same ragged edges: the same bits
    H51, S51 = ctx.create_HS2(h, sub, precision=HSGEN_PREC_OZAKI_INT8, target_rel_err=1e-12, tile=51)
    assert np.array_equal(H, H51) and np.array_equal(S, S51)
    Hr, Sr = po.reference_hs(s, sub)
    eh, es = rel_fro_lower(H, Hr), rel_fro_lower(S, Sr)
    print(f"{rep['slices_used']} moduli, tile {rep['tile_used']}: H {eh:.2e} S {es:.2e}", flush=True)
    assert eh <= 1e-12 and es <= 1e-12
    assert not np.any(np.triu(H, 1)) and not np.any(np.triu(S, 1))


def test_lmax_zero_atoms(ctx):
    s, kp = tiny_system(natoms=4, lmax=(0, 3, 1), ng_target=60, seed=11)
    assert 0 in s.lmax
    h = ctx.upload(s)
    A, B = ctx.construct_AB(h, kp)
    Ar, Br = po.construct_ab(s, kp)
    assert np.linalg.norm(A - Ar) <= TOL * np.linalg.norm(Ar)
    assert np.linalg.norm(B - Br) <= TOL * np.linalg.norm(Br)
    H, S = ctx.create_HS2(h, kp)
    Hr, Sr = po.reference_hs(s, kp)
    assert rel_fro_lower(H, Hr) <= TOL and rel_fro_lower(S, Sr) <= TOL


@pytest.mark.parametrize("natoms,rotate,own_rmt", [(16, True, True), (18, True, False), (17, False, True),
                                                   (40, False, False)])
def test_generator_rotation_and_rmt_classes(ctx, natoms, rotate, own_rmt):
    """Atoms with bitwise-equal rotations (radii) share Y_lm (Bessel values)
    in the generator; up to 16 classes of each take the class-shared kernel,
    more the general one (here 16/16, 18 rotations, 17 radii, and 40 atoms
    of one class in two chunks).  Both against the oracle."""
    s, kp = tiny_system(natoms=natoms, lmax=(2, 3, 1), ng_target=40, seed=13, rotate=rotate)
    if own_rmt:
        s.rmt = np.linspace(1.2, 1.9, natoms)
    print(f"natoms={natoms} rotate={rotate} own_rmt={own_rmt}", flush=True)
    h = ctx.upload(s)
    A, B = ctx.construct_AB(h, kp)
    Ar, Br = po.construct_ab(s, kp)
    assert np.linalg.norm(A - Ar) <= TOL * np.linalg.norm(Ar)
    assert np.linalg.norm(B - Br) <= TOL * np.linalg.norm(Br)
    H, S = ctx.create_HS2(h, kp)
    Hr, Sr = po.reference_hs(s, kp)
    assert rel_fro_lower(H, Hr) <= TOL and rel_fro_lower(S, Sr) <= TOL


@pytest.mark.parametrize("rotate", [False, True])
def test_high_lmax(ctx, rotate):
    """An l cutoff of 20 (L = 441, T blocks of 882 rows) next to lmax 7:
    high-l Bessel and Y_lm recurrences, smaller generator chunks, T-application
    row blocks of 4 tiles; FP64 and split precision against the oracle."""
    from hybridhsdla_amd.api import HSGEN_PREC_OZAKI_INT8
    s, kp = tiny_system(natoms=2, lmax=(20, 7), ng_target=80, seed=17, rotate=rotate)
    print(f"lmax {list(s.lmax)} ng={kp.ng} rotate={rotate}", flush=True)
    h = ctx.upload(s)
    A, B = ctx.construct_AB(h, kp)
    Ar, Br = po.construct_ab(s, kp)
    assert np.linalg.norm(A - Ar) <= TOL * np.linalg.norm(Ar)
    assert np.linalg.norm(B - Br) <= TOL * np.linalg.norm(Br)
    H, S = ctx.create_HS2(h, kp)
    Hr, Sr = po.reference_hs(s, kp)
    assert rel_fro_lower(H, Hr) <= TOL and rel_fro_lower(S, Sr) <= TOL
    H, S = ctx.create_HS2(h, kp, precision=HSGEN_PREC_OZAKI_INT8, target_rel_err=1e-12)
    assert rel_fro_lower(H, Hr) <= 1e-12 and rel_fro_lower(S, Sr) <= 1e-12


def test_max_lmax_zero_is_an_argument_error(ctx):
    from hybridhsdla_amd.api import HsgenError
    s, _ = tiny_system(natoms=2, lmax=(0, 0), ng_target=10)
    with pytest.raises(HsgenError) as e:
        ctx.upload(s)
    assert e.value.status == -1 and "max_lmax" in str(e.value)


def test_repeated_g_vectors(ctx, big_list):
    s, kp = big_list
    idx = [5, 17, 5, 40, 17, 17]
    sub = _subset(kp, idx)
    h = ctx.upload(s)
    H, S = ctx.create_HS2(h, sub, flags=1)          # HSGEN_FILL_UPPER: full Hermitian matrices
    Hr, Sr = po.reference_hs(s, sub)
    assert rel_fro_lower(H, Hr) <= TOL and rel_fro_lower(S, Sr) <= TOL
    # rows of one G-vector agree: bit for bit where both elements are computed
    # lower-triangle elements (columns left of both rows) and on the diagonal;
    # to rounding where one of them is the Hermitian mirror of another sum
    for a, b in ((0, 2), (1, 4), (1, 5), (4, 5)):
        assert np.array_equal(H[a, :a], H[b, :a]) and np.array_equal(S[a, :a], S[b, :a])
        assert H[a, a] == H[b, b] and S[a, a] == S[b, b]
        for M in (H, S):
            assert np.abs(M[a] - M[b]).max() <= 1e-14 * np.abs(M).max()
            assert np.abs(M[:, a] - M[:, b]).max() <= 1e-14 * np.abs(M).max()


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_unfolded_operands_when_a_norm_is_zero(ctx, precision):
    """A zero ||u_dot_l||^2 leaves B out of S for that l; the operand cannot
    carry sqrt(N) B then, so the library keeps Z, Zs = [A; N B] and
    W = M Z (S = Z^H Zs, H = Z^H W).  Against the oracle."""
    from hybridhsdla_amd.api import HSGEN_PREC_FP64, HSGEN_PREC_OZAKI_INT8
    s, kp = tiny_system(natoms=3, lmax=(3, 2, 4), ng_target=300, seed=9)
    s.match[0, 2, 4] = 0.0
    h = ctx.upload(s)
    prec, tol = (HSGEN_PREC_FP64, TOL) if precision == "fp64" else (HSGEN_PREC_OZAKI_INT8, 1e-12)
    H, S = ctx.create_HS2(h, kp, precision=prec, target_rel_err=1e-12)
    Hr, Sr = po.reference_hs(s, kp)
    eh, es = rel_fro_lower(H, Hr), rel_fro_lower(S, Sr)
    print(f"{precision}: H {eh:.2e} S {es:.2e}", flush=True)
    assert eh <= tol and es <= tol
    un = s.sqrt_u_norms()
    Hc, Sc = ctx.create_HS2_stripped_v2(s.L, s.natoms, kp.ng, *po.construct_ab(s, kp), s.taa, s.tab, s.tbb, s.lmax,
                                        s.max_lmax, un, precision=prec, target_rel_err=1e-12)
    assert rel_fro_lower(Hc, Hr) <= tol and rel_fro_lower(Sc, Sr) <= tol
    # the batched entry point on the same three-array operands: the single-k bits
    sub = _subset(kp, np.arange(0, kp.ng, 3))
    res, _ = ctx.create_HS2_batched(h, [kp, sub], precision=prec, target_rel_err=1e-12)
    for (Hb, Sb), k in zip(res, [kp, sub]):
        H1, S1 = ctx.create_HS2(h, k, precision=prec, target_rel_err=1e-12)
        assert np.array_equal(Hb, H1) and np.array_equal(Sb, S1), k.ng


def test_folded_and_unfolded_operands_agree(ctx, monkeypatch, big_list):
    """HSGEN_FOLD=0 keeps the three-array form for a system whose norms all
    allow folding: both forms agree with the oracle and with each other to
    FP64 rounding, in FP64 and in split precision, through create_hs and
    through the contraction-only entry point."""
    from hybridhsdla_amd.api import HSGEN_PREC_OZAKI_INT8
    s, kp = big_list
    sub = _subset(kp, np.arange(300) * 2)
    Hr, Sr = po.reference_hs(s, sub)
    A, B = po.construct_ab(s, sub)
    out = {}
    for fold in ("1", "0"):
        monkeypatch.setenv("HSGEN_FOLD", fold)
        h = ctx.upload(s)
        H, S = ctx.create_HS2(h, sub)
        Hs, Ss, rep = ctx.create_HS2(h, sub, precision=HSGEN_PREC_OZAKI_INT8, target_rel_err=1e-12, report=True)
        Hc, Sc = ctx.create_HS2_stripped_v2(s.L, s.natoms, sub.ng, A, B, s.taa, s.tab, s.tbb, s.lmax, s.max_lmax,
                                            s.sqrt_u_norms())
        h.close()
        print(f"fold={fold}: FP64 H {rel_fro_lower(H, Hr):.2e} S {rel_fro_lower(S, Sr):.2e}; split "
              f"({rep['slices_used']} moduli) H {rel_fro_lower(Hs, Hr):.2e} S {rel_fro_lower(Ss, Sr):.2e}", flush=True)
        assert rel_fro_lower(H, Hr) <= TOL and rel_fro_lower(S, Sr) <= TOL
        assert rel_fro_lower(Hc, Hr) <= TOL and rel_fro_lower(Sc, Sr) <= TOL
        assert rel_fro_lower(Hs, Hr) <= 1e-12 and rel_fro_lower(Ss, Sr) <= 1e-12
        out[fold] = (H, S)
    assert rel_fro_lower(out["1"][0], out["0"][0]) <= 1e-14 and rel_fro_lower(out["1"][1], out["0"][1]) <= 1e-14


@pytest.mark.parametrize("tile", [0, 11])
def test_batched_ragged_sizes_equal_single_calls(ctx, big_list, tile):
    """Host outputs, so that the empty k-point also goes through the
    device-to-host copy of the batched path; the default tile by size and
    the 128 x 128 tile."""
    s, kp = big_list
    kps = [_subset(kp, [7]), _subset(kp, []), _subset(kp, np.arange(129)), _subset(kp, np.arange(300) * 2),
           _subset(kp, np.arange(2) + 100)]
    h = ctx.upload(s)
    res, _ = ctx.create_HS2_batched(h, kps, tile=tile)
    for (Hb, Sb), k in zip(res, kps):
        H1, S1 = ctx.create_HS2(h, k)
        assert np.array_equal(Hb, H1) and np.array_equal(Sb, S1), k.ng
    assert res[1][0].shape == (0, 0)


@pytest.mark.parametrize("tile", [0, 51])
def test_batched_split_ragged_sizes_equal_single_calls(ctx, big_list, tile):
    from hybridhsdla_amd.api import HSGEN_PREC_OZAKI_INT8
    s, kp = big_list
    kps = [_subset(kp, [7]), _subset(kp, []), _subset(kp, np.arange(769)), _subset(kp, np.arange(300) * 2),
           _subset(kp, np.arange(2) + 100)]
    h = ctx.upload(s)
    res, reps = ctx.create_HS2_batched(h, kps, precision=HSGEN_PREC_OZAKI_INT8, target_rel_err=1e-12, tile=tile)
    for (Hb, Sb), k, r in zip(res, kps, reps):
        print(f"ng={k.ng}: {r['slices_used']} moduli", flush=True)
        H1, S1 = ctx.create_HS2(h, k, precision=HSGEN_PREC_OZAKI_INT8, target_rel_err=1e-12)
        assert np.array_equal(Hb, H1) and np.array_equal(Sb, S1), k.ng


def test_split_precision_deep_contraction(ctx):
    """Depth 2K >= 32768 (100 atoms with lmax 12): the split-precision FP64
    sample runs as depth chunks (two here) summed in chunk order.  The target
    holds against the oracle, the sampled error matches the measured one on
    the sampled columns' scale, and a batched call of two k-points gives the
    single calls' bits and report fields (the chunking depends on the depth
    alone)."""
    from hybridhsdla_amd.api import HSGEN_PREC_OZAKI_INT8
    s, kp = tiny_system(natoms=100, lmax=(12,), a_cell=14.0, ng_target=200, seed=4)
    assert 2 * s.kdim >= 32768
    kps = [kp, _subset(kp, np.arange(kp.ng // 2) * 2)]
    h = ctx.upload(s)
    res, reps = ctx.create_HS2_batched(h, kps, precision=HSGEN_PREC_OZAKI_INT8, target_rel_err=1e-12)
    for k, (Hb, Sb), rb in zip(kps, res, reps):
        H, S, r = ctx.create_HS2(h, k, precision=HSGEN_PREC_OZAKI_INT8, target_rel_err=1e-12, report=True)
        Hr, Sr = po.reference_hs(s, k)
        eh, es = rel_fro_lower(H, Hr), rel_fro_lower(S, Sr)
        print(f"ng={k.ng} K={s.kdim}: {r['slices_used']} moduli, rel err H {eh:.3e} S {es:.3e}, "
              f"sampled {r['err_smp_h']:.3e} / {r['err_smp_s']:.3e}", flush=True)
        assert eh <= 1e-12 and es <= 1e-12
        # 64 sampled columns of ~200 rows: the sampled error within a factor 2 of the whole one
        assert 0.5 * eh <= r["err_smp_h"] <= 2 * eh and 0.5 * es <= r["err_smp_s"] <= 2 * es
        assert np.array_equal(Hb, H) and np.array_equal(Sb, S), k.ng
        for key in ("slices_used", "err_est_h", "err_est_s", "err_smp_h", "err_smp_s"):
            assert rb[key] == r[key], (k.ng, key)
    # BF16 MFMA over a depth of 1057 stages (its FP32 accumulators reduced mod p
    # every 16): the INT8 bits
    from hybridhsdla_amd.api import HSGEN_PREC_OZAKI_BF16
    Hb, Sb = ctx.create_HS2(h, kp, precision=HSGEN_PREC_OZAKI_BF16, target_rel_err=1e-12)
    Hi, Si = ctx.create_HS2(h, kp, precision=HSGEN_PREC_OZAKI_INT8, target_rel_err=1e-12)
    assert np.array_equal(Hb, Hi) and np.array_equal(Sb, Si)


def test_argument_errors(ctx, big_list):
    from hybridhsdla_amd import _lib
    from hybridhsdla_amd.api import HsgenError, HSGEN_PREC_OZAKI_INT8
    s, kp = big_list
    sub = _subset(kp, np.arange(50))
    h = ctx.upload(s)
    H = np.zeros((50, 50), np.complex128, order="F")
    S = np.zeros_like(H)
    # leading dimension below ng
    with pytest.raises(HsgenError) as e:
        ctx.create_hs_device(h, sub, H.ctypes.data, S.ctypes.data, 49)
    assert e.value.status == -1 and "ld" in str(e.value)
    # unknown precision
    with pytest.raises(HsgenError) as e:
        ctx.create_HS2(h, sub, precision=7)
    assert e.value.status == -1
    # more moduli than the table holds
    with pytest.raises(HsgenError) as e:
        ctx.create_HS2(h, sub, precision=HSGEN_PREC_OZAKI_INT8, slices=99)
    assert e.value.status == -1
    # too few moduli for the contraction depth: an accuracy error, not a wrong result
    with pytest.raises(HsgenError) as e:
        ctx.create_HS2(h, sub, precision=HSGEN_PREC_OZAKI_INT8, slices=2)
    assert e.value.status == -6
    # a target below what 52 bits per element reach: an accuracy error, in
    # the single-k and in the batched entry point
    with pytest.raises(HsgenError) as e:
        ctx.create_HS2(h, sub, precision=HSGEN_PREC_OZAKI_INT8, target_rel_err=1e-20)
    assert e.value.status == -6 and "reaches rel err" in str(e.value)
    with pytest.raises(HsgenError) as e:
        ctx.create_HS2_batched(h, [sub, _subset(kp, np.arange(7))], precision=HSGEN_PREC_OZAKI_INT8,
                               target_rel_err=1e-20)
    assert e.value.status == -6
    # null outputs
    lib = _lib.load()
    with pytest.raises(HsgenError):
        ctx.create_hs_device(h, sub, 0, 0, 50)
    # the context still works after the errors
    H2, S2 = ctx.create_HS2(h, sub)
    Hr, Sr = po.reference_hs(s, sub)
    assert rel_fro_lower(H2, Hr) <= TOL and rel_fro_lower(S2, Sr) <= TOL
    assert lib.hsgen_status_string(-6) == b"accuracy target unreachable"


def test_contexts_in_parallel_threads(big_list):
    """Each context owns its stream and workspaces: two contexts driven from
    two host threads at once give the same bits as one context alone."""
    import threading
    from hybridhsdla_amd.api import Context, HSGEN_PREC_OZAKI_INT8
    s, kp = big_list
    subs = [_subset(kp, np.arange(400)), _subset(kp, np.arange(300) * 2 + 1)]
    with Context(0) as c0:
        h = c0.upload(s)
        ref = [c0.create_HS2(h, subs[0]), c0.create_HS2(h, subs[1], precision=HSGEN_PREC_OZAKI_INT8, slices=12)]
        h.close()
    out = [None, None]
    errs = []

    def work(i):
        try:
            with Context(0) as c:
                hh = c.upload(s)
                for _ in range(3):
                    if i == 0:
                        out[i] = c.create_HS2(hh, subs[0])
                    else:
                        out[i] = c.create_HS2(hh, subs[1], precision=HSGEN_PREC_OZAKI_INT8, slices=12)
                hh.close()
        except Exception as e:      # reported below, in the main thread
            errs.append(e)

    ts = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not errs, errs
    for (H, S), (Hr, Sr) in zip(out, ref):
        assert np.array_equal(H, Hr) and np.array_equal(S, Sr)
