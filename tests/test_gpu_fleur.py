"""End to end on the GPU against Fleur's own H and S (the reference's only
golden data, tests/golden/two_atoms/fleur_dump.h5): the dump is read and the
radial functions and final T computed on the host, then A/B generation and
the contraction run through the C ABI on the GPU.  The same comparison runs
in the C++ driver program (the reference's h_and_s executable over the C ABI).
"""
import re
import subprocess

import numpy as np
import pytest

from tests.helpers import driver_or_fail, open_dump_or_fail, rel_fro_lower

pytestmark = pytest.mark.gpu

FIXTURE = "tests/golden/two_atoms/fleur_dump.h5"


@pytest.fixture(scope="module")
def fleur():
    d = open_dump_or_fail(FIXTURE)
    s = d.system()
    kp = d.kpoint(0)
    H, S = d.fleur_hs(0)
    d.close()
    return s, kp, H, S


def test_fp64_matches_fleur(ctx, fleur):
    s, kp, Hf, Sf = fleur
    print(f"two_atoms: natoms={s.natoms} lmax={list(s.lmax)} ng={kp.ng}", flush=True)
    h = ctx.upload(s)
    H, S = ctx.create_HS2(h, kp)
    eh, es = rel_fro_lower(H, Hf), rel_fro_lower(S, Sf)
    print(f"GPU FP64 vs Fleur: H {eh:.3e} S {es:.3e}", flush=True)
    assert eh <= 2e-14 and es <= 2e-14


def test_contraction_only_matches_fleur(ctx, fleur):
    """construct_AB on the GPU, then the contraction-only entry point with
    the reference's argument layout."""
    s, kp, Hf, Sf = fleur
    h = ctx.upload(s)
    A, B = ctx.construct_AB(h, kp)
    H, S = ctx.create_HS2_stripped_v2(s.L, s.natoms, kp.ng, A, B, s.taa, s.tab, s.tbb, s.lmax, s.max_lmax,
                                      s.sqrt_u_norms())
    eh, es = rel_fro_lower(H, Hf), rel_fro_lower(S, Sf)
    print(f"GPU contract_hs vs Fleur: H {eh:.3e} S {es:.3e}", flush=True)
    assert eh <= 2e-14 and es <= 2e-14


@pytest.mark.parametrize("target", [1e-10, 1e-12])
def test_split_precision_matches_fleur(ctx, fleur, target):
    from hybridhsdla_amd.api import HSGEN_PREC_OZAKI_INT8
    s, kp, Hf, Sf = fleur
    h = ctx.upload(s)
    H, S, rep = ctx.create_HS2(h, kp, precision=HSGEN_PREC_OZAKI_INT8, target_rel_err=target, report=True)
    eh, es = rel_fro_lower(H, Hf), rel_fro_lower(S, Sf)
    print(f"GPU split ({rep['slices_used']} moduli) vs Fleur: H {eh:.3e} S {es:.3e}", flush=True)
    assert eh <= target and es <= target


def test_small_workload_through_the_dump_layout(ctx, tmp_path):
    """4 atoms / 2 types, lmax 8, N_G = 600: the inputs go through the dump
    layout (written, read back by the C reader), H and S come from the GPU
    and match the oracle evaluated on the original inputs, and the H/S store
    writes them out."""
    from hybridhsdla_amd import dump as dump_mod, make_config
    from oracle import pyoracle as po
    s, kps = make_config(0)
    path = tmp_path / "cfg0.h5"
    open_dump_or_fail(FIXTURE).close()      # fails where HDF5 is installed but the library is not
    dump_mod.write_dump(path, s, kps)
    with dump_mod.FleurDump(path) as d:
        s2 = d.system()
        kp2 = d.kpoint(0)
    h = ctx.upload(s2)
    H, S = ctx.create_HS2(h, kp2)
    Hr, Sr = po.reference_hs(s, kps[0])
    eh, es = rel_fro_lower(H, Hr), rel_fro_lower(S, Sr)
    print(f"N_G={kp2.ng}: H {eh:.2e} S {es:.2e}", flush=True)
    assert eh <= 1e-13 and es <= 1e-13
    with dump_mod.HSStore(tmp_path / "hs.h5") as st:
        st.add(0, H, S)
    assert (tmp_path / "hs.h5").stat().st_size > 2 * 16 * kp2.ng * (kp2.ng + 1) // 2


@pytest.mark.parametrize("args,tol", [([], 2e-14), (["--batched"], 2e-14),
                                      (["--precision", "int8", "--target", "1e-12"], 1e-12),
                                      (["--precision", "bf16", "--target", "1e-12"], 1e-12)])
def test_cpp_driver_reproduces_fleur(tmp_path, args, tol):
    """hsgen_driver reads the dump, builds H and S on the GPU, writes them in
    the reference's output layout and compares them with Fleur's own."""
    drv = driver_or_fail()
    out = tmp_path / "flapw_out.h5"
    cmd = [str(drv), FIXTURE, "--out", str(out), "--check", str(tol), *args]
    print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0, res.stderr
    m = re.search(r"rel H ([0-9.e+-]+) S ([0-9.e+-]+)", res.stdout)
    assert m, res.stdout
    assert float(m.group(1)) <= tol and float(m.group(2)) <= tol
    assert out.exists() and out.stat().st_size > 2 * 16 * 235 * 235


@pytest.fixture(scope="module")
def multi_k_dump(tmp_path_factory):
    """The 4-atom workload with three k-points of uneven size written into the
    dump layout, the oracle's H and S of each added as the reference matrices."""
    from hybridhsdla_amd import dump as dump_mod, make_config
    from hybridhsdla_amd.system import KPoint, gvectors, kmax_for_ng
    from oracle import pyoracle as po
    s, kps0 = make_config(0, ng=300)
    b33 = s.bmat33()
    kps = [kps0[0]]
    for k, ng in (((0.25, 0.0, 0.1), 220), ((0.5, 0.5, 0.0), 380)):
        k = np.array(k)
        kmax = kmax_for_ng(k, b33, ng)
        kps.append(KPoint(k=k, g=gvectors(k, b33, kmax), kmax=kmax))
    path = tmp_path_factory.mktemp("multi_k") / "dump.h5"
    open_dump_or_fail(FIXTURE).close()      # fails where HDF5 is installed but the library is not
    dump_mod.write_dump(path, s, kps)
    for i, kp in enumerate(kps):
        Hr, Sr = po.reference_hs(s, kp)
        dump_mod.write_hs_packed(path, i, Hr, Sr)
    return path, [kp.ng for kp in kps]


@pytest.mark.parametrize("args,tol", [([], 1e-13), (["--batched"], 1e-13),
                                      (["--precision", "int8", "--target", "1e-12", "--batched"], 1e-12)])
def test_cpp_driver_on_a_multi_k_dump(tmp_path, multi_k_dump, args, tol):
    """The driver over every k-point of a synthetic dump, one call per k-point
    or one batched call, against the oracle's H and S stored in the dump."""
    drv = driver_or_fail()
    path, ngs = multi_k_dump
    out = tmp_path / "flapw_out.h5"
    cmd = [str(drv), str(path), "--out", str(out), "--check", str(tol), *args]
    print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0, res.stderr
    errs = re.findall(r"k (\d+): ng (\d+).*rel H ([0-9.e+-]+) S ([0-9.e+-]+)", res.stdout)
    assert [int(n) for _, n, _, _ in errs] == ngs
    for _, _, eh, es in errs:
        assert float(eh) <= tol and float(es) <= tol
    assert out.exists()
