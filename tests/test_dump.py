"""Fleur dump I/O and the end-to-end pin (CPU).

The only golden data of the reference is tests/golden/two_atoms/fleur_dump.h5
(the reference's tests/sample_input/two_atoms, README.md:9-12): Fleur's
inputs for a 2-atom Si/Na cell and Fleur's own muffin-tin H and S for its one
k-point (235 plane waves).  From the dumped potential and energy parameters
the radial solver, the T assembly, the A/B generator and the contraction must
reproduce Fleur's H and S; the reference states ~1e-14 relative agreement
(README.md:35).
"""
import shutil
import subprocess

import numpy as np
import pytest

from hybridhsdla_amd import host, make_config
from oracle import pyoracle as po
from oracle import radial as orad
from tests.helpers import open_dump_or_fail, rel_fro_lower, require_dump_io

FIXTURE = "tests/golden/two_atoms/fleur_dump.h5"

dump_mod = pytest.importorskip("hybridhsdla_amd.dump")


@pytest.fixture(scope="module")
def dump():
    d = open_dump_or_fail(FIXTURE)
    yield d
    d.close()


def _h5dump_values(dataset):
    exe = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    if not shutil.which(exe) and not __import__("os").path.exists(exe):
        pytest.skip("h5dump not available")
    out = subprocess.run([exe, "-y", "-w", "0", "-m", "%.17g", "-d", dataset, FIXTURE], capture_output=True, text=True,
                         check=True).stdout
    body = out[out.index("DATA {") + 6: out.rindex("}")]
    body = body[: body.rindex("}")] if "ATTRIBUTE" in body else body
    toks = body.replace("{", " ").replace("}", " ").replace("[", " ").replace("]", " ").replace(",", " ").split()
    vals = []
    for t in toks:
        try:
            vals.append(float(t))
        except ValueError:
            break
    return np.array(vals)


def test_reader_matches_h5dump(dump):
    info = dump.info
    assert (info["num_atoms"], info["num_types"], info["max_lmax"], info["num_kpts"]) == (2, 2, 8, 1)
    assert np.array_equal(_h5dump_values("/real_basis"), np.array(info["real_basis"]))
    assert np.array_equal(_h5dump_values("/rec_basis"), np.array(info["rec_basis"]))
    E = _h5dump_values("/energy_params").reshape(2, 9)
    pot = _h5dump_values("/potential").reshape(2, 515)
    for t in range(2):
        td = dump.type_data(t)
        assert td["lmax"] == 8 and td["mesh_n"] == 515
        assert np.array_equal(td["E"], E[t])
        assert np.array_equal(td["vr"], pot[t])
    mesh = _h5dump_values("/mesh")
    assert np.array_equal(mesh, [515, 0.022, 2.13, 515, 0.021, 2.13])
    ty, pos, rot = dump.atoms()
    assert list(ty) == [0, 1]
    assert np.array_equal(pos, [[0.125, 0.125, 0.125], [-0.135, -0.125, -0.125]])
    kp = dump.kpoint(0)
    assert kp.ng == 235 and np.allclose(kp.k, 4 / 9, rtol=0, atol=1e-8)


def test_product_radial_and_tmat_equal_oracle_on_fixture(dump):
    s = dump.system()
    for t in range(2):
        td = dump.type_data(t)
        ddn = []
        for l in range(td["lmax"] + 1):
            m = host.radial_solve(l, td["E"][l], td["dx"], td["rmt"], td["vr"])
            mo = orad.radfun(l, td["E"][l], td["vr"], td["rmt"], td["dx"], td["mesh_n"])
            assert np.allclose(m, mo, rtol=1e-13, atol=0), (t, l, m, mo)
            assert np.array_equal(s.match[t, l], m)           # atom t has type t here
            ddn.append(m[4])
        oa, oab, ob = orad.tmat_assemble(td["lmax"], td["lnonsph"], td["t4"], td["E"], np.array(ddn))
        assert np.allclose(s.taa[t], oa, rtol=0, atol=1e-15 * np.abs(oa).max())
        assert np.allclose(s.tab[t], oab, rtol=0, atol=1e-15 * np.abs(oab).max())
        assert np.allclose(s.tbb[t], ob, rtol=0, atol=1e-15 * np.abs(ob).max())


def test_oracle_pipeline_reproduces_fleur_hs(dump):
    """Dumped potential -> radial solve -> T assembly -> A/B -> contraction
    (the reference's BLAS sequence) against Fleur's own H and S."""
    s = dump.system()
    kp = dump.kpoint(0)
    Hf, Sf = dump.fleur_hs(0)
    Hr, Sr = po.reference_hs(s, kp)
    eh, es = rel_fro_lower(Hr, Hf), rel_fro_lower(Sr, Sf)
    print(f"oracle vs Fleur: H {eh:.3e} S {es:.3e}")
    assert eh <= 2e-14 and es <= 2e-14
    # a plausible mistake is far off: e.g. the conjugate convention
    assert rel_fro_lower(Hr.conj(), Hf) > 1e-2


def test_unpack_lower_order():
    # row-wise packing: (0,0), (1,0), (1,1), (2,0), ...
    p = np.arange(6) + 1j * np.arange(6)
    M = dump_mod.unpack_lower(p)
    assert M[0, 0] == 0 and M[1, 0] == 1 + 1j and M[1, 1] == 2 + 2j and M[2, 0] == 3 + 3j
    assert M[2, 2] == 5 + 5j and M[0, 1] == 0


def test_write_read_round_trip(tmp_path):
    s, kps = make_config(0)
    path = tmp_path / "cfg0.h5"
    require_dump_io()
    dump_mod.write_dump(path, s, kps)
    with dump_mod.FleurDump(path) as d:
        assert d.info["has_final"] == 1
        s2 = d.system()
        kp2 = d.kpoint(0)
    assert np.array_equal(s2.lmax, s.lmax) and np.array_equal(s2.rmt, s.rmt)
    assert np.array_equal(s2.pos, s.pos) and np.array_equal(s2.rotmat, s.rotmat)
    assert np.array_equal(s2.match, s.match) and np.array_equal(s2.bmat, s.bmat)
    assert s2.omega == pytest.approx(s.omega, rel=1e-15)
    for a in range(s.natoms):
        assert np.array_equal(s2.taa[a], s.taa[a]) and np.array_equal(s2.tab[a], s.tab[a])
        assert np.array_equal(s2.tbb[a], s.tbb[a])
    assert np.array_equal(kp2.g, kps[0].g) and np.array_equal(kp2.k, kps[0].k)


def test_written_tmat_blocks_rebuild_the_final_t(tmp_path):
    """The /tmat blocks a synthetic dump carries give back the final T through
    the reference's assembly (zero energies, lnonsph = lmax), up to the
    rounding of the +-1/2 on the T_ab diagonal."""
    s, kps = make_config(0)
    path = tmp_path / "cfg0.h5"
    require_dump_io()
    dump_mod.write_dump(path, s, kps)
    with dump_mod.FleurDump(path) as d:
        for t in range(2):
            td = d.type_data(t)
            a = int(np.nonzero(s.types == t)[0][0])
            aa, ab, bb = host.tmat_assemble(td["lmax"], td["lnonsph"], td["t4"], td["E"], np.zeros(td["lmax"] + 1))
            hl = lambda m: np.tril(m) + np.tril(m, -1).conj().T
            assert np.array_equal(aa, hl(s.taa[a])) and np.array_equal(bb, hl(s.tbb[a]))
            assert np.allclose(ab, s.tab[a], rtol=0, atol=1e-15)


def test_hs_store(tmp_path):
    H = np.arange(9).reshape(3, 3, order="F") * (1 + 0.5j)
    S = np.eye(3) * (2 - 1j)
    path = tmp_path / "out.h5"
    require_dump_io()
    with dump_mod.HSStore(path) as st:
        st.add(0, H, S)
    exe = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    out = subprocess.run([exe, "-y", "-w", "0", "-m", "%.17g", "-d", "/matrix_hamilton/1", str(path)], capture_output=True,
                         text=True, check=True).stdout
    body = out[out.index("DATA {") + 6:]
    nums = [float(x) for x in body.replace("[", " ").replace("]", " ").replace(",", " ").replace("}", " ").split()]
    # file row r = matrix column r (the reference writes the column-major buffer)
    M = np.array(nums[0::2]) + 1j * np.array(nums[1::2])
    assert np.array_equal(M.reshape(3, 3), H.T)


def test_packed_hs_round_trip(tmp_path):
    """H and S added to a synthetic dump in Fleur's packed layout read back
    as Fleur's H and S (lower triangles, bit for bit); writing a k-point twice
    is an argument error."""
    s, kps = make_config(0, ng=40)
    path = tmp_path / "cfg0.h5"
    require_dump_io()
    dump_mod.write_dump(path, s, kps + kps)
    rng = np.random.default_rng(3)
    mats = []
    for k in range(2):
        n = kps[0].ng
        H = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        S = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        dump_mod.write_hs_packed(path, k, H, S)
        mats.append((H, S))
    with pytest.raises(RuntimeError):
        dump_mod.write_hs_packed(path, 1, mats[1][0], mats[1][1])
    with dump_mod.FleurDump(path) as d:
        assert d.info["num_hs"] == 2
        for k, (H, S) in enumerate(mats):
            Hf, Sf = d.fleur_hs(k)
            assert np.array_equal(Hf, np.tril(H)) and np.array_equal(Sf, np.tril(S))
