"""The allocation-free entry point (hsgen_create_hs_noalloc) through the C
ABI: its H and S equal hsgen_create_hs's bit for bit, over two calls in a row
(the second smaller, so it runs over what the first left in the scratch
array), with flags, and it refuses what it cannot serve with a message."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _subset(kp, idx):
    from hybridhsdla_amd.system import KPoint
    return KPoint(k=kp.k, g=np.ascontiguousarray(kp.g[np.asarray(idx, dtype=np.int64)]), kmax=kp.kmax)


def _pair(n):
    return (torch.zeros((n, n), dtype=torch.complex128, device="cuda"),
            torch.zeros((n, n), dtype=torch.complex128, device="cuda"))


def test_equals_create_hs_on_two_calls_in_a_row(ctx):
    from hybridhsdla_amd import make_config
    from oracle import pyoracle as po
    from tests.helpers import rel_fro_lower
    s, kps = make_config(0)                        # 4 atoms, lmax 8, N_G 600
    h = ctx.upload(s)
    kp = kps[0]
    for sub in (kp, _subset(kp, np.arange(257) * 2)):
        n = sub.ng
        print(f"N_G={n}", flush=True)
        H1, S1 = _pair(n)
        H2, S2 = _pair(n)
        r1 = ctx.create_hs_device(h, sub, H1.data_ptr(), S1.data_ptr(), n)
        r2 = ctx.create_hs_noalloc(h, sub, H2.data_ptr(), S2.data_ptr(), n)
        torch.cuda.synchronize()
        assert r2["tile_used"] == r1["tile_used"] and r2["ng"] == n
        assert torch.equal(H1, H2) and torch.equal(S1, S2)
        Hr, Sr = po.reference_hs(s, sub)
        # device buffers are column-major: the numpy transpose is the matrix
        H = H2.cpu().numpy().T
        S = S2.cpu().numpy().T
        assert rel_fro_lower(H, Hr) <= 1e-13 and rel_fro_lower(S, Sr) <= 1e-13
        assert not np.any(np.triu(H, 1)) and not np.any(np.triu(S, 1))


def test_flags_equal_create_hs(ctx):
    from hybridhsdla_amd import tiny_system
    from hybridhsdla_amd.api import HSGEN_FILL_UPPER, HSGEN_SKIP_H, HSGEN_SKIP_S
    s, kp = tiny_system(natoms=3, lmax=(3, 2, 4), ng_target=300, seed=5)
    h = ctx.upload(s)
    n = kp.ng
    for flags in (HSGEN_FILL_UPPER, HSGEN_SKIP_H, HSGEN_SKIP_S | HSGEN_FILL_UPPER):
        H1, S1 = _pair(n)
        H2, S2 = _pair(n)
        ctx.create_hs_device(h, kp, H1.data_ptr(), S1.data_ptr(), n, flags=flags)
        ctx.create_hs_noalloc(h, kp, H2.data_ptr(), S2.data_ptr(), n, flags=flags)
        torch.cuda.synchronize()
        assert torch.equal(H1, H2) and torch.equal(S1, S2), flags


def test_refusals(ctx):
    from hybridhsdla_amd import make_config
    from hybridhsdla_amd.api import HsgenError
    s, kps = make_config(1)                        # N_G 6002, K 3872: 2.2 GB of operands
    h = ctx.upload(s)
    kp = kps[0]
    H, S = _pair(kp.ng)
    with pytest.raises(HsgenError) as e:
        ctx.create_hs_noalloc(h, kp, H.data_ptr(), S.data_ptr(), kp.ng)
    assert e.value.status == -1 and "scratch" in str(e.value)
    sub = _subset(kp, np.arange(64))
    Hh = np.zeros((64, 64), np.complex128, order="F")
    Sh = np.zeros_like(Hh)
    with pytest.raises(HsgenError) as e:
        ctx.create_hs_noalloc(h, sub, Hh.ctypes.data, Sh.ctypes.data, 64)
    assert e.value.status == -1 and "device memory" in str(e.value)
    # the refused calls left nothing behind: a call that fits still works
    H1, S1 = _pair(64)
    H2, S2 = _pair(64)
    ctx.create_hs_device(h, sub, H1.data_ptr(), S1.data_ptr(), 64)
    ctx.create_hs_noalloc(h, sub, H2.data_ptr(), S2.data_ptr(), 64)
    torch.cuda.synchronize()
    assert torch.equal(H1, H2) and torch.equal(S1, S2)
