#!/usr/bin/env python3
"""Benchmark: seconds per H+S and effective TFLOP/s on one MI355X.

A step is one full H+S generation for one k-point of the workload (A/B
generation, per-atom T application, lower-triangle S and H contraction), with
the system description already resident in HBM and H, S written to HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2]
                    [--precision fp64|int8] [--no-cpu] [--no-check]

Prints ONE JSON line.  The default workload is the 256-atom, lmax 12,
N_G ~ 25000 system (FP64).  Effective TFLOP/s uses the algorithmic work
F = 16 G^2 K + 32 G sum_a L_a^2 (complex MAC = 8 flops, lower triangle).
After the FP64 run the same k-point runs in split precision at the targets
1e-10, 1e-12 (the headline, timed like the FP64 run) and 1e-13; the CPU leg
times the reference CPU path and checks every result against the oracle on a
stratified sample of 196 G-vectors.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "sec per H+S & TFLOP/s: FP64 MFMA vs split-precision at ≤1e-12 rel err"
FP64_PEAK = 78.6      # TFLOP/s, MI355X dense FP64 matrix
INT8_PEAK = 5000.0    # TOP/s, MI355X dense INT8
BF16_PEAK = 2500.0    # TFLOP/s, MI355X dense BF16
HBM_PEAK = 8000.0     # GB/s

WORKLOAD = {
    0: "cfg0: 4 atoms/2 types, lmax 8 (L=81), a=10 bohr, k=(0.1,0.2,0.3)",
    1: "cfg1: 32 atoms/4 types, lmax 10 (L=121), a=20 bohr, k=(0.1,0.2,0.3)",
    2: "cfg2: 256 atoms/8 types, lmax 12 (L=169), a=40 bohr, k=(0.1,0.2,0.3)",
    4: "cfg4: 48 atoms/3 types, lmax 8/10/12 by type, a=24 bohr, 4x4x4 Monkhorst-Pack k-points",
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_threads():
    """(threads used, CPUs in the affinity mask): OMP_NUM_THREADS where it is
    set (the host's CPU share for this GPU), else every CPU of the mask."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    try:
        t = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        t = 0
    return (min(t, n) if t > 0 else n), n


def stratified_sample(ng, block=128, seed=1):
    """One index drawn from every block of `block` consecutive G-vectors (the
    same sample the full-size parity tests use)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return np.array([b0 + int(rng.integers(0, min(block, ng - b0))) for b0 in range(0, ng, block)], dtype=np.int64)


def cpu_baseline(cfg: int, budget_s: float, samples=None):
    """The reference CPU path (oracle: serial construct_AB, tensor_AB_match.cpp:63-129,
    then the v2 BLAS sequence, h_and_s_stripped_v2.cpp:100-382, on the host
    OpenBLAS), timed stage by stage on the same system at a reduced plane-wave
    count sized to ~budget_s, each stage extrapolated to the full size by its
    own law (construct_AB ~ G, the BLAS sequence ~ 20 G^2 K + 24 G sum L^2);
    for cfg >= 2 also the medium workload (32 atoms, N_G 6002) timed whole at
    full size.  `samples` maps labels to (idx, H_sub, S_sub), principal
    submatrices of the GPU results on a stratified G sample; each is compared
    with the oracle evaluated on those G-vectors."""
    import numpy as np
    from hybridhsdla_amd.system import gvectors, kmax_for_ng, make_config, KPoint
    from oracle import pyoracle as po

    threads, ncpu = cpu_threads()
    if not po.blas_ready(threads):
        return None
    system, kps = make_config(cfg)
    kp0 = max(kps, key=lambda kp: kp.ng)
    full_ng = kp0.ng
    k = kp0.k
    b33 = system.bmat33()
    s = system

    def timed(sysm, kp):
        t0 = time.perf_counter()
        A, B = po.construct_ab(sysm, kp)
        t1 = time.perf_counter()
        po.create_hs2_v2(sysm.L, sysm.natoms, kp.ng, A, B, sysm.taa, sysm.tab, sysm.tbb, sysm.lmax, sysm.max_lmax,
                         sysm.sqrt_u_norms(), threads=threads)
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1

    def sub_kp(ng):
        km = kmax_for_ng(k, b33, ng)
        return KPoint(k=k, g=gvectors(k, b33, km), kmax=km)

    # calibrate on a small sample, then size the real sample to the budget
    kp_small = sub_kp(min(400, full_ng))
    ta, tc = timed(s, kp_small)
    rate = s.flops_ref(kp_small.ng) / max(tc, 1e-6)
    ab_per_g = ta / kp_small.ng
    ng_s = kp_small.ng
    for cand in (800, 1200, 1600, 2000, 2500, 3000, 4000, 5000, 6000, 8000):
        if cand > full_ng:
            break
        if s.flops_ref(cand) / rate + ab_per_g * cand <= budget_s:
            ng_s = cand
    if ng_s != kp_small.ng:
        kp_s = sub_kp(ng_s)
        ta, tc = timed(s, kp_s)
    else:
        kp_s = kp_small
    g_s = kp_s.ng
    ext_ab = ta * full_ng / g_s
    ext_c = tc * s.flops_ref(full_ng) / s.flops_ref(g_s)
    sec_full = ext_ab + ext_c
    out = {
        "value": round(s.flops_alg(full_ng) / sec_full / 1e12, 5),
        "unit": "TFLOP/s",
        "cores": threads,
        "host_cpus_in_affinity_mask": ncpu,
        "kind": "port",
        "sample": (f"{WORKLOAD.get(cfg, cfg)} system at N_G={g_s} of {full_ng} (smaller |k+G| cutoff): "
                   f"serial construct_AB {ta:.2f} s + v2 BLAS sequence {tc:.2f} s (OpenBLAS, {threads} threads); "
                   f"value = F_alg at N_G={full_ng} over the extrapolated time"),
        "ng_sample": g_s,
        "sec_construct_ab": round(ta, 4),
        "sec_contract": round(tc, 4),
        "sec_construct_ab_extrapolated": round(ext_ab, 2),
        "sec_contract_extrapolated": round(ext_c, 2),
        "sec_per_hs_extrapolated": round(sec_full, 2),
        "extrapolation": "construct_AB x G/G_s; BLAS sequence x F_ref(G)/F_ref(G_s), F_ref = 20 G^2 K + 24 G sum L_a^2",
    }
    if cfg >= 2:
        s1, k1 = make_config(1)
        ta1, tc1 = timed(s1, k1[0])
        out["measured_medium_full_size"] = {
            "workload": WORKLOAD[1] + f", N_G={k1[0].ng}, K={s1.kdim}",
            "sec_per_hs": round(ta1 + tc1, 3), "sec_construct_ab": round(ta1, 3), "sec_contract": round(tc1, 3),
            "tflops": round(s1.flops_alg(k1[0].ng) / (ta1 + tc1) / 1e12, 5), "threads": threads,
        }
    if samples:
        checks = {}
        for label, (idx, Hs, Ss) in samples.items():
            key = tuple(idx.tolist())
            if key not in checks:
                sub = KPoint(k=kp0.k, g=kp0.g[idx], kmax=kp0.kmax)
                checks[key] = po.reference_hs(s, sub, threads=threads)
            Hr, Sr = checks[key]
            rel = lambda x, r: float(np.linalg.norm(np.tril(x - r)) / np.linalg.norm(np.tril(r)))
            out.setdefault("oracle_sample_check", {})[label] = {
                "n_sample": int(len(idx)), "rel_err_H": rel(Hs, Hr), "rel_err_S": rel(Ss, Sr)}
    return out


def gather_sample(H, S, idx):
    """Principal submatrices of device H, S (column-major in torch tensors) on
    the indices idx, as column-major numpy arrays."""
    import torch
    ti = torch.as_tensor(idx, device=H.device)
    return H[ti][:, ti].cpu().numpy().T.copy(), S[ti][:, ti].cpu().numpy().T.copy()


def self_check(ctx, sysh, kp, H_dev, S_dev, nsample=48, seed=1):
    """Full-size consistency: the library recomputes H and S for a k-point
    made of nsample of the workload's G-vectors (one small tile), and the
    principal submatrix of the full-size result on those G-vectors must equal
    it (H_gg', S_gg' depend only on g and g'; every element is summed over the
    same depth in the same order, so the two agree to the last bit).  This
    uses the product only; parity against the oracle at this size is a GPU
    test (tests/test_gpu_fullsize.py)."""
    import numpy as np
    import torch
    from hybridhsdla_amd.system import KPoint

    rng = np.random.default_rng(seed)
    n = min(nsample, kp.ng)
    idx = np.sort(rng.choice(kp.ng, size=n, replace=False))
    sub = KPoint(k=kp.k, g=kp.g[idx], kmax=kp.kmax)
    Hs = torch.zeros((n, n), dtype=torch.complex128, device=H_dev.device)
    Ss = torch.zeros_like(Hs)
    ctx.create_hs_device(sysh, sub, Hs.data_ptr(), Ss.data_ptr(), n)
    ti = torch.as_tensor(idx, device=H_dev.device)
    # column-major matrices: the lower triangle is the torch tensor's upper one
    dh = torch.triu(H_dev[ti][:, ti] - Hs)
    ds = torch.triu(S_dev[ti][:, ti] - Ss)
    rel = lambda d, r: float(torch.linalg.norm(d) / torch.linalg.norm(torch.triu(r)))
    return {"n_sample": int(n), "recomputed_in": "fp64", "rel_diff_H": rel(dh, Hs), "rel_diff_S": rel(ds, Ss),
            "bitwise_equal": bool(torch.equal(dh, torch.zeros_like(dh)) and torch.equal(ds, torch.zeros_like(ds)))}


def lower_rel_err(X, ref, rows=2048):
    """||tril(X - ref)||_F / ||tril(ref)||_F for column-major device matrices
    held in torch tensors (torch row r = matrix column r, so the matrix's lower
    triangle is the tensor's upper triangle), in row blocks."""
    import torch
    num = 0.0
    den = 0.0
    n = ref.shape[0]
    for r0 in range(0, n, rows):
        r1 = min(n, r0 + rows)
        d = torch.triu(X[r0:r1] - ref[r0:r1], diagonal=r0)
        num += float(torch.sum(d.real ** 2 + d.imag ** 2))
        rr = torch.triu(ref[r0:r1], diagonal=r0)
        den += float(torch.sum(rr.real ** 2 + rr.imag ** 2))
    return (num / den) ** 0.5 if den > 0 else 0.0


def split_run(args, system, G, H, S, H64, S64, step, fp64_ms, F, target, steps, warmup, idx, ceil=None,
              bf16=False):
    """Split precision (Ozaki scheme II on INT8 MFMA, or with bf16 on BF16
    MFMA) on the same k-point at the error target, timed like the FP64 run,
    with its error against the FP64 result on the whole lower triangle;
    returns the entry and the result's principal submatrices on the
    stratified sample idx."""
    import numpy as np
    import torch
    from hybridhsdla_amd.api import HSGEN_PREC_OZAKI_BF16, HSGEN_PREC_OZAKI_INT8

    prec = HSGEN_PREC_OZAKI_BF16 if bf16 else HSGEN_PREC_OZAKI_INT8
    for _ in range(warmup):
        step(prec, target)
    torch.cuda.synchronize()
    reps = []
    t0 = time.perf_counter()
    for _ in range(steps):
        reps.append(step(prec, target))
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    eH = lower_rel_err(H, H64)
    eS = lower_rel_err(S, S64)
    smp = gather_sample(H, S, idx)
    r = reps[-1]
    nmod = int(r["slices_used"])
    ms_ig = float(np.mean([x["ms_igemm"] for x in reps]))
    ops = nmod * 16.0 * G * G * system.kdim      # INT8 ops (BF16 flops) of the nmod modulus GEMMs
    ach = ops / (ms_ig * 1e-3) / 1e12 if ms_ig > 0 else None
    pk = BF16_PEAK if bf16 else INT8_PEAK
    entry = {
        "target_rel_err": target,
        "moduli": nmod,
        "steps": steps, "warmup": warmup,
        "sec_per_hs": round(ms * 1e-3, 4),
        "value": round(F / (ms * 1e-3) / 1e12, 3),
        "unit": "TFLOP/s (FP64-equivalent)",
        "speedup_vs_fp64": round(fp64_ms / ms, 3),
        "rel_err_H_vs_fp64": eH,
        "rel_err_S_vs_fp64": eS,
        "err_model_H": r["err_est_h"], "err_model_S": r["err_est_s"],
        "err_sample_H": r["err_smp_h"], "err_sample_S": r["err_smp_s"],
        "breakdown_ms": {"contract": round(float(np.mean([x["ms_contract"] for x in reps])), 3),
                         "bf16_gemm" if bf16 else "int8_gemm": round(ms_ig, 3)},
        "roofline": {
            "bound": "mfma",
            "kernel": f"k_igemm2_crt ({'BF16' if bf16 else 'INT8'} residue GEMMs, one launch per modulus)",
            "achieved": round(ach, 1) if ach else None, "peak": pk, "unit": "TFLOP/s" if bf16 else "TOP/s",
            "frac": round(ach / pk, 4) if ach else None,
            "traffic": None if bf16 else (load_traffic("igemm2_crt", f"cfg{args.config}") or {}).get("bytes_per_launch"),
            "algorithmic_ops_per_launch": 16.0 * G * G * system.kdim,
            "avg_launch_ms": round(ms_ig / max(nmod, 1), 3),
        },
    }
    if bf16:
        entry["mfma"] = "bf16"
    else:
        with_ceiling(entry["roofline"], ceil)
    return entry, smp


def bench_batched(args) -> int:
    """64 k-points of uneven N_G through hsgen_create_hs_batched (FP64:
    grouped launches, graph captured on the warm-up call and replayed; split
    precision: INT8 launches shared by the k-points of a memory batch)."""
    import numpy as np
    import torch
    from hybridhsdla_amd import make_config
    from hybridhsdla_amd.api import Context, HSGEN_PREC_FP64, HSGEN_PREC_OZAKI_BF16, HSGEN_PREC_OZAKI_INT8

    t0 = time.perf_counter()
    system, kps = make_config(4)
    ngs = [kp.ng for kp in kps]
    log(f"[bench] workload cfg4: natoms={system.natoms} K={system.kdim} nk={len(kps)} "
        f"N_G {min(ngs)}..{max(ngs)} (setup {time.perf_counter() - t0:.1f} s)")
    torch.cuda.init()
    dev = torch.device("cuda:0")
    ctx = Context(0)
    sysh = ctx.upload(system)
    mats = [(torch.zeros((n, n), dtype=torch.complex128, device=dev),
             torch.zeros((n, n), dtype=torch.complex128, device=dev)) for n in ngs]
    outs = [(a.data_ptr(), b.data_ptr(), n) for (a, b), n in zip(mats, ngs)]
    prec = {"fp64": HSGEN_PREC_FP64, "int8": HSGEN_PREC_OZAKI_INT8, "bf16": HSGEN_PREC_OZAKI_BF16}[args.precision]

    def step():
        return ctx.create_HS2_batched(sysh, kps, precision=prec, target_rel_err=args.split_target,
                                      tile=args.tile, outputs=outs)[1]

    for i in range(args.warmup):
        r = step()
        log(f"[bench] warmup {i}: {r[0]['ms_total']:.1f} ms")
    torch.cuda.synchronize()
    reps = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        reps.append(step())
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t_start) * 1e3 / args.steps
    F = float(sum(system.flops_alg(n) for n in ngs))
    f_contract = float(sum(16.0 * n * n * system.kdim for n in ngs))
    ms_gpu = float(np.mean([r[0]["ms_total"] for r in reps]))
    tflops = F / (ms_step * 1e-3) / 1e12
    if prec == HSGEN_PREC_FP64:
        roof = {"bound": "mfma", "kernel": "batched graph: k_abgen per k, grouped T application, "
                "persistent grouped S/H contraction",
                "achieved": round(f_contract / (ms_gpu * 1e-3) / 1e12, 3), "peak": FP64_PEAK, "unit": "TFLOP/s",
                "frac": round(f_contract / (ms_gpu * 1e-3) / 1e12 / FP64_PEAK, 4), "traffic": None,
                "note": "contraction flops over the whole graph's GPU time"}
    else:
        # INT8 operations of the residue GEMMs (16 G^2 K per modulus and
        # k-point) over the GPU time of those launches (each k-point's report
        # holds its share of the launches it shared)
        ops = sum(r["slices_used"] * 16.0 * r["ng"] ** 2 * system.kdim for r in reps[-1])
        ms_ig = float(np.mean([sum(x["ms_igemm"] for x in r) for r in reps]))
        ach = ops / (ms_ig * 1e-3) / 1e12
        bf = prec == HSGEN_PREC_OZAKI_BF16
        pk = BF16_PEAK if bf else INT8_PEAK
        roof = {"bound": "mfma", "kernel": f"k_igemm2_crt ({'BF16' if bf else 'INT8'} residue GEMMs, one launch "
                "per modulus over the tiles of every k-point of a memory batch)",
                "achieved": round(ach, 1), "peak": pk, "unit": "TFLOP/s" if bf else "TOP/s",
                "frac": round(ach / pk, 4), "traffic": None, "ms_igemm_per_step": round(ms_ig, 2),
                "moduli": sorted({int(x["slices_used"]) for x in reps[-1]})}
        if not bf:
            with_ceiling(roof, int8_ceiling(ctx))
    check = None
    if not args.no_check:
        try:
            i = int(np.argmax(ngs))
            kp = kps[i]
            check = self_check(ctx, sysh, kp, mats[i][0], mats[i][1])
            check["k_index"] = i
            log(f"[bench] self check (k {i}, N_G {kp.ng}): {check}")
        except Exception as e:
            check = {"error": str(e)}
    # the same k-points one call each (no grouping, no graph), for comparison
    loop = None
    if not args.no_split:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for kp, (a, b) in zip(kps, mats):
            ctx.create_hs_device(sysh, kp, a.data_ptr(), b.data_ptr(), kp.ng, precision=prec,
                                 target_rel_err=args.split_target, tile=args.tile)
        torch.cuda.synchronize()
        loop_ms = (time.perf_counter() - t1) * 1e3
        loop = {"ms_per_step": round(loop_ms, 3), "batched_speedup": round(loop_ms / ms_step, 3)}
        log(f"[bench] one call per k-point: {loop}")
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(4, args.cpu_budget)
        except Exception as e:
            cpu = {"error": str(e)}
    out = {
        "metric": METRIC, "value": round(tflops, 3), "unit": "TFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "sec_per_hs": round(ms_step * 1e-3 / len(kps), 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": {HSGEN_PREC_FP64: "f64", HSGEN_PREC_OZAKI_INT8: "int8-slices/f64",
                                       HSGEN_PREC_OZAKI_BF16: "bf16-slices/f64"}[prec],
        "data": "synthetic (seeded system generator; random-init T-matrices and radial values)",
        "config": {"workload": WORKLOAD[4] + f", N_G {min(ngs)}..{max(ngs)}, K={system.kdim}",
                   "natoms": system.natoms, "n_k": len(kps), "n_g_total": int(sum(ngs)),
                   "k_dim": system.kdim, "precision": args.precision},
        "flops_per_step": F,
        "roofline": roof,
        "self_check": check,
        "per_k_calls": loop,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    sysh.close()
    ctx.close()
    return 0


def int8_ceiling(ctx):
    """The INT8 MFMA ceiling of this board: the split-precision GEMMs' MFMA
    instruction alone on every CU, random operands (hsgen_int8_mfma_ceiling),
    which the board's power limit holds below the nominal peak."""
    try:
        ctx.int8_mfma_ceiling(30000)               # warm-up
        tops, ms = ctx.int8_mfma_ceiling(0)
        return {"achieved": round(tops, 1), "unit": "TOP/s", "frac_of_peak": round(tops / INT8_PEAK, 4),
                "ms": round(ms, 1),
                "kernel": "k_int8_mfma_ceiling (v_mfma_i32_16x16x64_i8 only, random operands, every CU)"}
    except Exception as e:
        return {"error": str(e)}


def with_ceiling(roof, ceil):
    """Add the board's INT8 ceiling and the fraction of it reached to an INT8 roofline."""
    if roof and ceil and roof.get("achieved") and ceil.get("achieved"):
        roof["ceiling"] = ceil["achieved"]
        roof["frac_of_ceiling"] = round(roof["achieved"] / ceil["achieved"], 4)
    return roof


def load_traffic(kernel: str, workload: str):
    """HBM traffic per launch from the committed rocprofv3 --pmc summary."""
    p = ROOT / "profiles" / f"{kernel}_{workload}_traffic.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text())
    except Exception:
        return None


def tile_sweep(args) -> int:
    """Every tile configuration of one precision over problem sizes: the
    data behind the size-based default tiles (FP64 fixed; split precision at
    12 moduli)."""
    import numpy as np
    import torch
    from hybridhsdla_amd import make_config
    from hybridhsdla_amd.api import Context, HSGEN_PREC_FP64, HSGEN_PREC_OZAKI_BF16, HSGEN_PREC_OZAKI_INT8

    prec = {"fp64": HSGEN_PREC_FP64, "int8": HSGEN_PREC_OZAKI_INT8, "bf16": HSGEN_PREC_OZAKI_BF16}[args.precision]
    tiles = {"fp64": (11, 12, 13, 14), "int8": (51, 11, 12, 13, 14, 15, 16), "bf16": (61, 62, 63)}[args.precision]
    torch.cuda.init()
    ctx = Context(0)
    for cfg, ngs in ((0, [600]), (1, [600, 1200, 2000, 3000, 4000, 6002])):
        for ng in ngs:
            s, kps = make_config(cfg, ng=ng)
            h = ctx.upload(s)
            kp = kps[0]
            kw = {} if prec == HSGEN_PREC_FP64 else {"precision": prec, "slices": 12}
            _, _, rep = ctx.create_HS2(h, kp, report=True, **kw)
            default = int(rep["tile_used"])
            out = {}
            for t in tiles:
                ms = []
                for r in range(6):
                    _, _, rep = ctx.create_HS2(h, kp, tile=t, report=True, **kw)
                    if r >= 2:
                        ms.append(rep["ms_total"])
                out[str(t)] = round(float(np.median(ms)), 4)
            print(json.dumps({"tile_sweep": args.precision, "workload": f"cfg{cfg}", "n_g": kp.ng, "k_dim": s.kdim,
                              "ms_total_by_tile": out, "default_tile": default}), flush=True)
            h.close()
    ctx.close()
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1, help="number of GPUs; only 1 is measured")
    ap.add_argument("--steps", type=int, default=3, help="timed H+S evaluations")
    ap.add_argument("--warmup", type=int, default=1, help="untimed H+S evaluations before the timed ones")
    ap.add_argument("--config", type=int, default=2, choices=[0, 1, 2, 4],
                    help="workload; 4 = 64 k-points through the batched entry point")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "int8", "bf16"],
                    help="arithmetic of the measured run (int8, bf16: split precision on INT8 or BF16 MFMA)")
    ap.add_argument("--slices", type=int, default=0, help="split precision: modulus count (0: from the error model)")
    ap.add_argument("--tile", type=int, default=0, help="tile configuration id (0: default or autotuned)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-check", action="store_true",
                    help="skip the self check (a sampled submatrix recomputed as a small k-point)")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU work for the baseline sample")
    ap.add_argument("--no-split", action="store_true",
                    help="skip the split-precision comparison run after the FP64 run")
    ap.add_argument("--split-target", type=float, default=1e-12,
                    help="split precision: relative error target of the headline run")
    ap.add_argument("--no-bf16", action="store_true",
                    help="skip the split-precision run on BF16 MFMA at --split-target")
    ap.add_argument("--split-sweep", default="1e-10,1e-12,1e-13",
                    help="split precision: further targets, timed with at most 3 steps each")
    ap.add_argument("--eig", action="store_true",
                    help="also time the generalized eigensolve H x = lambda S x of the result")
    ap.add_argument("--host-out", action="store_true",
                    help="also time one H+S with host (numpy) outputs, i.e. including the PCIe copies")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="GiB: run one device-to-device copy of this size first (traffic calibration)")
    ap.add_argument("--tile-sweep", action="store_true",
                    help="instead of the benchmark: whole-call GPU time of every tile of --precision at N_G "
                         "600..6002 (the 32-atom system) and 600 (the 4-atom one), median of 4, the default "
                         "tile marked; one JSON line per size")
    args = ap.parse_args(argv)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (the H/S path is benchmarked on a "
              "single MI355X); --gpus must be 1", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from hybridhsdla_amd import make_config
    from hybridhsdla_amd.api import Context, HSGEN_PREC_FP64, HSGEN_PREC_OZAKI_BF16, HSGEN_PREC_OZAKI_INT8

    if args.config == 4:
        return bench_batched(args)
    if args.tile_sweep:
        return tile_sweep(args)

    t0 = time.perf_counter()
    system, kps = make_config(args.config)
    kp = kps[0]
    G = kp.ng
    log(f"[bench] workload cfg{args.config}: natoms={system.natoms} K={system.kdim} G={G} "
        f"kmax={kp.kmax:.4f} (setup {time.perf_counter() - t0:.1f} s)")

    torch.cuda.init()
    dev = torch.device("cuda:0")
    ctx = Context(0)
    sysh = ctx.upload(system)
    H = torch.zeros((G, G), dtype=torch.complex128, device=dev)
    S = torch.zeros((G, G), dtype=torch.complex128, device=dev)
    prec = {"fp64": HSGEN_PREC_FP64, "int8": HSGEN_PREC_OZAKI_INT8, "bf16": HSGEN_PREC_OZAKI_BF16}[args.precision]
    if args.copy_probe > 0:
        n = int(args.copy_probe * 2**30) // 8
        a = torch.empty(n, dtype=torch.float64, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: {n * 8} bytes read, {n * 8} bytes written")
        del a, b

    def step(precision=prec, target=0.0):
        return ctx.create_hs_device(sysh, kp, H.data_ptr(), S.data_ptr(), G, precision=precision,
                                    slices=args.slices, target_rel_err=target, tile=args.tile)

    for i in range(args.warmup):
        r = step()
        log(f"[bench] warmup {i}: {r['ms_total']:.1f} ms (abgen {r['ms_abgen']:.2f}, tmul "
            f"{r['ms_tmul']:.2f}, contract {r['ms_contract']:.1f})")
    torch.cuda.synchronize()
    reps = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        reps.append(step())
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_step = elapsed * 1e3 / args.steps
    log(f"[bench] {args.steps} steps: {ms_step:.1f} ms/step")

    F = system.flops_alg(G)
    tflops = F / (ms_step * 1e-3) / 1e12
    ms_contract = float(np.mean([r["ms_contract"] for r in reps]))
    ms_abgen = float(np.mean([r["ms_abgen"] for r in reps]))
    ms_tmul = float(np.mean([r["ms_tmul"] for r in reps]))
    f_contract = 16.0 * G * G * system.kdim
    contract_tf = f_contract / (ms_contract * 1e-3) / 1e12
    peak = FP64_PEAK if prec == HSGEN_PREC_FP64 else INT8_PEAK
    workload_tag = f"cfg{args.config}"
    tile_used = int(reps[-1]["tile_used"])
    # tiles 9..12 are the pipelined LDS-DMA kernel (k_zgemm2_lh), 1..8 the
    # register-staged one (k_zgemm_lh)
    kname = "zgemm2_lh" if tile_used >= 9 else "zgemm_lh"
    traffic = load_traffic(f"{kname}_contract", workload_tag)
    roofline = {
        "bound": "mfma",
        "kernel": f"k_{kname} tile {tile_used} (S and H lower-triangle tiles, one grouped launch)",
        "achieved": round(contract_tf, 3),
        "peak": peak,
        "unit": "TFLOP/s",
        "frac": round(contract_tf / peak, 4),
        "traffic": (traffic or {}).get("bytes_per_launch"),
        "algorithmic_flops_per_launch": f_contract,
        "avg_launch_ms": round(ms_contract, 3),
    }
    if prec != HSGEN_PREC_FP64:
        # split precision as the main mode: the residue GEMMs dominate
        nmod = int(reps[-1]["slices_used"])
        ms_ig = float(np.mean([r["ms_igemm"] for r in reps]))
        ach = nmod * f_contract / (ms_ig * 1e-3) / 1e12
        bf = prec == HSGEN_PREC_OZAKI_BF16
        pk = BF16_PEAK if bf else INT8_PEAK
        roofline = {
            "bound": "mfma",
            "kernel": f"k_igemm2_crt tile {tile_used} ({'BF16' if bf else 'INT8'} residue GEMMs, one launch per modulus)",
            "achieved": round(ach, 1), "peak": pk, "unit": "TFLOP/s" if bf else "TOP/s", "frac": round(ach / pk, 4),
            "traffic": None if bf else (load_traffic("igemm2_crt", workload_tag) or {}).get("bytes_per_launch"),
            "algorithmic_ops_per_launch": f_contract, "avg_launch_ms": round(ms_ig / nmod, 3),
        }
    ceil = int8_ceiling(ctx) if (prec == HSGEN_PREC_OZAKI_INT8 or (prec == HSGEN_PREC_FP64 and not args.no_split)) \
        else None
    if prec == HSGEN_PREC_OZAKI_INT8:
        with_ceiling(roofline, ceil)
    abgen_bytes = float(reps[-1]["bytes_abgen"])
    abgen = {
        "bound": "hbm", "kernel": "k_abgen_cls (A/B generator, Y_lm and Bessel values shared by rotation / rmt class)",
        "achieved": round(abgen_bytes / (ms_abgen * 1e-3) / 1e9, 1) if ms_abgen > 0 else None,
        "peak": HBM_PEAK, "unit": "GB/s",
        "frac": round(abgen_bytes / (ms_abgen * 1e-3) / 1e9 / HBM_PEAK, 4) if ms_abgen > 0 else None,
        "bytes_per_launch": abgen_bytes, "avg_launch_ms": round(ms_abgen, 4),
    }

    check = None
    if not args.no_check:
        try:
            check = self_check(ctx, sysh, kp, H, S)
            log(f"[bench] self check: {check}")
        except Exception as e:  # the check must not hide the measurement
            check = {"error": str(e)}

    eig = None
    if args.eig:
        try:
            w, _, ms_eig = ctx.eigensolve(H.data_ptr(), S.data_ptr(), n=G, ld=G)
            eig = {"n": G, "ms": round(ms_eig, 3), "lowest": float(w[0]), "solver": "rocsolver_zhegvd (eigenvalues)"}
        except Exception as e:      # a muffin-tin-only S need not be positive definite
            eig = {"n": G, "error": str(e)}
        log(f"[bench] eigensolve: {eig}")

    host_out = None
    if args.host_out:
        import numpy as np
        Hh = np.zeros((G, G), np.complex128, order="F")
        Sh = np.zeros((G, G), np.complex128, order="F")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ctx.create_HS2(sysh, kp, precision=prec, slices=args.slices, tile=args.tile, H=Hh, S=Sh)
        host_out = {"sec_per_hs": round(time.perf_counter() - t1, 4),
                    "note": "host numpy H, S: lower triangles copied over PCIe"}
        del Hh, Sh
        log(f"[bench] host outputs: {host_out}")

    # principal submatrices on a stratified G sample, for the oracle check of
    # the CPU leg
    idx = stratified_sample(G)
    samples = {"fp64": (idx, *gather_sample(H, S, idx))} if prec == HSGEN_PREC_FP64 else {}
    split = None
    sweep = None
    split_bf16 = None
    if prec == HSGEN_PREC_FP64 and not args.no_split:
        try:
            H64 = H.clone()
            S64 = S.clone()
            sweep = []
            targets = sorted({float(t) for t in args.split_sweep.split(",")} | {args.split_target}, reverse=True)
            for target in targets:
                main_t = target == args.split_target
                # the headline target is timed like the FP64 run, the others
                # with fewer steps
                st, wu = (args.steps, args.warmup) if main_t else (min(args.steps, 3), min(args.warmup, 1))
                entry, smp = split_run(args, system, G, H, S, H64, S64, step, ms_step, F, target, st, wu, idx,
                                       ceil)
                samples[f"split_{target:.0e}"] = (idx, *smp)
                sweep.append(entry)
                if main_t:
                    split = entry
                log(f"[bench] split precision {target:.0e}: {entry['moduli']} moduli, {entry['sec_per_hs']} s, "
                    f"{entry['speedup_vs_fp64']}x, err H {entry['rel_err_H_vs_fp64']:.2e} S "
                    f"{entry['rel_err_S_vs_fp64']:.2e}")
            if not args.no_bf16:
                # the same slices on BF16 MFMA: same moduli, same bits as INT8
                target = args.split_target
                entry, smp = split_run(args, system, G, H, S, H64, S64, step, ms_step, F, target,
                                       min(args.steps, 3), min(args.warmup, 1), idx, bf16=True)
                ref = samples.get(f"split_{target:.0e}")
                entry["sample_bitwise_equal_to_int8"] = bool(
                    ref is not None and np.array_equal(smp[0], ref[1]) and np.array_equal(smp[1], ref[2]))
                split_bf16 = entry
                log(f"[bench] split precision on BF16 MFMA {target:.0e}: {entry['moduli']} moduli, "
                    f"{entry['sec_per_hs']} s, {entry['speedup_vs_fp64']}x, same bits as INT8 on the sample: "
                    f"{entry['sample_bitwise_equal_to_int8']}")
            del H64, S64
        except Exception as e:
            split = {"error": str(e)}

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(args.config, args.cpu_budget, samples)
            log(f"[bench] cpu baseline: {cpu}")
        except Exception as e:
            cpu = {"error": str(e)}
    if cpu and sweep:
        for e in sweep:
            e["rel_err_vs_oracle_sample"] = cpu.get("oracle_sample_check", {}).get(f"split_{e['target_rel_err']:.0e}")

    out = {
        "metric": METRIC,
        "value": round(tflops, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "sec_per_hs": round(ms_step * 1e-3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {HSGEN_PREC_FP64: "f64", HSGEN_PREC_OZAKI_INT8: "int8-slices/f64",
                  HSGEN_PREC_OZAKI_BF16: "bf16-slices/f64"}[prec],
        "data": "synthetic (seeded system generator; random-init T-matrices and radial values)",
        "config": {
            "workload": WORKLOAD[args.config] + f", N_G={G} (Kmax={kp.kmax:.4f}), K={system.kdim}",
            "natoms": system.natoms, "n_g": G, "k_dim": system.kdim,
            "precision": args.precision, "tile": reps[-1]["tile_used"],
        },
        "flops_per_step": F,
        "breakdown_ms": {"abgen": round(ms_abgen, 3), "tmul": round(ms_tmul, 3),
                         "contract": round(ms_contract, 3)},
        "roofline": roofline,
        "abgen_roofline": abgen,
        "self_check": check,
        "split_precision": split,
        "split_sweep": sweep,
        "split_bf16": split_bf16,
        "int8_mfma_ceiling": ceil,
        "host_outputs": host_out,
        "eigensolve": eig,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    sysh.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
