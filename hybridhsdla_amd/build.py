"""Build the hsgen HIP library (gfx950) and the CPU oracle, in-tree.

    python -m hybridhsdla_amd.build            # library + oracle
    python -m hybridhsdla_amd.build --lib-only

The library is compiled with hipcc for gfx950 only into
hybridhsdla_amd/lib/libhsgen.so; the oracle (test infrastructure) is compiled
with gcc into oracle/liborc.so.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libhsgen.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liborc.so"

HIP_SOURCES = ["abgen.hip", "zgemm.hip", "ozaki.hip", "context.hip", "fp64_path.hip", "split_path.hip",
               "tune_eig.hip", "noalloc.hip"]
# host-only set-up code (radial solver, T assembly), linked into the same library
HOST_SOURCES = ["radial.cpp", "tmat.cpp"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-ffp-contract=off"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the hsgen library needs ROCm's hipcc for gfx950")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(
        ln for ln in (res.stdout + res.stderr).splitlines()
        if "packed-fp32-ops" not in ln  # host pass does not know the feature: harmless
    )
    if res.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def build_lib(jobs: int = 4) -> Path:
    hipcc = _hipcc()
    LIBDIR.mkdir(exist_ok=True)
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in HIP_SOURCES if (CSRC / s).exists()]
    objs = []
    procs = []
    for src in srcs:
        obj = objdir / (src.stem + ".o")
        objs.append(obj)
        cmd = [hipcc, *HIPCC_FLAGS, "-I", str(ROOT / "include"), "-c", str(src), "-o", str(obj)]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        if len(procs) >= jobs:
            _drain(procs)
    for src in HOST_SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        objs.append(obj)
        cmd = [shutil.which("g++") or "g++", *HOST_FLAGS, "-I", str(ROOT / "include"), "-c",
               str(CSRC / src), "-o", str(obj)]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        if len(procs) >= jobs:
            _drain(procs)
    _drain(procs)
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas"])
    os.replace(tmp, LIB)
    return LIB


HDF5_ROOT = Path(os.environ.get("HSGEN_HDF5_ROOT", "/opt/conda"))
IO_LIB = LIBDIR / "libhsgen_io.so"


def hdf5_available() -> bool:
    """True when the HDF5 headers and shared library the dump I/O library
    builds against are installed (HSGEN_HDF5_ROOT, default /opt/conda)."""
    inc, lib = HDF5_ROOT / "include", HDF5_ROOT / "lib"
    return (inc / "hdf5.h").exists() and bool(list(lib.glob("libhdf5.so*")))


def build_io() -> Path | None:
    """libhsgen_io.so: the Fleur dump reader / H-S writer (host C++, HDF5),
    linked against libhsgen.so.  Skipped when no HDF5 is installed."""
    inc, lib = HDF5_ROOT / "include", HDF5_ROOT / "lib"
    if not hdf5_available():
        print(f"HDF5 not found under {HDF5_ROOT}: libhsgen_io.so not built", file=sys.stderr)
        return None
    tmp = IO_LIB.with_suffix(".so.tmp")
    _run([shutil.which("g++") or "g++", *HOST_FLAGS, "-shared", "-I", str(ROOT / "include"), "-isystem", str(inc),
          "-o", str(tmp), str(CSRC / "dump_io.cpp"), "-L", str(LIBDIR), "-lhsgen", "-L", str(lib), "-lhdf5",
          # RUNPATH (not RPATH) and a static C++ runtime: the HDF5 directory is
          # searched for this library's own dependencies only, so its older
          # libstdc++ is never picked up for the HIP runtime
          "-static-libstdc++", "-static-libgcc", "-Wl,--enable-new-dtags", f"-Wl,-rpath,$ORIGIN:{lib}"])
    os.replace(tmp, IO_LIB)
    return IO_LIB


DRIVER = LIBDIR / "hsgen_driver"


def build_driver() -> Path | None:
    """hsgen_driver: the reference's h_and_s executable over the C ABI (plain
    C++11 against include/, linked to libhsgen_io.so and libhsgen.so).
    Skipped when the dump I/O library is not built (no HDF5)."""
    if not IO_LIB.exists():
        print("libhsgen_io.so not built: hsgen_driver not built", file=sys.stderr)
        return None
    tmp = DRIVER.with_suffix(".tmp")
    _run([shutil.which("g++") or "g++", "-O2", "-std=c++11", "-Wall", "-Wextra", "-Werror", "-I", str(ROOT / "include"),
          "-o", str(tmp), str(CSRC / "hsgen_driver.cpp"), "-L", str(LIBDIR), "-lhsgen_io", "-lhsgen",
          "-Wl,--enable-new-dtags", "-Wl,-rpath,$ORIGIN:/opt/rocm/lib"])
    os.replace(tmp, DRIVER)
    return DRIVER


def _drain(procs) -> None:
    while procs:
        cmd, p = procs.pop(0)
        out, _ = p.communicate()
        out = "\n".join(ln for ln in out.splitlines() if "packed-fp32-ops" not in ln)
        if p.returncode != 0:
            raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
        if out.strip():
            print(out, file=sys.stderr)


def build_oracle() -> Path:
    cc = shutil.which("gcc") or "gcc"
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run([cc, "-O2", "-std=c99", "-fPIC", "-shared", "-Wall", "-o", str(tmp),
          str(ORACLE_DIR / "hs_oracle.c"), "-ldl", "-lm"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    ap.add_argument("--oracle-only", action="store_true")
    args = ap.parse_args(argv)
    if not args.oracle_only:
        print("built", build_lib())
        io = build_io()
        if io:
            print("built", io)
            drv = build_driver()
            if drv:
                print("built", drv)
    if not args.lib_only:
        print("built", build_oracle())
    return 0


if __name__ == "__main__":
    sys.exit(main())
