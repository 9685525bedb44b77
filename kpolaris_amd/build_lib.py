"""Build the HIP imaging library for gfx950 (in-tree, no JIT cache).

`python -m kpolaris_amd.build_lib` or `build()` compiles the host code
(kp_image.hip, kp_host.cpp) and the kernel parts of kp_kernels.hip (one
object per KP_PART, compiled in parallel) into
kpolaris_amd/_lib/libkpolaris_amd.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "_lib"
LIB = OUT_DIR / "libkpolaris_amd.so"
OPCOUNT = OUT_DIR / "opcount.json"

SOURCES = ["kp_image.hip", "kp_host.cpp"]
KERNELS = "kp_kernels.hip"
NPARTS = 17  # KP_PART values of kp_kernels.hip

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # multiply-adds are fused within an expression only (llvm.fmuladd), so
    # every instantiation of the lane code rounds the same way: a batched
    # image equals the single image bit for bit, and the host build of the
    # lanes (with FMA) reproduces the device's fused operations
    "-ffp-contract=on",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the imaging library needs ROCm's hipcc")


def jobs() -> int:
    """Parallel compiles: the host's CPUs, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def build(verbose: bool = False, resource_usage: bool = False, defines: dict | None = None,
          lib: Path | None = None) -> Path:
    """Compile and link the library.  `defines` sets the kernel's tuning macros
    (KP_WAVES_PER_EU, KP_LANE_LDS); `lib` redirects the output for a tuning
    build that must not replace the in-tree library."""
    OUT_DIR.mkdir(parents=True, exist_ok=True)
    lib = Path(lib) if lib else LIB
    tag = lib.stem
    dflags = [f"-D{k}={v}" for k, v in (defines or {}).items()]
    units = [(CSRC / src, OUT_DIR / f"{tag}.{Path(src).stem}.o", []) for src in SOURCES]
    units += [(CSRC / KERNELS, OUT_DIR / f"{tag}.kp_kernels.{k}.o", [f"-DKP_PART={k}"]) for k in range(NPARTS)]

    def compile_unit(u):
        src, obj, extra = u
        cmd = [hipcc(), *HIPCC_FLAGS, *dflags, *extra, "-x", "hip", "-c", str(src), "-o", str(obj)]
        if resource_usage and src.suffix == ".hip":
            cmd.append("-Rpass-analysis=kernel-resource-usage")
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        return u, res

    # the largest parts first
    order = sorted(units, key=lambda u: 0 if u[0].name == KERNELS else 1)
    failed = []
    with ThreadPoolExecutor(max_workers=jobs()) as ex:
        for u, res in ex.map(compile_unit, order):
            out = "\n".join(l for l in (res.stdout + res.stderr).splitlines()
                            if "packed-fp32-ops" not in l)
            if out.strip() and (verbose or res.returncode != 0 or resource_usage):
                print(out, flush=True)
            if res.returncode != 0:
                failed.append(str(u[0]) + " " + " ".join(u[2]))
    if failed:
        raise RuntimeError("compilation failed: " + "; ".join(failed))
    objs = [str(u[1]) for u in units]
    tmp = lib.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, lib)
    for o in objs:
        Path(o).unlink(missing_ok=True)
    if lib == LIB:
        build_opcount(verbose)
        build_io(verbose)
    return lib


IO_LIB = OUT_DIR / "libkpolaris_io.so"
HDF5_PREFIX = Path(os.environ.get("KP_HDF5_PREFIX", "/opt/conda"))


def build_io(verbose: bool = False) -> Path | None:
    """HDF5 input/output library (kp_io.cpp, host only, HDF5 C API), linked
    against the imaging library.  Skipped with a message when HDF5's headers
    are absent."""
    if not (HDF5_PREFIX / "include" / "hdf5.h").exists():
        print(f"kpolaris_amd: no HDF5 under {HDF5_PREFIX}; libkpolaris_io.so not built", flush=True)
        return None
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra", f"-I{HDF5_PREFIX / 'include'}",
           str(CSRC / "kp_io.cpp"), "-o", str(IO_LIB), f"-L{HDF5_PREFIX / 'lib'}", "-lhdf5",
           f"-Wl,-rpath,{HDF5_PREFIX / 'lib'}", f"-L{OUT_DIR}", "-lkpolaris_amd", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return IO_LIB


def build_opcount(verbose: bool = False) -> Path:
    """Host-only op-counting tool; writes the per-unit FLOP table."""
    exe = OUT_DIR / "kp_opcount"
    cmd = [hipcc(), "-x", "hip", "--offload-host-only", "-O1", "-std=c++17",
           str(CSRC / "kp_opcount.cpp"), "-o", str(exe)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, stderr=subprocess.DEVNULL if not verbose else None)
    res = subprocess.run([str(exe)], check=True, capture_output=True, text=True)
    OPCOUNT.write_text(res.stdout)
    return OPCOUNT


if __name__ == "__main__":
    build(verbose=True, resource_usage="--usage" in sys.argv)
    print(LIB)
