"""Build the HIP imaging library for gfx950 (in-tree, no JIT cache).

`python -m kpolaris_amd.build_lib` or `build()` compiles
kpolaris_amd/csrc/*.{hip,cpp} into kpolaris_amd/_lib/libkpolaris_amd.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "_lib"
LIB = OUT_DIR / "libkpolaris_amd.so"
OPCOUNT = OUT_DIR / "opcount.json"

SOURCES = ["kp_image.hip", "kp_host.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=fast",
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


def build(verbose: bool = False, resource_usage: bool = False, defines: dict | None = None,
          lib: Path | None = None) -> Path:
    """Compile and link the library.  `defines` sets the kernel's tuning macros
    (KP_WAVES_PER_EU, KP_LANE_LDS); `lib` redirects the output for a tuning
    build that must not replace the in-tree library."""
    OUT_DIR.mkdir(parents=True, exist_ok=True)
    lib = Path(lib) if lib else LIB
    dflags = [f"-D{k}={v}" for k, v in (defines or {}).items()]
    objs = []
    for src in SOURCES:
        obj = OUT_DIR / (Path(src).stem + ".o")
        cmd = [hipcc(), *HIPCC_FLAGS, *dflags, "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)]
        if resource_usage and src.endswith(".hip"):
            cmd.append("-Rpass-analysis=kernel-resource-usage")
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(str(obj))
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
    return lib


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
