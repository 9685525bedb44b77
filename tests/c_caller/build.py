"""Compiles the plain-C caller of include/kp_image.h against the built
library (gcc, C99, no HIP headers: the ABI is plain C)."""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
EXE = HERE / "_build" / "kp_image_caller"
LIBDIR = ROOT / "kpolaris_amd" / "_lib"


def build() -> Path:
    EXE.parent.mkdir(parents=True, exist_ok=True)
    cmd = ["gcc", "-std=c99", "-O1", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}",
           str(HERE / "kp_image_caller.c"), "-o", str(EXE), f"-L{LIBDIR}", "-lkpolaris_amd",
           "-Wl,-rpath,$ORIGIN/../../../kpolaris_amd/_lib", "-Wl,-rpath-link,/opt/rocm/lib", "-Wl,--allow-shlib-undefined"]
    subprocess.run(cmd, check=True)
    return EXE


if __name__ == "__main__":
    print(build())
