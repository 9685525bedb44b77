"""Builds the CPU emulation of the kernel's lane logic (test harness only)."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
LIB = HERE / "_build" / "libkp_lane_emul.so"


def build() -> Path:
    """Host-only build (no device code) of the lanes and the host layer."""
    from kpolaris_amd.build_lib import HIPCC_FLAGS, hipcc
    LIB.parent.mkdir(parents=True, exist_ok=True)
    srcs = [HERE / "kp_lane_emul.cpp", ROOT / "kpolaris_amd" / "csrc" / "kp_host.cpp"]
    flags = [f for f in HIPCC_FLAGS if not f.startswith("--offload-arch")]
    # -mfma: the host lowers the fused multiply-adds the device forms
    cmd = [hipcc(), *flags, "-mfma", "--offload-host-only", "-x", "hip", "-shared", *map(str, srcs), "-o",
           str(LIB)]
    subprocess.run(cmd, check=True, stderr=subprocess.DEVNULL)
    return LIB


if __name__ == "__main__":
    print(build())
