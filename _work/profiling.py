"""Profiling helpers used under rocprofv3.

``python -m kpolaris_amd.profiling copy --mib 512 --reps 5`` runs a plain
device-to-device copy of a known size: its FETCH_SIZE / WRITE_SIZE readings
calibrate the counters' units for the imaging kernels' HBM traffic.
"""
from __future__ import annotations

import argparse
import sys


def copy_probe(mib: int, reps: int) -> None:
    import torch
    n = mib * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x {reps}", flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=512)
    c.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()
    if args.cmd == "copy":
        copy_probe(args.mib, args.reps)
    return 0


if __name__ == "__main__":
    sys.exit(main())
