"""Build libmmnorm_hip.so in-tree with hipcc for gfx950.

    python -m mmnorm_amd.build           # compile (always), report the output
    python -m mmnorm_amd.build --asm     # also write gfx950 assembly to _build/

All HIP sources are compiled with packed FP32 instructions disabled
(``-packed-fp32-ops``): some forms of the packed FP32 add, multiply and FMA
have given wrong results on MI355X now and then, without an error, so the kernels use only
ordinary FP32 instructions (the host pass warns that it does not know the
feature; that is harmless).
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OUT = os.path.join(PKG, "libmmnorm_hip.so")
ARCH = "gfx950"

SOURCES = ["capi.hip", "bp_forward.hip", "bp_profile.hip", "sn_vote.hip", "bp_backward.hip",
           "rsdf_bfs.hip", "segment.hip", "raymarch.hip"]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-gpu-rdc", "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def build(verbose: bool = False, asm: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = OUT + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", INCLUDE, "-I", CSRC, *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    if asm:
        outdir = os.path.join(ROOT, "_build")
        os.makedirs(outdir, exist_ok=True)
        for s in srcs:
            base = os.path.splitext(os.path.basename(s))[0]
            subprocess.run([hipcc(), *[f for f in HIPCC_FLAGS if f not in ("-shared", "-fPIC")],
                            "-I", INCLUDE, "-I", CSRC, "--cuda-device-only", "-S", s,
                            "-o", os.path.join(outdir, base + ".s")], check=True)
    return OUT


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--asm", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    out = build(verbose=a.verbose, asm=a.asm)
    print(out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
