"""Build libmotion_hip.so for gfx950 in-tree (hipcc; no JIT cache, so the library travels with
the source tree).

    python -m <package>.build         (or: python 3fold-motion-engine_amd/build.py)
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "libmotion_hip.so"
SOURCES = ["motion_hip.hip"]

# -ffp-contract=off: no fused multiply-add but those the source writes out (fmaf, as the CPU oracle
# writes the same ones) -- bit parity with the oracle;
# IEEE fp32 div/sqrt and denormals kept; packed fp32 ops disabled (some packed forms have
# returned wrong results on these machines).
HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose: bool = False) -> Path:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc, *HIPFLAGS, "-I", str(HERE.parent / "include"), "-o", str(tmp),
           *[str(CSRC / s) for s in SOURCES]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libmotion_hip.so")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
