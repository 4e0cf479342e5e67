"""Build libmjg.so in-tree with hipcc for gfx950.

``python -m mjolnir_amd.build`` (or ``__graft_entry__.build()``) compiles every HIP source of the
engine into ``mjolnir_amd/libmjg.so``.  Flags:

* ``--offload-arch=gfx950``: MI355X (CDNA4) code objects only;
* ``-ffp-contract=off``: no fused multiply-add contraction, so distance tests and FP64 forces
  follow the reference's operation order;
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 instructions.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmjg.so")
SOURCES = ["engine.hip", "batch.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose: bool = False) -> str:
    cmd = [HIPCC] + FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libmjg.so")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
