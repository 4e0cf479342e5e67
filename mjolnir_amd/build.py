"""Build libmjg.so in-tree with hipcc for gfx950.

``python -m mjolnir_amd.build`` (or ``__graft_entry__.build()``) compiles every HIP source of the
engine into ``mjolnir_amd/libmjg.so``.  The translation units are compiled in parallel into
``mjolnir_amd/_build/`` and linked; the replica segment kernel is compiled once per arithmetic
type, energy flag and 3SPN2 flag (``batch_seg.hip``).  Flags:

* ``--offload-arch=gfx950``: MI355X (CDNA4) code objects only;
* ``-ffp-contract=off``: no fused multiply-add contraction, so distance tests and FP64 forces
  follow the reference's operation order (the FP32 replica-kernel units allow contraction: they
  are held to the mixed-precision force bar, and their distance tests keep the FP64 order);
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 instructions.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libmjg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def units():
    """(object name, source, extra flags) of every translation unit."""
    out = [("engine", "engine.hip", []), ("batch", "batch.hip", [])]
    for real in ("float", "double"):
        # the FP32 (mixed) replica kernels may fuse multiply-adds: their results are checked
        # against the FP64 oracle at the mixed-precision bar, not bit for bit; the FP64 ones keep
        # the reference's operation order
        contract = ["-ffp-contract=fast"] if real == "float" else []
        for energy in (0, 1):
            for tsp in (0, 1):
                out.append((f"batch_seg_{real}_e{energy}_t{tsp}", "batch_seg.hip",
                            [f"-DMJG_SEG_REAL={real}", f"-DMJG_SEG_ENERGY={energy}", f"-DMJG_SEG_TSP={tsp}"] + contract))
    return out


def _newest_source(src: str) -> float:
    """Latest modification time of a unit's source, every header in csrc/ and the C header."""
    paths = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")]
    paths += [os.path.join(CSRC, src), os.path.join(HERE, "..", "include", "mjg.h")]
    return max(os.path.getmtime(p) for p in paths if os.path.isfile(p))


def _compile(unit, verbose, incremental=False):
    name, src, extra = unit
    obj = os.path.join(OBJ, name + ".o")
    if incremental and os.path.exists(obj) and os.path.getmtime(obj) >= _newest_source(src):
        return name, None
    cmd = [HIPCC] + FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", obj + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return name, r.stdout + r.stderr
    os.replace(obj + ".tmp", obj)
    return name, None


def build(verbose: bool = False, jobs: int = 0, incremental: bool = False) -> str:
    """Compile and link libmjg.so.  incremental: skip translation units whose object is newer
    than every source (a header change rebuilds all of them)."""
    os.makedirs(OBJ, exist_ok=True)
    us = units()
    jobs = jobs or max(1, min(len(us), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda u: _compile(u, verbose, incremental), us))
    failed = [(n, log) for n, log in results if log is not None]
    if failed:
        for n, log in failed:
            sys.stderr.write(f"---- {n} ----\n{log}")
        raise RuntimeError("hipcc failed building libmjg.so: " + ", ".join(n for n, _ in failed))
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + [os.path.join(OBJ, u[0] + ".o") for u in us] + \
        ["-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("linking libmjg.so failed")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True, incremental="--incremental" in sys.argv))
