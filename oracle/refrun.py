"""Run the reference build in oracle/_ref on a SystemSpec and parse what it prints.

TEST INFRASTRUCTURE ONLY (golden-vector generation and the CPU baseline of the benchmark).
"""
from __future__ import annotations

import json
import os
import subprocess
import tempfile

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_DRIVER = os.path.join(_HERE, "_ref", "ref_driver")
REF_DRIVER_OMP = os.path.join(_HERE, "_ref", "ref_driver_omp")


def available(omp: bool = False) -> bool:
    return os.path.exists(REF_DRIVER_OMP if omp else REF_DRIVER)


def _write(spec, d):
    from mjolnir_amd.spec import write_mjspec
    p = os.path.join(d, "in.mjspec")
    write_mjspec(spec, p)
    return p


def _read_vec3(lines, k):
    tag, n = lines[k].split()
    n = int(n)
    arr = np.array([[float(x) for x in lines[k + 1 + i].split()] for i in range(n)]).reshape(n, 3)
    return arr, k + 1 + n


def forces(spec):
    """Per-interaction {name, energy, force, pairs} and the total force/energy."""
    with tempfile.TemporaryDirectory() as d:
        p = _write(spec, d)
        out = os.path.join(d, "out.txt")
        subprocess.run([REF_DRIVER, "forces", p, out], check=True, stdout=subprocess.DEVNULL)
        lines = open(out).read().splitlines()
    res = {"interactions": []}
    k = 0
    while k < len(lines):
        w = lines[k].split()
        if w[0] == "interaction":
            item = {"name": w[1], "energy": float(w[3])}
            item["force"], k = _read_vec3(lines, k + 1)
            if k < len(lines) and lines[k].startswith("pairs"):
                n = int(lines[k].split()[1])
                item["pairs"] = np.array([[int(x) for x in lines[k + 1 + i].split()] for i in range(n)],
                                         dtype=np.int64).reshape(n, 2)
                k += 1 + n
            res["interactions"].append(item)
        elif w[0] == "total":
            res["energy"] = float(w[2])
            res["force"], k = _read_vec3(lines, k + 1)
        elif w[0] == "end":
            break
        else:
            raise ValueError(lines[k])
    return res


def run(spec, nsteps: int, every: int = 0):
    """Integrate with the reference integrator; returns list of (step, x, v)."""
    with tempfile.TemporaryDirectory() as d:
        p = _write(spec, d)
        out = os.path.join(d, "out.txt")
        subprocess.run([REF_DRIVER, "run", p, str(nsteps), out, str(every)], check=True,
                       stdout=subprocess.DEVNULL)
        lines = open(out).read().splitlines()
    frames = []
    k = 0
    while k < len(lines):
        w = lines[k].split()
        if w[0] == "step":
            step = int(w[1])
            x, k = _read_vec3(lines, k + 1)
            v, k = _read_vec3(lines, k)
            frames.append((step, x, v))
        elif w[0] == "end":
            break
        else:
            raise ValueError(lines[k])
    return frames


def bench(spec, nsteps: int, threads: int = 0, timeout: float = 600):
    """Time the reference OpenMP build's step loop; returns the parsed JSON line."""
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    env.setdefault("OMP_PROC_BIND", "close")
    with tempfile.TemporaryDirectory() as d:
        p = _write(spec, d)
        r = subprocess.run([REF_DRIVER_OMP, "bench", p, str(nsteps)], check=True, capture_output=True,
                           text=True, env=env, timeout=timeout)
    line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1]
    return json.loads(line)


def bench_many(specs, nsteps, procs: int, timeout: float = 600):
    """Time independent systems on the reference OpenMP build, one single-threaded process per
    system and `procs` processes at a time (how a CPU runs a replica ensemble best).  `nsteps` is
    one step count per system.  Returns (particle-steps, wall seconds)."""
    from concurrent.futures import ThreadPoolExecutor
    import time

    def one(k):
        return bench(specs[k], int(nsteps[k]), threads=1, timeout=timeout)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=procs) as ex:
        res = list(ex.map(one, range(len(specs))))
    wall = time.perf_counter() - t0
    ps = sum(int(nsteps[k]) * specs[k].n for k in range(len(specs)))
    return ps, wall, res


def sim(spec, total: int, save: int) -> str:
    """Text of the .ene file the reference's EnergyObserver writes over a `total`-step run."""
    return sim_files(spec, total, save)[".ene"]


def sim_files(spec, total: int, save: int) -> dict:
    """Every file the reference's observers write over a `total`-step run (output every `save`):
    ".ene" (text), "_position.xyz", "_velocity.xyz" (text), "_position.dcd", "_velocity.dcd"
    (bytes)."""
    out = {}
    with tempfile.TemporaryDirectory() as d:
        p = _write(spec, d)
        prefix = os.path.join(d, "sim")
        subprocess.run([REF_DRIVER, "sim", p, str(total), str(save), prefix], check=True, stdout=subprocess.DEVNULL)
        for suffix in (".ene", "_position.xyz", "_velocity.xyz"):
            with open(prefix + suffix) as f:
                out[suffix] = f.read()
        for suffix in ("_position.dcd", "_velocity.dcd"):
            with open(prefix + suffix, "rb") as f:
                out[suffix] = f.read()
    return out


def save(spec) -> bytes:
    """Bytes of the <prefix>_system.msg file the reference's MsgPackSaver writes for the
    initialized system (forces from one calc_force)."""
    with tempfile.TemporaryDirectory() as d:
        p = _write(spec, d)
        prefix = os.path.join(d, "ckpt")
        subprocess.run([REF_DRIVER, "save", p, prefix], check=True, stdout=subprocess.DEVNULL)
        with open(prefix + "_system.msg", "rb") as f:
            return f.read()
