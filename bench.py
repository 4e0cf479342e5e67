"""Benchmark of the force-and-integrate step on one MI355X.

``python bench.py [--steps K] [--warmup W] [--workload NAME]`` prints one JSON line:
particle-steps/s of the timed K steps (inputs resident in HBM), the roofline position of the
dominant force kernel measured live with HIP events, and the reference CPU build timed on this
host's cores on a bounded sample of the same workload.

Workloads (seeded synthetic inputs, ``mjolnir_amd.workloads``):
  chain1000  protein-like CG chain, N = 1000, FP64 (bonds, Gaussian angles, Clementi dihedrals,
             Go contacts, excluded volume, Debye-Hueckel; BAOAB gamma = 0.1, dt = 0.01)
  lj4096     LJ fluid, N = 4096, rho sigma^3 = 0.6, FP64 (BAOAB gamma = 0.1, dt = 0.01)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def make_workload(name: str):
    from mjolnir_amd import workloads
    if name == "chain1000":
        return workloads.cg_chain(n=1000, seed=2)
    if name == "lj4096":
        return workloads.lj_fluid(n_side=16, seed=1)
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(spec, target_s: float = 12.0):
    """The reference's OpenMP build (oracle/_ref/ref_driver_omp) on this host: a short probe,
    then a run sized to about target_s seconds of CPU work."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refrun
    threads = os.cpu_count() or 1
    if not refrun.available(omp=True):
        return None
    probe = refrun.bench(spec, 50, threads=threads)
    steps = max(50, int(target_s * probe["particle_steps_per_s"] / spec.n))
    r = refrun.bench(spec, steps, threads=threads, timeout=300)
    return {"value": r["particle_steps_per_s"], "unit": "particle-steps/s", "cores": r["threads"],
            "kind": "reference",
            "sample": f"{spec.name}: {steps} BAOAB steps of the reference OpenMP build "
                      f"(OMP_NUM_THREADS={r['threads']}), step loop only, {r['seconds']:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--workload", default="chain1000")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (replicas and jobs "
              "shard across GPUs with no exchange; see the design notes)", file=sys.stderr)
        sys.exit(2)

    from mjolnir_amd import capi
    spec = make_workload(args.workload)
    print(f"[bench] workload {spec.name} N={spec.n}", file=sys.stderr, flush=True)
    eng = capi.Engine(spec)
    eng.step(args.warmup)
    s0, r0 = eng.counters()
    t0 = time.perf_counter()
    eng.step(args.steps)           # returns after the device work is complete
    t1 = time.perf_counter()
    s1, r1 = eng.counters()
    assert s1 - s0 == args.steps
    sec = t1 - t0
    value = spec.n * args.steps / sec

    # dominant force kernel: the pair interaction with the most list entries
    best = None
    for k in range(eng.n_local, eng.num_interactions):
        ms, hp = eng.profile_interaction(k, reps=200)
        if best is None or ms > best[1]:
            best = (k, ms, hp)
    k, ms, hp = best
    g = spec.global_[k - eng.n_local]
    np_ = len(g.participants)
    # algorithmic bytes per launch: 4 B per half-list entry, 8 B CSR offset/count, 24 B FP64
    # position read and 24 B force write per participant
    alg_bytes = 4 * hp + (8 + 24 + 24) * np_
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None,
                "kernel": f"k_pair_full<{g.potential}>", "avg_us": round(ms * 1e3, 3),
                "half_pairs": hp, "participants": np_}

    cpu = None if args.no_cpu_baseline else cpu_baseline(spec)
    out = {
        "metric": "particle-steps/s", "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec * 1e3 / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": spec.name, "particles": spec.n, "precision": "fp64",
                   "integrator": spec.integrator.type, "dt": spec.integrator.dt,
                   "rebuilds_in_timed_region": int(r1 - r0)},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
