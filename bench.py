"""Benchmark of the force-and-integrate step on one MI355X.

``python bench.py [--steps K] [--warmup W] [--workload NAME]`` prints one JSON line: whole-job
particle-steps/s of the timed K steps (inputs resident in HBM), the roofline position of the
dominant kernel measured live with HIP events, and the reference CPU build timed on this host's
cores on a bounded sample of the same workload.

Workloads (seeded synthetic inputs, ``mjolnir_amd.workloads``):
  replicas512  (default) 512 independent CG chains, lengths log-uniform in [50, 2000] beads,
               temperatures uniform in [280, 400] K (bonds, Gaussian angles, Clementi dihedrals,
               Go contacts, excluded volume, Debye-Hueckel; BAOAB gamma = 0.1, dt = 0.01), run
               as one batch: FP32 pair/local arithmetic, int64 fixed-point force sums, FP64
               integrator, one persistent launch per 100-step segment
  chain1000    one CG chain, N = 1000, FP64 engine
  lj4096       LJ fluid, N = 4096, rho sigma^3 = 0.6, FP64 engine
  lj262144     LJ fluid, N = 262144 (the largest LJ job of the serving stream), FP64 engine
  dna200       3SPN2 double-stranded DNA, 200 bp (1198 beads), G-JF dt = 0.02, FP64 engine
  dna200x128   128 independent 200-bp duplexes (random sequences) in the replica kernel, FP32-bulk mode
  serve        a seeded Poisson stream (50 jobs/s) of LJ fluids (1e3..2.6e5 particles), CG chains
               (50..5000 beads) and 3SPN2 DNA (20..400 bp), 1e3..1e5 steps each, energy trace every
               100 steps, through the continuous-batching server (``mjolnir_amd.serve``)

The default run also serves a short stream (``--serve-jobs``, 0 to skip) and reports its
p50/p99 job latency under "serving".
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


def host_cores() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def _stratified(specs, k):
    order = sorted(range(len(specs)), key=lambda i: specs[i].n)
    step = max(1, len(order) // k)
    return [specs[order[i]] for i in range(step // 2, len(order), step)][:k]


# ------------------------------------------------------------------------------------------
# single system (FP64 engine)
# ------------------------------------------------------------------------------------------
def cpu_baseline_single(spec, target_s: float = 12.0):
    """The reference's OpenMP build (oracle/_ref/ref_driver_omp) on this host: a short probe,
    then a run sized to about target_s seconds of CPU work."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refrun
    threads = host_cores()
    if not refrun.available(omp=True):
        return None
    probe = refrun.bench(spec, 50, threads=threads)
    steps = max(50, int(target_s * probe["particle_steps_per_s"] / spec.n))
    r = refrun.bench(spec, steps, threads=threads, timeout=300)
    return {"value": r["particle_steps_per_s"], "unit": "particle-steps/s", "cores": r["threads"],
            "kind": "reference",
            "sample": f"{spec.name}: {steps} {spec.integrator.type} steps of the reference OpenMP build "
                      f"(OMP_NUM_THREADS={r['threads']}), step loop only, {r['seconds']:.1f} s"}


def bench_single(args, spec):
    from mjolnir_amd import capi
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
                "kernel": ("k_tsp_pairs" if g.potential in ("3SPN2BasePair", "3SPN2CrossStacking") else "k_pairs")
                + f" ({g.potential})", "avg_us": round(ms * 1e3, 3),
                "half_pairs": hp, "participants": np_}
    cpu = None if args.no_cpu_baseline else cpu_baseline_single(spec)
    return {
        "metric": "particle-steps/s", "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec * 1e3 / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": spec.name, "particles": spec.n, "precision": "fp64",
                   "integrator": spec.integrator.type, "dt": spec.integrator.dt,
                   "rebuilds_in_timed_region": int(r1 - r0)},
        "roofline": roofline, "cpu_baseline": cpu,
    }


# ------------------------------------------------------------------------------------------
# replica batch (headline)
# ------------------------------------------------------------------------------------------
SEGMENT = 100


def measured_traffic(workload: str, segment_steps: int):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this
    workload (scaled by the copy-kernel calibration), or None if there is none."""
    p = os.path.join(ROOT, "profiles", "k_replica_segment_replicas512_traffic.json")
    try:
        with open(p) as f:
            t = json.load(f)
    except OSError:
        return None
    if t.get("workload") != workload or t.get("segment_steps") != segment_steps:
        return None
    return t["traffic_bytes_per_launch"]


def batch_alg_bytes(b, ids, specs, steps):
    """Algorithmic bytes of one segment launch of `steps` steps over the batch.

    Per step and replica: 36 B per local term (8 B packed indices, 4 B kind, 24 B parameters),
    4 B per neighbour-list entry, 24 B per particle (inverse mass and two integrator
    coefficients).  Per launch: 288 B per particle (FP64 position, velocity, force read as
    double4, written back, and copied once as the segment-start backup)."""
    total = 0
    for r, s in zip(ids, specs):
        n_terms = sum(len(l.indices) for l in s.local)
        entries = int(sum(b.list_sizes(r)))
        total += steps * (36 * n_terms + 4 * entries + 24 * s.n) + 288 * s.n
    return total


def cpu_baseline_batch(specs, target_s: float = 15.0):
    """The reference's OpenMP build, one single-threaded process per replica and one process per
    core, on a stratified sample of the ensemble; steps per replica sized to about target_s s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refrun
    if not refrun.available(omp=True):
        return None
    procs = host_cores()
    # largest first, so the longest runs do not start last
    sample = sorted(_stratified(specs, 2 * procs), key=lambda s: -s.n)
    # a short pass over the whole sample sizes the timed one
    tot_n = sum(s.n for s in sample)
    _, wall0, res0 = refrun.bench_many(sample, [200] * len(sample), procs)
    loop0 = max(r["seconds"] for r in res0)
    steps = max(200, int(200 * target_s / max(wall0, loop0, 1e-3)))
    ps, wall, _ = refrun.bench_many(sample, [steps] * len(sample), procs)
    return {"value": round(ps / wall, 1), "unit": "particle-steps/s", "cores": procs, "kind": "reference",
            "sample": f"{len(sample)} replicas of the ensemble (stratified by length, {tot_n} particles), "
                      f"{steps} {sample[0].integrator.type} steps each on the reference OpenMP build with OMP_NUM_THREADS=1, "
                      f"{procs} processes at a time, step loops only, {wall:.1f} s wall"}


def fp64_engine_rate(specs, steps: int = 300):
    """The plain FP64 engine run replica by replica on a stratified sample (particle-steps/s)."""
    from mjolnir_amd import capi
    tps, tt = 0, 0.0
    for s in _stratified(specs, 12):
        e = capi.Engine(s, capi.FP64)
        e.step(50)
        t = time.perf_counter()
        e.step(steps)
        tt += time.perf_counter() - t
        tps += s.n * steps
        e.close()
    return tps / tt


def bench_batch(args):
    from mjolnir_amd import capi, workloads
    t = time.perf_counter()
    specs = workloads.replica_ensemble(512)
    print(f"[bench] replicas512: {len(specs)} replicas, {sum(s.n for s in specs)} particles "
          f"(generated in {time.perf_counter() - t:.1f} s)", file=sys.stderr, flush=True)
    b = capi.Batch(capi.FP32_FIXED)
    ids = [b.add(s) for s in specs]
    b.initialize()
    b.step(args.warmup)
    ps0, rb0 = b.counters()
    t0 = time.perf_counter()
    b.step(args.steps)             # returns after the device work is complete
    t1 = time.perf_counter()
    ps1, rb1 = b.counters()
    sec = t1 - t0
    value = (ps1 - ps0) / sec
    print(f"[bench] timed {args.steps} steps in {sec:.3f} s", file=sys.stderr, flush=True)

    # dominant kernel: the segment kernel (all size classes of one launch), HIP events
    ms, ps_launch = b.profile(SEGMENT, reps=20)
    alg = batch_alg_bytes(b, ids, specs, SEGMENT)
    achieved = alg / (ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": measured_traffic("replicas512", SEGMENT),
                "kernel": "k_replica_segment<float> (3 size classes, one launch each per segment, concurrent)",
                "avg_ms": round(ms, 4), "segment_steps": SEGMENT, "alg_bytes_per_launch": int(alg),
                "kernel_particle_steps_per_s": round(ps_launch / (ms * 1e-3), 1)}
    fp64 = None if args.no_fp64_sample else fp64_engine_rate(specs)
    cpu = None if args.no_cpu_baseline else cpu_baseline_batch(specs)
    return {
        "metric": "particle-steps/s", "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec * 1e3 / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+i64fixed+f64",
        "data": "synthetic",
        "config": {"workload": "replicas512", "replicas": len(specs), "particles": int(sum(s.n for s in specs)),
                   "lengths": "log-uniform [50, 2000]", "temperatures": "uniform [280, 400] K",
                   "precision": "fp32 pair/local, int64 fixed-point sums, fp64 integrator",
                   "integrator": "BAOABLangevin", "dt": 0.01, "segment_steps": SEGMENT,
                   "rebuilds_in_timed_region": int(rb1 - rb0)},
        "fp64_engine": None if fp64 is None else {
            "value": round(fp64, 1), "unit": "particle-steps/s",
            "sample": "12 replicas (stratified by length), FP64 engine one at a time",
            "speedup_of_batch": round(value / fp64, 2)},
        "roofline": roofline, "cpu_baseline": cpu,
    }


# ------------------------------------------------------------------------------------------
# 3SPN2 duplex ensemble: 200-bp duplexes in the FP32-bulk mode, batched
# ------------------------------------------------------------------------------------------
def bench_dna_ensemble(args, n_rep: int = 128):
    from mjolnir_amd import capi, workloads
    specs = [workloads.dna_duplex(n_bp=200, seed=100 + k) for k in range(n_rep)]
    b = capi.Batch(capi.FP32_FIXED)
    ids = [b.add(s) for s in specs]
    b.initialize()
    b.step(args.warmup)
    ps0, rb0 = b.counters()
    t0 = time.perf_counter()
    b.step(args.steps)
    sec = time.perf_counter() - t0
    ps1, rb1 = b.counters()
    value = (ps1 - ps0) / sec
    ms, ps_launch = b.profile(SEGMENT, reps=3)
    alg = batch_alg_bytes(b, ids, specs, SEGMENT)
    achieved = alg / (ms * 1e-3) / 1e9
    e = capi.Engine(specs[0], capi.FP64)
    e.step(100)
    t = time.perf_counter()
    e.step(500)
    fp64 = specs[0].n * 500 / (time.perf_counter() - t)
    e.close()
    cpu = None if args.no_cpu_baseline else cpu_baseline_batch(specs)
    return {
        "metric": "particle-steps/s", "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec * 1e3 / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+i64fixed+f64",
        "data": "synthetic",
        "config": {"workload": "dna200x128", "replicas": n_rep, "particles": int(sum(s.n for s in specs)),
                   "model": "3SPN2 duplex, 200 bp, random sequences", "integrator": "GJFNVTLangevin", "dt": 0.02,
                   "precision": "fp32 pair/local, int64 fixed-point sums, fp64 integrator",
                   "rebuilds_in_timed_region": int(rb1 - rb0)},
        "fp64_engine": {"value": round(fp64, 1), "unit": "particle-steps/s",
                        "sample": "one duplex on the FP64 engine", "speedup_of_batch": round(value / fp64, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None,
                     "kernel": "k_replica_segment<float, false, 1024, 2, true>", "avg_ms": round(ms, 4),
                     "segment_steps": SEGMENT, "alg_bytes_per_launch": int(alg)},
        "cpu_baseline": cpu,
    }


# ------------------------------------------------------------------------------------------
# serving stream
# ------------------------------------------------------------------------------------------
def run_serving(n_jobs: int, seed: int = 7):
    from mjolnir_amd import serve
    t = time.perf_counter()
    jobs = serve.make_jobs(serve.job_trace(n_jobs, rate=50.0, seed=seed), procs=host_cores())
    print(f"[bench] serving: {n_jobs} jobs built in {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    out = serve.Server(workers=4).run(jobs)
    out.update({"rate_jobs_per_s": 50.0, "seed": seed,
                "mix": "LJ fluid N log-uniform [1e3, 2.6e5], CG chain 50..5000 beads, 3SPN2 DNA 20..400 bp; "
                       "steps log-uniform [1e3, 1e5]; energy trace every 100 steps",
                "lanes": "batch: replicas <= 2048 particles (FP32 + int64 fixed-point sums); "
                         "engine: FP64 engine per job, 4 at a time"})
    print(f"[bench] serving: {json.dumps(out)}", file=sys.stderr, flush=True)
    return out


def bench_serve(args):
    s = run_serving(args.serve_jobs)
    return {
        "metric": "particle-steps/s (serving)", "value": s["particle_steps_per_s"], "unit": "particle-steps/s",
        "n_gpus": 1, "steps": None, "warmup": 0, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+i64fixed+f64 (batch lane), f64 (engine lane)",
        "data": "synthetic",
        "config": {"workload": "serve", "jobs": args.serve_jobs, "rate_jobs_per_s": 50.0},
        "serving": s, "roofline": None, "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--workload", default="replicas512", choices=["replicas512", "chain1000", "lj4096", "lj262144", "dna200", "dna200x128", "serve"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-sample", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=40,
                    help="jobs of the serving stream (workload serve, and the default run's serving entry)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run one device-to-device copy kernel of this size (the reference "
                         "kernel that scales hardware byte counters in a profiler pass)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (replicas and jobs "
              "shard across GPUs with no exchange)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy_mib > 0:
        import torch
        nbytes = args.calibrate_copy_mib << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] calibration copy: {nbytes} B read, {nbytes} B written", file=sys.stderr, flush=True)
        del src, dst
    if args.workload == "replicas512":
        out = bench_batch(args)
        out["serving"] = run_serving(args.serve_jobs) if args.serve_jobs > 0 else None
    elif args.workload == "serve":
        out = bench_serve(args)
    elif args.workload == "dna200x128":
        out = bench_dna_ensemble(args)
    else:
        from mjolnir_amd import workloads
        spec = {"chain1000": lambda: workloads.cg_chain(n=1000, seed=2),
                "lj4096": lambda: workloads.lj_fluid(n_side=16, seed=1),
                "lj262144": lambda: workloads.lj_fluid(n_side=64, seed=1),
                "dna200": lambda: workloads.dna_duplex(n_bp=200)}[args.workload]()
        print(f"[bench] workload {spec.name} N={spec.n}", file=sys.stderr, flush=True)
        out = bench_single(args, spec)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
