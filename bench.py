"""Benchmark of the force-and-integrate step on one MI355X.

``python bench.py [--steps K] [--warmup W] [--workload NAME]`` prints one JSON line: whole-job
particle-steps/s of the timed K steps (inputs resident in HBM), the roofline position of the
dominant kernel measured live with HIP events, and the reference CPU build timed on this host's
cores on a bounded sample of the same workload.

Workloads (seeded synthetic inputs, ``mjolnir_amd.workloads``):
  replicas512  (default) config 4: 512 independent CG chains, lengths log-uniform in [50, 2000]
               beads, temperatures uniform in [280, 400] K (bonds, Gaussian angles, Clementi
               dihedrals, Go contacts, excluded volume, Debye-Hueckel; BAOAB gamma = 0.1,
               dt = 0.01), one batch: FP32 pair/local arithmetic, int64 fixed-point force sums,
               FP64 integrator, one persistent launch per 100-step segment.  The config's whole
               job (10^4 steps from the generated state) is timed first; `value` is the timed K
               steps that follow it (the equilibrated ensemble, list rebuilds at their steady rate)
  chain1000    config 2: one CG chain, N = 1000 (``--precision fp32`` for the mixed engine)
  lj4096       config 1: LJ fluid, N = 4096, rho sigma^3 = 0.6
  lj262144     LJ fluid, N = 262144 (the largest LJ job of the serving stream)
  dna200       config 3: 3SPN2 double-stranded DNA, 200 bp (1198 beads), rise 3.38 A, G-JF dt = 0.02
  dna200x128   128 independent 200-bp duplexes in the replica kernel, FP32-bulk mode
  serve        config 5: a seeded Poisson stream (50 jobs/s) of LJ fluids (1e3..2.6e5 particles),
               CG chains (50..5000 beads) and 3SPN2 DNA (20..400 bp), 1e3..1e5 steps each, energy
               trace every 100 steps, through the continuous-batching server (``mjolnir_amd.serve``),
               beside the same jobs one at a time on the plain FP64 engine

The default run also serves the config-5 stream (``--serve-jobs``, 0 to skip) and reports its
p50/p99 job latency under "serving".

Force errors are measured against the FP64 GPU path on the same state (that path agrees with the
reference to 1e-10 per particle; tests/test_gpu_parity.py pins it).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s mixed vs FP64 on GPU at fixed force err; serve p50/p99 lat"
HBM_PEAK_GBS = 8000.0
# one wave64 VALU instruction per CU per clock (4 SIMDs x 16 lanes; packed FP32 is not used):
# 256 CUs x 2.4 GHz
VALU_PEAK_GINST = 256 * 2.4
CONFIG4_STEPS = 10000
SEGMENT = 100


def host_info():
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": cores, "cpu_model": model}


def host_cores() -> int:
    return max(1, host_info()["nproc"])


def _stratified(specs, k):
    order = sorted(range(len(specs)), key=lambda i: specs[i].n)
    step = max(1, len(order) // k)
    return [order[i] for i in range(step // 2, len(order), step)][:k]


def force_error(f, ref):
    """RMS and max per-particle force error relative to the RMS force."""
    import numpy as np
    rms = float(np.sqrt(np.mean(np.sum(ref ** 2, axis=1))))
    d = np.linalg.norm(f - ref, axis=1)
    return {"rms": float(np.sqrt(np.mean(d ** 2)) / rms), "max": float(d.max() / rms)}


def fp64_reference_force(spec, x):
    """Total force of the FP64 engine at positions x (agrees with the reference to 1e-10)."""
    import copy
    from mjolnir_amd import capi, workloads
    s = copy.deepcopy(workloads.zero_temperature(spec))
    s.pos = x
    e = capi.Engine(s, capi.FP64, with_integrator=False)
    e.calc_force()
    _, _, f = e.state()
    e.close()
    return f


CPU_NOTE = ("reference OpenMP build with the reference's release flags except -march: x86-64-v3 instead of "
            "native, so the binary runs on any AVX2 host")


# ------------------------------------------------------------------------------------------
# single system (engine)
# ------------------------------------------------------------------------------------------
def cpu_baseline_single(spec, target_s: float = 12.0):
    """The reference's OpenMP build (oracle/_ref/ref_driver_omp) on this host: a short probe,
    then a run sized to about target_s seconds of CPU work."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refrun
    threads = host_cores()
    if not refrun.available(omp=True):
        return None
    print(f"[bench] CPU baseline: probing the reference build on {spec.name} ({threads} threads)", file=sys.stderr,
          flush=True)
    probe = refrun.bench(spec, 50, threads=threads)
    steps = max(50, int(target_s * probe["particle_steps_per_s"] / spec.n))
    print(f"[bench] CPU baseline: {steps} steps", file=sys.stderr, flush=True)
    r = refrun.bench(spec, steps, threads=threads, timeout=300)
    return {"value": r["particle_steps_per_s"], "unit": "particle-steps/s", "cores": r["threads"],
            "kind": "reference", "host": host_info(), "build": CPU_NOTE,
            "sample": f"{spec.name}: {steps} {spec.integrator.type} steps of the reference OpenMP build "
                      f"(OMP_NUM_THREADS={r['threads']}), step loop only, {r['seconds']:.1f} s"}


def _time_engine(spec, precision, steps, warmup):
    from mjolnir_amd import capi
    eng = capi.Engine(spec, precision)
    eng.step(warmup)
    s0, r0 = eng.counters()
    t0 = time.perf_counter()
    eng.step(steps)                # returns after the device work is complete
    sec = time.perf_counter() - t0
    s1, r1 = eng.counters()
    assert s1 - s0 == steps
    return eng, sec, int(r1 - r0)


def bench_single(args, spec):
    from mjolnir_amd import capi
    prec = capi.FP32_FIXED if args.precision == "fp32" else capi.FP64
    eng, sec, rebuilds = _time_engine(spec, prec, args.steps, args.warmup)
    value = spec.n * args.steps / sec
    # the other precision on the same workload, and the force error of the mixed path
    other = capi.FP64 if prec == capi.FP32_FIXED else capi.FP32_FIXED
    e2, sec2, _ = _time_engine(spec, other, args.steps, args.warmup)
    e2.close()
    x, _, _ = eng.state()
    em = capi.Engine(_at(spec, x), capi.FP32_FIXED, with_integrator=False)
    em.calc_force()
    _, _, fm = em.state()
    em.close()
    ferr = force_error(fm, fp64_reference_force(spec, x))

    # dominant force kernel: the pair interaction with the longest kernel time
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
    tiled = (spec.boundary == "periodic" and g.partition == "CellList" and np_ >= 16384 and
             g.potential not in ("3SPN2BasePair", "3SPN2CrossStacking") and os.environ.get("MJG_NO_TILE") != "1")
    if tiled:
        kern = "k_pairs_tiled<mixed>" if prec == capi.FP32_FIXED else "k_pairs_tiled<fp64>"
    else:
        kern = ("k_tsp_pairs" if g.potential in ("3SPN2BasePair", "3SPN2CrossStacking") else
                ("k_pairs_mixed" if prec == capi.FP32_FIXED else "k_pairs"))
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None,
                "kernel": f"{kern} ({g.potential})", "avg_us": round(ms * 1e3, 3),
                "half_pairs": hp, "participants": np_}
    cpu = None if args.no_cpu_baseline else cpu_baseline_single(spec)
    v2 = spec.n * args.steps / sec2
    return {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec * 1e3 / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+i64fixed+f64" if prec == capi.FP32_FIXED else "f64",
        "data": "synthetic",
        "config": {"workload": spec.name, "particles": spec.n, "precision": args.precision,
                   "integrator": spec.integrator.type, "dt": spec.integrator.dt,
                   "rebuilds_in_timed_region": rebuilds},
        "other_precision": {"precision": "fp64" if prec == capi.FP32_FIXED else "fp32", "value": round(v2, 1),
                            "ratio_fp32_over_fp64": round((value / v2) if prec == capi.FP32_FIXED else (v2 / value), 3)},
        "force_error_fp32_vs_fp64": ferr,
        "roofline": roofline, "cpu_baseline": cpu,
    }


def _at(spec, x):
    import copy
    s = copy.deepcopy(spec)
    s.pos = x
    return s


# ------------------------------------------------------------------------------------------
# replica batch (headline, config 4)
# ------------------------------------------------------------------------------------------
def committed_profile(name: str, workload: str, segment_steps: int):
    """A per-launch figure of this workload from a committed rocprofv3 summary (profiles/)."""
    p = os.path.join(ROOT, "profiles", name)
    try:
        with open(p) as f:
            t = json.load(f)
    except OSError:
        return None
    if t.get("workload") != workload or t.get("segment_steps") != segment_steps:
        return None
    return t


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
    core (all the host's cores), on a stratified sample of the ensemble; steps per replica sized
    to about target_s s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refrun
    if not refrun.available(omp=True):
        return None
    procs = host_cores()
    # largest first, so the longest runs do not start last
    sample = sorted([specs[k] for k in _stratified(specs, 2 * procs)], key=lambda s: -s.n)
    tot_n = sum(s.n for s in sample)
    _, wall0, res0 = refrun.bench_many(sample, [200] * len(sample), procs)
    loop0 = max(r["seconds"] for r in res0)
    steps = max(200, int(200 * target_s / max(wall0, loop0, 1e-3)))
    ps, wall, _ = refrun.bench_many(sample, [steps] * len(sample), procs)
    return {"value": round(ps / wall, 1), "unit": "particle-steps/s", "cores": procs, "kind": "reference",
            "host": host_info(), "build": CPU_NOTE,
            "sample": f"{len(sample)} replicas of the ensemble (stratified by length, {tot_n} particles), "
                      f"{steps} {sample[0].integrator.type} steps each on the reference OpenMP build with "
                      f"OMP_NUM_THREADS=1, {procs} processes at a time (all host cores), step loops only, "
                      f"{wall:.1f} s wall"}


def fp64_engine_rate(specs, steps: int = 300):
    """The plain FP64 engine run replica by replica on a stratified sample (particle-steps/s)."""
    from mjolnir_amd import capi
    tps, tt = 0, 0.0
    for k in _stratified(specs, 12):
        s = specs[k]
        e = capi.Engine(s, capi.FP64)
        e.step(50)
        t = time.perf_counter()
        e.step(steps)
        tt += time.perf_counter() - t
        tps += s.n * steps
        e.close()
    return tps / tt


def run_batch(specs, precision, steps_job, warmup, steps):
    """The config-4 job (steps_job steps from the generated state), then `warmup` more steps and
    the timed `steps`.  Returns the batch, ids and the measured figures."""
    from mjolnir_amd import capi
    b = capi.Batch(precision)
    ids = [b.add(s) for s in specs]
    failed = list(b.initialize())
    t0 = time.perf_counter()
    failed += b.step(steps_job)
    sec_job = time.perf_counter() - t0
    ps_job, _ = b.counters()
    failed += b.step(warmup)
    ps0, rb0 = b.counters()
    t0 = time.perf_counter()
    failed += b.step(steps)         # returns after the device work is complete
    sec = time.perf_counter() - t0
    ps1, rb1 = b.counters()
    active = len(specs) - len(failed)
    return b, ids, {"failed": failed, "job_sec": sec_job, "job_value": ps_job / sec_job, "sec": sec,
                    "value": (ps1 - ps0) / sec,
                    "rebuilds_per_1000_replica_steps": round(1000.0 * (rb1 - rb0) / max(1, active * steps), 3)}


def bench_batch(args):
    import numpy as np
    from mjolnir_amd import capi, workloads
    t = time.perf_counter()
    specs = workloads.replica_ensemble(512)
    print(f"[bench] replicas512: {len(specs)} replicas, {sum(s.n for s in specs)} particles "
          f"(generated in {time.perf_counter() - t:.1f} s)", file=sys.stderr, flush=True)
    b, ids, m = run_batch(specs, capi.FP32_FIXED, args.job_steps, args.warmup, args.steps)
    print(f"[bench] FP32 batch: config-4 job {m['job_value']:.4g} particle-steps/s over {args.job_steps} steps; "
          f"timed {args.steps} steps in {m['sec']:.3f} s", file=sys.stderr, flush=True)

    # dominant kernel: the segment kernel (all size classes of one launch), HIP events
    ms, ps_launch = b.profile(SEGMENT, reps=20)
    alg = batch_alg_bytes(b, ids, specs, SEGMENT)
    counters = committed_profile("k_replica_segment_replicas512_counters.json", "replicas512", SEGMENT)
    traffic = committed_profile("k_replica_segment_replicas512_traffic.json", "replicas512", SEGMENT)
    kern_rate = ps_launch / (ms * 1e-3)
    hbm = {"achieved": round(alg / (ms * 1e-3) / 1e9, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 6),
           "traffic": traffic["traffic_bytes_per_launch"] if traffic else None,
           "counter_bytes_frac": (round(traffic["traffic_bytes_per_launch"] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 6)
                                  if traffic else None)}
    if counters:
        # the binding resource of this kernel: VALU issue (wave64 instructions per launch from the
        # committed SQ_INSTS_VALU pass, divided by the live launch time)
        ach = counters["valu_insts_per_launch"] / (ms * 1e-3) / 1e9
        roofline = {"bound": "valu", "achieved": round(ach, 3), "peak": VALU_PEAK_GINST, "unit": "G wave-instr/s",
                    "frac": round(ach / VALU_PEAK_GINST, 6), "traffic": hbm["traffic"], "hbm": hbm}
    else:
        roofline = dict(bound="hbm", **hbm)
    roofline.update({"kernel": "k_replica_segment<float> (3 size classes, one launch each per segment, concurrent)",
                     "avg_ms": round(ms, 4), "segment_steps": SEGMENT, "alg_bytes_per_launch": int(alg),
                     "kernel_particle_steps_per_s": round(kern_rate, 1)})

    # force error of the mixed path against the FP64 path at the current state, 16 stratified replicas
    errs = []
    for k in _stratified(specs, 16):
        x, _, _ = b.state(ids[k])
        _, _, f32 = b.evaluate(ids[k])
        errs.append(force_error(f32, fp64_reference_force(specs[k], x)))
    b.close()
    ferr = {"rms_worst": max(e["rms"] for e in errs), "rms_median": float(np.median([e["rms"] for e in errs])),
            "max_worst": max(e["max"] for e in errs), "sample": "16 replicas stratified by length, after "
            f"{args.job_steps + args.warmup + args.steps} steps"}
    per_config = {} if args.no_config_errors else config_force_errors()

    # the same ensemble in the FP64 replica kernel (isolates what precision buys), and the plain
    # FP64 engine one replica at a time
    m64 = None
    if not args.no_fp64_batch:
        b64, _, m64 = run_batch(specs, capi.FP64, args.job_steps, args.warmup, args.steps)
        b64.close()
    fp64 = None if args.no_fp64_sample else fp64_engine_rate(specs)
    cpu = None if args.no_cpu_baseline else cpu_baseline_batch(specs)
    return {
        "metric": METRIC, "value": round(m["value"], 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(m["sec"] * 1e3 / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+i64fixed+f64",
        "data": "synthetic",
        "config": {"workload": "replicas512", "replicas": len(specs), "particles": int(sum(s.n for s in specs)),
                   "lengths": "log-uniform [50, 2000]", "temperatures": "uniform [280, 400] K",
                   "precision": "fp32 pair/local, int64 fixed-point sums, fp64 integrator",
                   "integrator": "BAOABLangevin", "dt": 0.01, "segment_steps": SEGMENT,
                   "timed_after_steps": args.job_steps + args.warmup,
                   "rebuilds_per_1000_replica_steps": m["rebuilds_per_1000_replica_steps"],
                   "failed_replicas": len(m["failed"])},
        "config4_job": {"steps": args.job_steps, "seconds": round(m["job_sec"], 4),
                        "value": round(m["job_value"], 1), "unit": "particle-steps/s",
                        "note": "the config's whole 10^4-step run from the generated state, 100-step segments"},
        "force_error": ferr,
        "force_error_per_config": per_config,
        "fp64_batch": None if m64 is None else {
            "value": round(m64["value"], 1), "config4_job_value": round(m64["job_value"], 1),
            "unit": "particle-steps/s", "ratio_fp32_over_fp64": round(m["value"] / m64["value"], 3),
            "failed_replicas": len(m64["failed"]),
            "note": "the same ensemble and windows in the replica kernel's FP64 mode"},
        "fp64_engine": None if fp64 is None else {
            "value": round(fp64, 1), "unit": "particle-steps/s",
            "sample": "12 replicas (stratified by length), plain FP64 engine one at a time",
            "speedup_of_batch": round(m["value"] / fp64, 2)},
        "roofline": roofline, "cpu_baseline": cpu,
    }


def config_force_errors(thermalize: int = 2000):
    """RMS / max force error of the mixed path against the FP64 path, configs 1-3 after
    `thermalize` Langevin steps of the FP64 engine from their generated states (a generated chain
    or duplex sits near its native structure, where every bonded term is near its minimum and the
    net force is a small remainder that no relative error measures): config 1 (LJ 4096) and the
    262,144-particle LJ box on the mixed engine, config 2 (chain 1000) and config 3 (200-bp duplex)
    in the replica kernel and on the mixed engine."""
    from mjolnir_amd import capi, workloads
    out = {}
    cases = {"config1_lj4096": workloads.lj_fluid(n_side=16, seed=1),
             "config2_chain1000": workloads.cg_chain(n=1000, seed=2),
             "config3_dna200": workloads.dna_duplex(n_bp=200, seed=3),
             "lj262144": workloads.lj_fluid(n_side=64, seed=1)}
    for name, spec in cases.items():
        e64 = capi.Engine(spec, capi.FP64)
        e64.step(thermalize)
        x, v, _ = e64.state()
        e64.close()
        spec = _at(spec, x)
        spec.vel = v
        ref = fp64_reference_force(spec, spec.pos)
        e = capi.Engine(spec, capi.FP32_FIXED, with_integrator=False)
        e.calc_force()
        _, _, f = e.state()
        e.close()
        row = {"engine_fp32": force_error(f, ref)}
        if spec.n <= 2048:
            bb = capi.Batch(capi.FP32_FIXED)
            r = bb.add(spec)
            bb.initialize()
            _, _, fb = bb.evaluate(r)
            bb.close()
            row["batch_fp32"] = force_error(fb, ref)
        row["state"] = f"after {thermalize} FP64 Langevin steps"
        out[name] = row
    return out


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
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec * 1e3 / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+i64fixed+f64",
        "data": "synthetic",
        "config": {"workload": "dna200x128", "replicas": n_rep, "particles": int(sum(s.n for s in specs)),
                   "model": "3SPN2 duplex, 200 bp, random sequences, rise 3.38 A", "integrator": "GJFNVTLangevin",
                   "dt": 0.02, "precision": "fp32 pair/local, int64 fixed-point sums, fp64 integrator",
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
# serving stream (config 5)
# ------------------------------------------------------------------------------------------
def run_serving(n_jobs: int, seed: int = 7, baseline: bool = True):
    from mjolnir_amd import capi, serve
    t = time.perf_counter()
    jobs = serve.make_jobs(serve.job_trace(n_jobs, rate=50.0, seed=seed), procs=host_cores())
    print(f"[bench] serving: {n_jobs} jobs built in {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    out = serve.Server(workers=4).run(jobs)
    out.update({"rate_jobs_per_s": 50.0, "seed": seed,
                "mix": "LJ fluid N log-uniform [1e3, 2.6e5], CG chain 50..5000 beads, 3SPN2 DNA 20..400 bp; "
                       "steps log-uniform [1e3, 1e5]; energy trace every 100 steps",
                "lanes": "batch: replicas <= 2048 particles (FP32 + int64 fixed-point sums); "
                         "engine: mixed-precision engine per job (FP32 pairs + int64 fixed-point sums), "
                         "4 at a time"})
    print(f"[bench] serving: {json.dumps(out)}", file=sys.stderr, flush=True)
    if baseline:
        base = serve.one_at_a_time(jobs, capi.FP64)
        out["one_at_a_time_fp64"] = {k: base[k] for k in ("particle_steps_per_s", "makespan_s", "latency_p50_s",
                                                           "latency_p99_s", "failed_jobs")}
        out["one_at_a_time_fp64"]["note"] = "the same jobs and arrivals, one at a time on the plain FP64 engine"
        out["throughput_ratio_vs_one_at_a_time_fp64"] = round(out["particle_steps_per_s"] /
                                                              base["particle_steps_per_s"], 3)
        print(f"[bench] serving baseline: {json.dumps(out['one_at_a_time_fp64'])}", file=sys.stderr, flush=True)
    return out


def bench_serve(args):
    s = run_serving(args.serve_jobs)
    return {
        "metric": METRIC, "value": s["particle_steps_per_s"], "unit": "particle-steps/s",
        "n_gpus": 1, "steps": None, "warmup": 0, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+i64fixed+f64",
        "data": "synthetic",
        "config": {"workload": "serve", "jobs": args.serve_jobs, "rate_jobs_per_s": 50.0},
        "serving": s, "roofline": None, "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--job-steps", type=int, default=CONFIG4_STEPS,
                    help="replicas512: steps of the config-4 job timed before the warm-up and timed window")
    ap.add_argument("--workload", default="replicas512",
                    choices=["replicas512", "chain1000", "lj4096", "lj262144", "dna200", "dna200x128", "serve"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"],
                    help="single-system workloads: the engine's arithmetic (fp32: MJG_FP32_FIXED)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-sample", action="store_true")
    ap.add_argument("--no-config-errors", action="store_true")
    ap.add_argument("--no-fp64-batch", action="store_true",
                    help="replicas512: skip the FP64 replica-kernel run of the same ensemble")
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
