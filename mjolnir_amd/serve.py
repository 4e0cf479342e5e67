"""Serving loop: a stream of MD jobs on one GPU with continuous batching.

A job is a ``SystemSpec``, a step count and an arrival time.  Jobs are admitted while others are
on the device, and every job returns an energy trace (per-interaction energies and the kinetic
energy, as the reference's EnergyObserver writes them, EnergyObserver.hpp:89-150) every
``segment`` steps and at its last step.

Two lanes share the GPU:

* batch lane: jobs of at most ``batch_max_n`` particles whose interactions the replica kernel
  covers run as replicas of one ``capi.Batch`` (FP32 arithmetic, int64 fixed-point force sums,
  FP64 integrator).  Each segment advances every resident replica in one launch per size class
  and then records every replica's energies in one evaluation launch.  Between segments the
  finished replicas are retired and the waiting jobs admitted.
* engine lane: larger systems (and constrained ones) run on engines, one per job and ``workers``
  at a time, in ``engine_precision`` (default MJG_FP32_FIXED: FP32 pair arithmetic with int64
  fixed-point sums, FP64 local terms and integrator; MJG_FP64 for the plain FP64 path).  Each engine has its own HIP stream and host thread (the library's C
  calls release the GIL), so their segments overlap on the device with the batch lane's.

Admission checks the slot counts and a device-memory budget (``hbm_budget`` bytes against an
estimated footprint per job).  Noise depends only on a job's seed, step and particle index, and
a replica's arithmetic only on its own size, so a job's trajectory is the same alone and in any
mix; the GPU tests check this.
"""
from __future__ import annotations

import argparse
import json
import math
import queue
import threading
import time
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from .spec import SystemSpec

# interactions the replica kernel covers
_BATCH_LOCALS = {("BondLength", "Harmonic"), ("BondLength", "GoContact"), ("Contact", "GoContact"),
                 ("BondAngle", "Harmonic"), ("BondAngle", "Gaussian"), ("DihedralAngle", "ClementiDihedral"),
                 ("BondLength", "3SPN2Bond"), ("DihedralAngle", "Gaussian"), ("3SPN2BaseStacking", "3SPN2")}
_BATCH_GLOBALS = {"LennardJones", "ExcludedVolume", "DebyeHuckel", "3SPN2ExcludedVolume", "3SPN2BasePair",
                  "3SPN2CrossStacking"}

# footprint estimate per particle (state, backups, coefficients, lists) and per engine
_BYTES_PER_PARTICLE = 1024
_ENGINE_OVERHEAD = 64 << 20


@dataclass
class Job:
    id: int
    kind: str                 # "lj", "chain", "dna" or any label
    spec: SystemSpec
    steps: int
    arrival: float            # seconds after the start of the run
    # filled in by the server
    lane: str = ""
    admitted: float = math.nan
    finished: float = math.nan
    trace: List[Tuple[int, np.ndarray, float]] = field(default_factory=list)
    final_pos: Optional[np.ndarray] = None
    final_vel: Optional[np.ndarray] = None
    failed: str = ""          # why the job stopped early (non-finite state); "" for a finished job
    steps_done: int = 0

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def footprint(self) -> int:
        base = self.spec.n * _BYTES_PER_PARTICLE
        if self.lane == "engine":
            return base + _ENGINE_OVERHEAD
        # a batch replica holds each pair list at its bound, every participant pair (4 B an entry)
        return base + sum(4 * len(g.participants) * (len(g.participants) - 1) // 2 for g in self.spec.global_)


def batch_eligible(spec: SystemSpec, max_n: int = 2048) -> bool:
    return (spec.n <= max_n and len(spec.local) <= 8 and len(spec.global_) <= 4 and spec.constraint is None
            and spec.integrator is not None
            and spec.integrator.type in ("BAOABLangevin", "GJFNVTLangevin", "UnderdampedLangevin")
            and all((l.interaction, l.potential) in _BATCH_LOCALS for l in spec.local)
            and all(g.potential in _BATCH_GLOBALS for g in spec.global_))


# ------------------------------------------------------------------------------------------
# the job stream
# ------------------------------------------------------------------------------------------
def job_trace(n_jobs: int, rate: float = 50.0, seed: int = 7, lj_n=(1000, 260000), chain_n=(50, 5000),
              dna_bp=(20, 400), steps=(1000, 100000)):
    """Seeded Poisson arrivals at ``rate`` jobs/s; kind uniform over LJ fluid / CG chain / 3SPN2
    DNA; sizes and step counts log-uniform in the given ranges.  Returns a list of
    (arrival, kind, size, steps, job seed)."""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0

    def logu(lo, hi):
        return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))

    for k in range(n_jobs):
        t += rng.exponential(1.0 / rate)
        kind = ("lj", "chain", "dna")[int(rng.integers(3))]
        size = logu(*{"lj": lj_n, "chain": chain_n, "dna": dna_bp}[kind])
        out.append((t, kind, size, logu(*steps), int(seed) * 100003 + k))
    return out


def _make_spec(item) -> SystemSpec:
    from . import workloads
    _, kind, size, _, s = item
    if kind == "lj":
        return workloads.lj_fluid(n=size, seed=s)
    if kind == "chain":
        return workloads.cg_chain(n=size, seed=s, max_bend_deg=150.0)
    return workloads.dna_duplex(n_bp=size, seed=s)


def make_jobs(trace, procs: int = 8) -> List[Job]:
    """Build the jobs' systems (the requests' payloads) before the clock starts."""
    if procs > 1 and len(trace) > 1:
        import gc
        import multiprocessing as mp
        # forked children never call into the HIP runtime (see workloads.replica_ensemble)
        gc.collect()
        with mp.get_context("fork").Pool(min(procs, len(trace))) as pool:
            specs = pool.map(_make_spec, trace, chunksize=1)
    else:
        specs = [_make_spec(it) for it in trace]
    return [Job(id=k, kind=it[1], spec=sp, steps=it[3], arrival=it[0]) for k, (it, sp) in enumerate(zip(trace, specs))]


# ------------------------------------------------------------------------------------------
# the server
# ------------------------------------------------------------------------------------------
class Server:
    def __init__(self, hbm_budget: float = 200e9, batch_slots: int = 1024, workers: int = 4, segment: int = 100,
                 batch_max_n: int = 2048, device: int = 0, engine_precision: Optional[int] = None):
        from . import capi
        self.engine_precision = capi.FP32_FIXED if engine_precision is None else engine_precision
        self.hbm_budget = hbm_budget
        self.batch_slots = batch_slots
        self.workers = workers
        self.segment = segment
        self.batch_max_n = batch_max_n
        self.device = device
        self._lock = threading.Lock()
        self._used = 0
        self._errors: List[BaseException] = []

    # -- memory budget -------------------------------------------------------------------
    def _reserve(self, job: Job) -> bool:
        with self._lock:
            if self._used + job.footprint > self.hbm_budget and self._used > 0:
                return False
            self._used += job.footprint
            return True

    def _release(self, job: Job):
        with self._lock:
            self._used -= job.footprint

    # -- engine lane -------------------------------------------------------------------------
    def _engine_worker(self, q: "queue.Queue", t0: float):
        from . import capi
        while True:
            job = q.get()
            if job is None:
                return
            try:
                job.admitted = time.perf_counter() - t0
                e = capi.Engine(job.spec, self.engine_precision, device=self.device)
                try:
                    self._record_engine(job, e, 0)
                    done = 0
                    while done < job.steps:
                        k = min(self.segment, job.steps - done)
                        e.step(k)
                        done += k
                        job.steps_done = done
                        self._record_engine(job, e, done)
                    x, v, _ = e.state()
                    job.final_pos, job.final_vel = x, v
                except capi.MjgError as exc:
                    # a job whose state stopped being finite fails alone; the server goes on
                    if exc.status != capi.NAN_DETECTED:
                        raise
                    job.failed = str(exc)
                finally:
                    e.close()
                job.finished = time.perf_counter() - t0
            except BaseException as exc:   # reported by run()
                with self._lock:
                    self._errors.append(exc)
            finally:
                self._release(job)

    @staticmethod
    def _record_engine(job: Job, e, step: int):
        per, _ = e.energies()
        job.trace.append((step, np.asarray(per, dtype=np.float64), e.kinetic_energy()))

    def _fail_replicas(self, batch, resident, bad, tnow):
        """Replicas the library stopped with non-finite state: their jobs fail alone (retired,
        budget released); every other resident replica keeps running."""
        for r in bad:
            j = resident.pop(r)[0]
            j.steps_done = batch.info(r)["step"]
            j.failed = f"non-finite state at step {j.steps_done}"
            batch.retire(r)
            j.finished = tnow
            self._release(j)

    # -- main loop ---------------------------------------------------------------------------
    def run(self, jobs: List[Job]) -> dict:
        """Serve ``jobs`` (arrival times relative to the start of this call); returns a summary.
        The jobs are filled in with their lane, admission/finish times, energy traces and final
        states."""
        from . import capi
        pending = sorted(jobs, key=lambda j: j.arrival)
        for j in pending:
            j.lane = "batch" if batch_eligible(j.spec, self.batch_max_n) else "engine"
        eq: "queue.Queue" = queue.Queue()
        t0 = time.perf_counter()
        threads = [threading.Thread(target=self._engine_worker, args=(eq, t0), daemon=True)
                   for _ in range(self.workers)]
        for th in threads:
            th.start()
        batch = capi.Batch(capi.FP32_FIXED, device=self.device) if any(j.lane == "batch" for j in pending) else None
        waiting_b: List[Job] = []
        waiting_e: List[Job] = []
        resident = {}                   # replica id -> [job, steps done]
        k = 0
        try:
            while True:
                now = time.perf_counter() - t0
                while k < len(pending) and pending[k].arrival <= now:
                    (waiting_b if pending[k].lane == "batch" else waiting_e).append(pending[k])
                    k += 1
                # engine lane: hand over every arrived job the budget admits
                still = []
                for j in waiting_e:
                    if self._reserve(j):
                        eq.put(j)
                    else:
                        still.append(j)
                waiting_e = still
                # batch lane: admit between segments
                fresh = []
                still = []
                for j in waiting_b:
                    if len(resident) + len(fresh) < self.batch_slots and self._reserve(j):
                        fresh.append(j)
                    else:
                        still.append(j)
                waiting_b = still
                if fresh:
                    ids = [batch.add(j.spec) for j in fresh]
                    bad = batch.initialize()
                    tnow = time.perf_counter() - t0
                    for r, j in zip(ids, fresh):
                        j.admitted = tnow
                        resident[r] = [j, 0]
                        if r not in bad:
                            batch.set_steps(r, j.steps)
                    self._fail_replicas(batch, resident, bad, tnow)
                    batch.observe()
                    for r, j in zip(ids, fresh):
                        per, ke = batch.energies(r)
                        j.trace.append((0, per.copy(), ke))
                if resident:
                    bad = batch.advance(self.segment)
                    self._fail_replicas(batch, resident, bad, time.perf_counter() - t0)
                    batch.observe()
                    done_ids = []
                    for r, rec in resident.items():
                        j = rec[0]
                        rec[1] = min(j.steps, rec[1] + self.segment)
                        j.steps_done = rec[1]
                        per, ke = batch.energies(r)
                        j.trace.append((rec[1], per.copy(), ke))
                        if rec[1] >= j.steps:
                            done_ids.append(r)
                    tnow = time.perf_counter() - t0
                    for r in done_ids:
                        j = resident.pop(r)[0]
                        x, v, _ = batch.state(r)
                        j.final_pos, j.final_vel = x, v
                        batch.retire(r)
                        j.finished = tnow
                        self._release(j)
                    continue
                with self._lock:
                    if self._errors:
                        break
                if k >= len(pending) and not waiting_b and not waiting_e:
                    break
                # idle: wait for the next arrival (or for memory to free up)
                nxt = pending[k].arrival if k < len(pending) else now + 0.002
                time.sleep(min(0.002, max(0.0, nxt - (time.perf_counter() - t0))))
        finally:
            for _ in threads:
                eq.put(None)
            for th in threads:
                th.join()
            if batch is not None:
                batch.close()
        if self._errors:
            raise self._errors[0]
        return summarize(jobs)


def one_at_a_time(jobs: List[Job], precision: int, segment: int = 100, device: int = 0) -> dict:
    """The same jobs run one after another on single engines of the given precision, each with its
    energy trace every `segment` steps, as a client without the server would run them: every job
    waits for the ones before it.  Arrival times are kept, so latencies include that wait."""
    from . import capi
    t0 = time.perf_counter()
    done = []
    for job in sorted(jobs, key=lambda j: j.arrival):
        now = time.perf_counter() - t0
        if now < job.arrival:
            time.sleep(job.arrival - now)
        j = Job(id=job.id, kind=job.kind, spec=job.spec, steps=job.steps, arrival=job.arrival, lane="engine")
        j.admitted = time.perf_counter() - t0
        e = capi.Engine(j.spec, precision, device=device)
        try:
            Server._record_engine(j, e, 0)
            k = 0
            while k < j.steps:
                n = min(segment, j.steps - k)
                e.step(n)
                k += n
                j.steps_done = k
                Server._record_engine(j, e, k)
        except capi.MjgError as exc:
            if exc.status != capi.NAN_DETECTED:
                raise
            j.failed = str(exc)
        finally:
            e.close()
        j.finished = time.perf_counter() - t0
        done.append(j)
    return summarize(done)


def summarize(jobs: List[Job]) -> dict:
    """Aggregate throughput (particle-steps of the work done, failed jobs counting the steps they
    took) and latency percentiles over the finished jobs."""
    ok = [j for j in jobs if not j.failed]
    lat = np.array([j.latency for j in ok] or [math.nan])
    ps = float(sum(j.spec.n * (j.steps if not j.failed else j.steps_done) for j in jobs))
    makespan = max(j.finished for j in jobs)
    last_arrival = max(j.arrival for j in jobs)
    out = {
        "jobs": len(jobs),
        "particle_steps": ps,
        "makespan_s": round(makespan, 4),
        "particle_steps_per_s": round(ps / makespan, 1),
        "offered_particle_steps_per_s": round(ps / max(last_arrival, 1e-9), 1),
        "latency_p50_s": round(float(np.percentile(lat, 50)), 4),
        "latency_p99_s": round(float(np.percentile(lat, 99)), 4),
        "drain_after_last_arrival_s": round(makespan - last_arrival, 4),
        "failed_jobs": len(jobs) - len(ok),
    }
    for lane in ("batch", "engine"):
        sel = [j for j in ok if j.lane == lane]
        if sel:
            l = np.array([j.latency for j in sel])
            out[lane] = {"jobs": len(sel), "particle_steps": float(sum(j.spec.n * j.steps for j in sel)),
                         "latency_p50_s": round(float(np.percentile(l, 50)), 4),
                         "latency_p99_s": round(float(np.percentile(l, 99)), 4)}
    return out


def main():
    ap = argparse.ArgumentParser(description="serve a seeded stream of MD jobs on one GPU")
    ap.add_argument("--jobs", type=int, default=100)
    ap.add_argument("--rate", type=float, default=50.0)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--segment", type=int, default=100)
    args = ap.parse_args()
    jobs = make_jobs(job_trace(args.jobs, args.rate, args.seed))
    print(json.dumps(Server(workers=args.workers, segment=args.segment).run(jobs)), flush=True)


if __name__ == "__main__":
    main()
