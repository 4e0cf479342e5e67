"""Host logic of the serving loop (no GPU): the seeded job stream, lane choice, summaries."""
import math

import numpy as np

from mjolnir_amd import serve, workloads


def test_job_trace_is_seeded_and_in_range():
    a = serve.job_trace(300, rate=50.0, seed=7)
    b = serve.job_trace(300, rate=50.0, seed=7)
    assert a == b
    assert serve.job_trace(300, seed=8) != a
    t = np.array([x[0] for x in a])
    assert np.all(np.diff(t) > 0)
    # Poisson at 50 jobs/s: mean gap 20 ms (300 gaps: standard error ~1.2 ms)
    assert abs(np.mean(np.diff(np.concatenate([[0.0], t]))) - 0.02) < 0.005
    kinds = [x[1] for x in a]
    for k in ("lj", "chain", "dna"):
        assert 60 < kinds.count(k) < 140
    for _, kind, size, steps, _ in a:
        lo, hi = {"lj": (1000, 260000), "chain": (50, 5000), "dna": (20, 400)}[kind]
        assert lo <= size <= hi
        assert 1000 <= steps <= 100000
    # log-uniform sizes: the median is near the geometric mean of the range
    lj = sorted(x[2] for x in a if x[1] == "lj")
    assert 0.5 * math.sqrt(1000 * 260000) < lj[len(lj) // 2] < 2.0 * math.sqrt(1000 * 260000)


def test_lane_choice():
    assert serve.batch_eligible(workloads.cg_chain(n=120, seed=3, max_bend_deg=150.0))
    assert serve.batch_eligible(workloads.lj_fluid(n=1500, seed=3))
    assert not serve.batch_eligible(workloads.lj_fluid(n=3000, seed=3))
    assert serve.batch_eligible(workloads.dna_duplex(n_bp=20, seed=3))
    assert not serve.batch_eligible(workloads.dna_duplex(n_bp=400, seed=3))      # 2398 beads
    assert not serve.batch_eligible(workloads.constrained_chain(60, seed=3))


def test_summary():
    jobs = []
    for k in range(10):
        j = serve.Job(id=k, kind="chain", spec=workloads.lj_fluid(n=1000, seed=1), steps=100, arrival=0.1 * k)
        j.lane = "batch" if k % 2 else "engine"
        j.finished = j.arrival + 1.0 + k
        jobs.append(j)
    s = serve.summarize(jobs)
    assert s["jobs"] == 10
    assert s["particle_steps"] == 10 * 1000 * 100
    assert abs(s["makespan_s"] - (0.9 + 10.0)) < 1e-9
    assert abs(s["latency_p50_s"] - 5.5) < 1e-9
    assert s["batch"]["jobs"] == 5 and s["engine"]["jobs"] == 5
