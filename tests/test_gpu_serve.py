"""Serving loop on the GPU: a small mixed stream through both lanes.

Bars:
* every job's result is the same alone and in the mix: final positions and velocities bit for
  bit; energy-trace entries at the same steps, equal within 1e-12 relative (the replica kernel
  sums its energies with LDS atomics, whose order may vary; forces use exact fixed-point sums);
* jobs overlap on the device: engine-lane jobs run while batch-lane replicas are resident;
* the batch lane's step-0 energies agree with the FP64 engine's within 1e-4 relative
  (FP32 arithmetic).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _jobs():
    from mjolnir_amd import workloads
    from mjolnir_amd.serve import Job
    rows = [
        ("chain", lambda: workloads.cg_chain(n=80, seed=11, max_bend_deg=150.0), 300, 0.0),
        ("dna", lambda: workloads.dna_duplex(n_bp=12, seed=14), 300, 0.002),
        ("chain", lambda: workloads.cg_chain(n=300, seed=12, max_bend_deg=150.0), 450, 0.004),
        ("chain", lambda: workloads.cg_chain(n=2100, seed=15, max_bend_deg=150.0), 200, 0.006),
        ("lj", lambda: workloads.lj_fluid(n=1000, seed=13), 250, 0.008),
        ("chain", lambda: workloads.cg_chain(n=600, seed=16, max_bend_deg=150.0), 350, 0.012),
    ]
    return [Job(id=k, kind=kd, spec=mk(), steps=st, arrival=a) for k, (kd, mk, st, a) in enumerate(rows)]


def _fresh(job):
    from mjolnir_amd.serve import Job
    return Job(id=job.id, kind=job.kind, spec=job.spec, steps=job.steps, arrival=0.0)


def test_serving_alone_vs_mix():
    from mjolnir_amd.serve import Server
    jobs = _jobs()
    print("[gpu] serving mix of", len(jobs), "jobs", flush=True)
    summary = Server(workers=2).run(jobs)
    print("[gpu] mix summary", summary, flush=True)
    lanes = {j.lane for j in jobs}
    assert lanes == {"batch", "engine"}
    for j in jobs:
        steps = [t[0] for t in j.trace]
        want = list(range(0, j.steps, 100)) + [j.steps]
        assert steps == want, (j.id, steps)
        assert np.isfinite(j.final_pos).all()
    # overlap: an engine job was on the device while a batch replica was resident
    eng = [j for j in jobs if j.lane == "engine"]
    bat = [j for j in jobs if j.lane == "batch"]
    assert any(e.admitted < b.finished and b.admitted < e.finished for e in eng for b in bat)
    for j in jobs:
        alone = _fresh(j)
        print(f"[gpu] job {j.id} ({j.kind}, N={j.spec.n}, {j.lane}) alone", flush=True)
        Server(workers=1).run([alone])
        assert alone.lane == j.lane
        assert np.array_equal(alone.final_pos, j.final_pos), j.id
        assert np.array_equal(alone.final_vel, j.final_vel), j.id
        assert [t[0] for t in alone.trace] == [t[0] for t in j.trace]
        for (s, e1, k1), (_, e2, k2) in zip(alone.trace, j.trace):
            scale = np.maximum(np.abs(e1), 1.0)
            assert np.all(np.abs(e1 - e2) <= 1e-12 * scale), (j.id, s, e1, e2)
            assert abs(k1 - k2) <= 1e-12 * max(abs(k1), 1.0), (j.id, s, k1, k2)


def test_batch_lane_energies_vs_fp64_engine():
    from mjolnir_amd import capi
    from mjolnir_amd.serve import Server
    job = _jobs()[2]      # 300-bead chain, batch lane
    Server(workers=1).run([job])
    assert job.lane == "batch"
    e = capi.Engine(job.spec, capi.FP64, with_integrator=False)
    per, _ = e.energies()
    e.close()
    s0, e0, k0 = job.trace[0]
    assert s0 == 0
    scale = np.maximum(np.abs(per), 1.0)
    assert np.all(np.abs(e0 - per) <= 1e-4 * scale), (e0, per)
    m = job.spec.mass
    ke = 0.5 * float(np.sum(m * np.sum(job.spec.vel ** 2, axis=1)))
    assert abs(k0 - ke) <= 1e-12 * ke
