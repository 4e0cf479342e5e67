"""MsgPack checkpoints (MsgPackSaver.hpp / MsgPackLoader.hpp format).

CPU: the reference's own saver output (tests/golden/*_system.msg, written by oracle/_ref from
the same spec) decodes to the spec's state and the oracle's forces bit for bit, and our encoder
reproduces those files byte for byte.  GPU: an engine restored from its own checkpoint
continues the trajectory of the uninterrupted run; a batch replica continues it bit for bit."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
import make_golden
import oracle
from mjolnir_amd import checkpoint


@pytest.mark.parametrize("name", list(make_golden.checkpoint_cases().keys()))
def test_reference_checkpoint_decodes_and_reencodes_bytewise(name):
    spec = make_golden.checkpoint_cases()[name]()
    with open(os.path.join(GOLDEN, name + "_system.msg"), "rb") as f:
        data = f.read()
    d = checkpoint.decode_system(data)
    assert np.array_equal(d["mass"], spec.mass)
    assert np.array_equal(d["pos"], spec.pos)
    assert np.array_equal(d["vel"], spec.vel)
    o = oracle.OracleEngine(spec, with_integrator=False)
    f_ref, _ = o.force_and_energy()
    assert np.array_equal(d["force"], f_ref)
    assert d["names"] == ["X"] * spec.n and d["groups"] == ["NONE"] * spec.n
    assert d["attributes"] == spec.attributes
    periodic = spec.boundary == "periodic"
    assert (d["lower"] is not None) == periodic
    again = checkpoint.encode_system(d["mass"], d["pos"], d["vel"], d["force"],
                                     spec.lower if periodic else None, spec.upper if periodic else None,
                                     spec.attributes, d["names"], d["groups"], d["virial"])
    assert again == data


def test_large_counts_and_rng_state():
    # more than 15 particles and attributes switch the headers to array16 / map16
    n = 70
    rng = np.random.default_rng(1)
    x = rng.normal(size=(n, 3))
    attrs = {f"a{k:02d}": float(k) for k in range(17)}
    data = checkpoint.encode_system(np.ones(n), x, x, x, None, None, attrs)
    d = checkpoint.decode_system(data)
    assert np.array_equal(d["pos"], x) and d["attributes"] == attrs
    assert checkpoint.decode_rng(checkpoint.encode_rng(12345, 9876543210)) == (12345, 9876543210)


@pytest.mark.gpu
def test_engine_checkpoint_restart_continues_trajectory(tmp_path):
    from mjolnir_amd import capi, workloads
    spec = make_golden.perturb(workloads.cg_chain(n=200, seed=31), 0.1, 31)
    a = capi.Engine(spec)
    print("[gpu] checkpoint restart, chain N=200", flush=True)
    a.step(300)
    prefix = str(tmp_path / "ck")
    checkpoint.save_engine(a, prefix)
    x0, v0, f0 = a.state()
    d = checkpoint.decode_system(open(prefix + "_system.msg", "rb").read())
    assert np.array_equal(d["pos"], x0) and np.array_equal(d["vel"], v0) and np.array_equal(d["force"], f0)
    b = checkpoint.load_engine(spec, prefix)
    assert b.counters()[0] == 300
    a.step(100)
    b.step(100)
    xa, va, _ = a.state()
    xb, vb, _ = b.state()
    # same noise stream; only the neighbour-list history differs (rebuilt at the restart), which
    # changes the order of a few force sums
    assert np.abs(xa - xb).max() < 1e-8, np.abs(xa - xb).max()
    assert np.abs(va - vb).max() < 1e-6, np.abs(va - vb).max()


@pytest.mark.gpu
def test_underdamped_checkpoint_restart_continues_the_noise_stream(tmp_path):
    """UnderdampedLangevin keeps an acceleration that carries the noise of the last force call.
    A restored engine redraws it at the saved step count (as the reference's initialize does
    after loading), which is the draw the uninterrupted run made there, so the two runs continue
    together."""
    from mjolnir_amd import capi, workloads
    spec = make_golden.underdamped(make_golden.perturb(workloads.lj_fluid(n_side=6, seed=3), 0.1, 3), 1.0)
    a = capi.Engine(spec)
    print("[gpu] checkpoint restart, underdamped LJ N=216", flush=True)
    a.step(300)
    prefix = str(tmp_path / "ul")
    checkpoint.save_engine(a, prefix)
    b = checkpoint.load_engine(spec, prefix)
    a.step(100)
    b.step(100)
    xa, va, _ = a.state()
    xb, vb, _ = b.state()
    w = spec.upper - spec.lower
    d = xa - xb
    d = d - w * np.round(d / w)
    assert np.abs(d).max() < 1e-8, np.abs(d).max()
    assert np.abs(va - vb).max() < 1e-6, np.abs(va - vb).max()


@pytest.mark.gpu
def test_batch_replica_checkpoint_restart_is_bitwise(tmp_path):
    """A replica of the batch engine (FP32 mode: exact fixed-point force sums, so the rebuilt lists
    change nothing) saved at step 300 and restored into a new batch continues bit for bit, for
    BAOAB and UnderdampedLangevin."""
    from mjolnir_amd import capi, workloads
    specs = [make_golden.perturb(workloads.cg_chain(n=300, seed=51), 0.1, 51),
             make_golden.underdamped(make_golden.perturb(workloads.cg_chain(n=200, seed=52), 0.1, 52), 0.5)]
    a = capi.Batch(capi.FP32_FIXED)
    ids = [a.add(s) for s in specs]
    a.initialize()
    print("[gpu] batch checkpoint restart, chains N=300 (BAOAB), 200 (UnderdampedLangevin)", flush=True)
    a.step(300)
    b = capi.Batch(capi.FP32_FIXED)
    restored = []
    for k, (s, r) in enumerate(zip(specs, ids)):
        prefix = str(tmp_path / f"rep{k}")
        checkpoint.save_batch_replica(a, r, prefix)
        restored.append(checkpoint.load_batch_replica(b, s, prefix))
    a.step(100)
    b.step(100)
    for r, q in zip(ids, restored):
        xa, va, _ = a.state(r)
        xb, vb, _ = b.state(q)
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
        assert a.info(r)["step"] == b.info(q)["step"] == 400
