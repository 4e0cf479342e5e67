"""Batched replicas (one persistent workgroup per replica) against the FP64 oracle.

Bars:
* mixed precision (FP32 arithmetic, int64 fixed-point sums): RMS force error over all
  particles <= 1e-5 of the RMS force, per replica, against the FP64 oracle on the same inputs;
  per-interaction energies within 1e-5 relative (floor 1e-4 kcal/mol);
* FP64 batch: forces within 1e-9 (|F_ref,i| + F_rms) per particle (the fixed-point sums round
  each contribution to 2^-32 kcal/mol/A); short zero-temperature trajectories within 1e-8 A;
* a replica's trajectory is bit-identical alone and in a mix of other replicas;
* a list-capacity overflow is recovered by re-running the segment with a larger list.
"""
import copy
import os

import numpy as np
import pytest

from conftest import base_pair_energy_paths, force_tolerance
import make_golden

pytestmark = pytest.mark.gpu


def _chains(sizes, seed0=100, temps=None):
    from mjolnir_amd import workloads
    out = []
    for k, n in enumerate(sizes):
        T = 300.0 if temps is None else temps[k]
        s = make_golden.perturb(workloads.cg_chain(n=n, seed=seed0 + k, T=T), 0.1, seed0 + k)
        out.append(s)
    return out


def _batch(specs, precision):
    from mjolnir_amd import capi
    print(f"[gpu] batch of {len(specs)} replicas, precision {precision}", flush=True)
    b = capi.Batch(precision)
    ids = [b.add(s) for s in specs]
    b.initialize()
    return b, ids


def test_mixed_precision_force_error_vs_oracle():
    import oracle
    from mjolnir_amd import capi
    specs = _chains([60, 250, 700, 1000, 1500], temps=[280.0, 320.0, 360.0, 300.0, 400.0])
    b, ids = _batch(specs, capi.FP32_FIXED)
    for spec, r in zip(specs, ids):
        o = oracle.OracleEngine(spec, with_integrator=False)
        f_ref, _ = o.force_and_energy()
        e_ref = [o.energy_of(k) for k in range(len(spec.interaction_names()))]
        per, tot, f = b.evaluate(r)
        rms_ref = np.sqrt(np.mean(np.sum(f_ref ** 2, axis=1)))
        rms_err = np.sqrt(np.mean(np.sum((f - f_ref) ** 2, axis=1)))
        print(f"[gpu] N={spec.n}: RMS force error {rms_err / rms_ref:.2e} (relative), "
              f"max {np.abs(f - f_ref).max():.2e}", flush=True)
        assert rms_err <= 1e-5 * rms_ref
        for k, (a, c) in enumerate(zip(per, e_ref)):
            assert abs(a - c) <= 1e-5 * abs(c) + 1e-4, (spec.interaction_names()[k], a, c)


def test_fp64_batch_forces_vs_oracle():
    import oracle
    from mjolnir_amd import capi
    specs = _chains([80, 400])
    b, ids = _batch(specs, capi.FP64)
    for spec, r in zip(specs, ids):
        o = oracle.OracleEngine(spec, with_integrator=False)
        f_ref, _ = o.force_and_energy()
        _, _, f = b.evaluate(r)
        err = np.linalg.norm(f - f_ref, axis=1)
        assert np.all(err <= force_tolerance(f_ref, 1e-9)), err.max()


def test_fp64_batch_zero_temperature_trajectory_vs_oracle():
    """The 1200-bead chain's excluded-volume list is built on the cell grid and its
    Debye-Hueckel list by the sliced row build; both are rebuilt during the run with the
    replica kernel's own skin, so a pair missing from a rebuilt list would show here."""
    import oracle
    from mjolnir_amd import capi, workloads
    specs = [workloads.zero_temperature(s) for s in _chains([90, 300, 1200])]
    b, ids = _batch(specs, capi.FP64)
    b.step(200)
    for spec, r in zip(specs, ids):
        o = oracle.OracleEngine(spec)
        o.step(200)
        x, v, _ = b.state(r)
        xo, vo, _ = o.state()
        assert np.abs(x - xo).max() < 1e-8, np.abs(x - xo).max()
        assert b.info(r)["step"] == 200
        if spec.n == 1200:
            assert b.info(r)["rebuilds"] > len(spec.global_), b.info(r)   # rebuilt after set-up


def test_fp64_batch_langevin_and_gjf_vs_oracle_same_noise():
    import oracle
    from mjolnir_amd import capi
    a, c = _chains([120, 200], seed0=300)
    c = copy.deepcopy(c)
    c.integrator.type = "GJFNVTLangevin"
    c.integrator.coefs = np.full(c.n, 0.1)
    b, ids = _batch([a, c], capi.FP64)
    b.step(120)
    for spec, r in zip([a, c], ids):
        o = oracle.OracleEngine(spec)
        o.step(120)
        x, v, _ = b.state(r)
        xo, vo, _ = o.state()
        assert np.abs(x - xo).max() < 1e-7
        assert np.abs(v - vo).max() < 1e-7


def test_alone_equals_in_the_mix_bitwise():
    from mjolnir_amd import capi
    target = _chains([333], seed0=500)[0]
    others = _chains([50, 700, 1200, 90, 410], seed0=600)
    b1, (r1,) = _batch([target], capi.FP32_FIXED)
    b2, ids = _batch(others[:2] + [target] + others[2:], capi.FP32_FIXED)
    b1.step(250)
    b2.step(250)
    x1, v1, f1 = b1.state(r1)
    x2, v2, f2 = b2.state(ids[2])
    assert np.array_equal(x1, x2) and np.array_equal(v1, v2) and np.array_equal(f1, f2)


def test_list_overflow_is_recovered(monkeypatch):
    from mjolnir_amd import capi
    specs = _chains([150, 420], seed0=700)
    ref, rids = _batch(specs, capi.FP32_FIXED)
    monkeypatch.setenv("MJG_BATCH_INITIAL_CAP", "8")
    small, sids = _batch(specs, capi.FP32_FIXED)
    ref.step(200)
    small.step(200)
    for a, c in zip(rids, sids):
        xa, _, _ = ref.state(a)
        xc, _, _ = small.state(c)
        # both runs build the same lists at the same steps, so they agree exactly
        assert np.array_equal(xa, xc)
        assert small.info(c)["footprint"] > 0


def test_thermostat_mixed_precision_kinetic_temperature():
    """Kinetic temperature of 8 + 8 LJ-fluid replicas (216 particles, gamma = 0.5) at two
    temperatures, pooled per temperature.  The FP64 oracle gives 0.998 for this set-up (BAOAB's
    kinetic-temperature bias at dt = 0.01 is below 0.5 % here).  The config-4 chains, which sit a
    few per cent above 1 in either precision, are compared mixed against FP64 in
    test_gpu_thermostat.py.
    The kinetic energy of 216 particles fluctuates by sqrt(2 / (3 N)) = 5.5 % and decorrelates
    within about 100 steps; 160 samples 50 steps apart from 8 replicas give a pooled mean with a
    spread near 0.3 %; the bar is 2 %."""
    from mjolnir_amd import capi, workloads
    temps = [300.0] * 8 + [350.0] * 8
    specs = [workloads.lj_fluid(n_side=6, seed=900 + k, T=t, gamma=0.5) for k, t in enumerate(temps)]
    b, ids = _batch(specs, capi.FP32_FIXED)
    b.step(4000)
    ratio = {300.0: [], 350.0: []}
    for _ in range(160):
        b.step(50)
        for r, s in zip(ids, specs):
            _, v, _ = b.state(r)
            T = s.attributes["temperature"]
            ratio[T].append(0.5 * np.sum(s.mass[:, None] * v * v) / (1.5 * s.n * workloads.KB_KCAL * T))
    for T, x in ratio.items():
        print(f"[gpu] T={T}: <KE>/(3/2 N kB T) = {np.mean(x):.4f}", flush=True)
        assert abs(np.mean(x) - 1.0) < 0.02, (T, np.mean(x))


def test_periodic_lj_batch_vs_oracle():
    """Periodic boundary (minimum image, wrapped positions, grid with wrap-around) in both
    precisions: FP64 forces at 1e-9, FP32 RMS error at 1e-5, FP64 zero-T trajectory."""
    import oracle
    from mjolnir_amd import capi, workloads
    specs = [make_golden.perturb(workloads.lj_fluid(n_side=k, seed=20 + k), 0.2, 20 + k) for k in (6, 12)]
    for prec in (capi.FP64, capi.FP32_FIXED):
        b, ids = _batch(specs, prec)
        for spec, r in zip(specs, ids):
            o = oracle.OracleEngine(spec, with_integrator=False)
            f_ref, _ = o.force_and_energy()
            _, _, f = b.evaluate(r)
            if prec == capi.FP64:
                err = np.linalg.norm(f - f_ref, axis=1)
                assert np.all(err <= force_tolerance(f_ref, 1e-9)), err.max()
            else:
                rms_ref = np.sqrt(np.mean(np.sum(f_ref ** 2, axis=1)))
                rms_err = np.sqrt(np.mean(np.sum((f - f_ref) ** 2, axis=1)))
                assert rms_err <= 1e-5 * rms_ref, rms_err / rms_ref
    zt = [workloads.zero_temperature(s) for s in specs]
    b, ids = _batch(zt, capi.FP64)
    b.step(200)
    for spec, r in zip(zt, ids):
        o = oracle.OracleEngine(spec)
        o.step(200)
        x, _, _ = b.state(r)
        xo, _, _ = o.state()
        d = np.abs(x - xo)
        d = np.minimum(d, (spec.upper - spec.lower) - d)   # wrapped coordinates
        assert d.max() < 1e-8, d.max()


def _dna_specs():
    from mjolnir_amd import workloads
    return [make_golden.perturb(workloads.dna_duplex(12, seed=5, rise=3.6), 0.2, 11),
            make_golden.perturb(workloads.dna_duplex(200, seed=3, rise=3.6), 0.1, 12),
            workloads.dna_duplex(200, seed=3)]          # config 3 as stated: B-form, rise 3.38 A


def test_dna_batch_forces_vs_oracle():
    """3SPN2 in the replica kernel: FP32 mode against the FP64 oracle at the mixed-precision bar
    (RMS force error <= 1e-5 of the RMS force, energies 1e-5), FP64 mode at 1e-9 per particle."""
    import oracle
    from mjolnir_amd import capi
    specs = _dna_specs()
    b32, ids32 = _batch(specs, capi.FP32_FIXED)
    b64, ids64 = _batch(specs, capi.FP64)
    for spec, r32, r64 in zip(specs, ids32, ids64):
        o = oracle.OracleEngine(spec, with_integrator=False)
        f_ref, e_tot = o.force_and_energy()
        e_ref = [o.energy_of(k) for k in range(len(spec.interaction_names()))]
        bp = base_pair_energy_paths(spec.interaction_names(), e_ref, e_tot)
        per, _, f = b32.evaluate(r32)
        rms_ref = np.sqrt(np.mean(np.sum(f_ref ** 2, axis=1)))
        rms_err = np.sqrt(np.mean(np.sum((f - f_ref) ** 2, axis=1)))
        per64, _, f64 = b64.evaluate(r64)
        err = np.linalg.norm(f64 - f_ref, axis=1)
        print(f"[gpu] DNA N={spec.n}: FP32 RMS force error {rms_err / rms_ref:.2e}; FP64 worst force error "
              f"{np.max(err / force_tolerance(f_ref, 1.0)):.2e} of the per-particle scale", flush=True)
        for k, name in enumerate(spec.interaction_names()):
            print(f"[gpu]   {name}: reference {e_ref[k]:.10g}, FP32 {per[k]:.10g}, FP64 {per64[k]:.10g}", flush=True)
        assert rms_err <= 1e-5 * rms_ref
        for k, (a, c) in enumerate(zip(per, e_ref)):
            assert abs(a - c) <= 1e-5 * abs(c) + 1e-4, (spec.interaction_names()[k], a, c)
        assert np.all(err <= force_tolerance(f_ref, 1e-9)), err.max()
        for k, (a, c) in enumerate(zip(per64, e_ref)):
            refs = bp.get(k, (c,))
            assert min(abs(a - r) for r in refs) <= 1e-9 * max(abs(c), 1.0), (spec.interaction_names()[k], a, refs)


def test_dna_batch_fp64_trajectory_and_mix():
    """FP64 replica kernel on the 12-bp duplex: 150 zero-temperature G-JF steps against the
    oracle; and a DNA replica in a batch with chains (two kernel variants in one class) ends
    bit-identical to the same replica alone."""
    import oracle
    from mjolnir_amd import capi, workloads
    dna = workloads.zero_temperature(_dna_specs()[0])
    b, (r,) = _batch([dna], capi.FP64)
    b.step(150)
    o = oracle.OracleEngine(dna)
    o.step(150)
    x, _, _ = b.state(r)
    xo, _, _ = o.state()
    assert np.abs(x - xo).max() < 1e-8, np.abs(x - xo).max()
    target = _dna_specs()[0]
    b1, (r1,) = _batch([target], capi.FP32_FIXED)
    b2, ids = _batch(_chains([60, 90], seed0=900) + [target] + _chains([180], seed0=950), capi.FP32_FIXED)
    b1.step(200)
    b2.step(200)
    x1, v1, _ = b1.state(r1)
    x2, v2, _ = b2.state(ids[2])
    assert np.array_equal(x1, x2) and np.array_equal(v1, v2)


def test_fp64_batch_underdamped_vs_oracle_same_noise():
    """UnderdampedLangevin in the replica kernel (FP64 mode): the stored acceleration, its noise
    at the count after each step, the drift and both velocity updates against the oracle with
    the same noise stream; a periodic LJ box and a chain in one batch."""
    import oracle
    from mjolnir_amd import capi, workloads
    specs = [make_golden.underdamped(make_golden.perturb(workloads.lj_fluid(n_side=6, seed=41), 0.1, 41), 1.0),
             make_golden.underdamped(_chains([150], seed0=420)[0], 0.5)]
    b, ids = _batch(specs, capi.FP64)
    b.step(200)
    for spec, r in zip(specs, ids):
        o = oracle.OracleEngine(spec)
        o.step(200)
        x, v, _ = b.state(r)
        xo, vo, _ = o.state()
        d = x - xo
        if spec.boundary == "periodic":
            w = spec.upper - spec.lower
            d = d - w * np.round(d / w)
        assert np.abs(d).max() < 1e-7, (spec.name, np.abs(d).max())
        assert np.abs(v - vo).max() < 1e-7, (spec.name, np.abs(v - vo).max())


def test_nonfinite_replica_is_isolated():
    """A replica whose state stops being finite is stopped alone: one whose bond force overflows the
    fixed point after its first step (a particle launched at 1e300 A per time unit; its poisoned
    force sum is seen when the next step reads it), and one with two coincident LJ particles (no
    finite force: flagged at the first evaluation).  Every other replica of the mix ends bitwise equal to its
    run alone, and the failed ones report their status."""
    from mjolnir_amd import capi, workloads
    good = _chains([150, 420, 900], seed0=800)
    blow = copy.deepcopy(_chains([300], seed0=810)[0])
    blow.vel[5] = [1e300, 0.0, 0.0]
    coinc = workloads.lj_fluid(n_side=4, seed=3)
    coinc.pos[1] = coinc.pos[0]
    b, ids = _batch([good[0], blow, good[1], coinc, good[2]], capi.FP32_FIXED)
    rb, rc = ids[1], ids[3]
    assert rc in b.failed and rb not in b.failed          # no finite force at set-up
    assert b.step(250) == [rb]
    assert b.status(rb) == capi.NAN_DETECTED and b.status(rc) == capi.NAN_DETECTED
    assert b.info(rb)["step"] == 1 and b.info(rc)["step"] == 0
    for spec, r in zip(good, [ids[0], ids[2], ids[4]]):
        assert b.status(r) == capi.OK and b.info(r)["step"] == 250
        a, (ra,) = _batch([spec], capi.FP32_FIXED)
        a.step(250)
        xa, va, fa = a.state(ra)
        xm, vm, fm = b.state(r)
        assert np.array_equal(xa, xm) and np.array_equal(va, vm) and np.array_equal(fa, fm)
    b.retire(rb)
    b.retire(rc)
    assert b.step(50) == []


def test_engine_coincident_particles_report_nonfinite():
    """The FP64 engine with two coincident LJ particles: the step reports MJG_NAN_DETECTED."""
    from mjolnir_amd import capi, workloads
    spec = workloads.lj_fluid(n_side=4, seed=3)
    spec.pos[1] = spec.pos[0]
    print("[gpu] engine with coincident particles", flush=True)
    e = capi.Engine(spec)
    with pytest.raises(capi.MjgError) as exc:
        e.step(10)
    assert exc.value.status == capi.NAN_DETECTED


def test_restored_replica_keeps_its_forced_rebuild_until_it_steps():
    """A replica restored at a step count (forced list rebuild and, for UnderdampedLangevin, the
    acceleration redrawn at that count) that takes part in a launch with no steps left keeps the
    forced restart for its next real step: load -> advance -> set_steps -> advance equals the
    uninterrupted run bitwise."""
    from mjolnir_amd import capi
    spec = make_golden.underdamped(_chains([240], seed0=820)[0], 0.5)
    other = _chains([100], seed0=821)[0]
    ref, (r0,) = _batch([spec], capi.FP32_FIXED)
    ref.step(300)
    x300, v300, f300 = ref.state(r0)
    ref.step(200)
    want = ref.state(r0)
    s2 = copy.deepcopy(spec)
    s2.pos, s2.vel = x300, v300
    b, (ro, rr) = _batch([other, s2], capi.FP32_FIXED)
    b.set_step_count(rr, 300)
    b.set_steps(ro, 100)
    b.set_steps(rr, 0)
    b.advance(100)               # the restored replica has no steps left in this launch
    b.set_steps(rr, 200)
    b.advance(100)
    b.advance(100)
    got = b.state(rr)
    assert b.info(rr)["step"] == 500
    for a, c in zip(want, got):
        assert np.array_equal(a, c)


DAMPED_BATCH = ["lj512_damped", "lj512_gjf_damped", "lj512_ul_damped", "chain200_damped", "chain200_gjf_damped",
                "dna12_damped"]


def test_fp64_batch_damped_trajectories_vs_reference():
    """Friction at zero temperature (deterministic damped trajectories, per-particle friction,
    non-unit masses): the reference's own BAOAB, G-JF and UnderdampedLangevin runs reproduced by FP64
    replicas, all in one batch, within 1e-8 A every 100 steps for 1000 steps."""
    from conftest import golden_case
    from mjolnir_amd import capi
    cases = [golden_case(n) for n in DAMPED_BATCH]
    b, ids = _batch([c[0] for c in cases], capi.FP64)
    steps = [int(s) for s in cases[0][1]["zt_frame_steps"]]
    done = 0
    for k, step in enumerate(steps):
        b.step(step - done)
        done = step
        for (spec, g), r in zip(cases, ids):
            assert [int(s) for s in g["zt_frame_steps"]] == steps
            x, _, _ = b.state(r)
            d = x - g["zt_frames"][k]
            if spec.boundary == "periodic":
                w = spec.upper - spec.lower
                d = d - w * np.round(d / w)
            assert np.abs(d).max() <= 1e-8, (spec.name, step, np.abs(d).max())
    for (spec, g), r in zip(cases, ids):
        _, v, _ = b.state(r)
        assert np.abs(v - g["zt_vel"]).max() <= 1e-8, spec.name


def test_fp64_batch_config3_duplex_vs_reference():
    """Config 3 as stated (200 bp, ideal B-form at rise 3.38 A, strained stacking at the start):
    the reference's own 1000-step zero-temperature G-JF run reproduced by an FP64 replica within
    1e-8 A (the reference's sensitivity here is about 1e-10 A)."""
    from conftest import golden_case
    from mjolnir_amd import capi
    spec, g = golden_case("dna200_b338")
    b, (r,) = _batch([spec], capi.FP64)
    b.step(int(g["zt_steps"]))
    x, v, _ = b.state(r)
    print(f"[gpu] dna200_b338 FP64 replica: max|dx| {np.abs(x - g['zt_pos']).max():.2e}", flush=True)
    assert np.abs(x - g["zt_pos"]).max() <= 1e-8
    assert np.abs(v - g["zt_vel"]).max() <= 1e-8


@pytest.mark.parametrize("kind", ["chain", "lj"])
def test_tuning_changes_no_result(kind):
    """The autotuner's knobs (list skins, cell-grid resolution, lanes per grid row) change no
    result, because forces are exact fixed-point sums over the pairs inside the cutoff: one replica
    under four fixed tunings and under the autotuner ends bitwise equal after 400 steps."""
    from mjolnir_amd import capi, workloads
    spec = _chains([1500], seed0=830)[0] if kind == "chain" else workloads.lj_fluid(n=1000, seed=31)
    settings = [(0.2, 0.05, 3, 1), (0.6, 0.3, 4, 4), (0.4, 0.1, 4, 2), (0.3, 0.2, 3, 4), None]
    states = []
    for st in settings:
        b = capi.Batch(capi.FP32_FIXED)
        r = b.add(spec)
        if st is not None:
            b.set_tuning(r, *st)
        b.initialize()
        b.step(400)
        states.append(b.state(r))
        print(f"[gpu] {kind} tuning {b.tuning(r)}: rebuilds {b.info(r)['rebuilds']}", flush=True)
        b.close()
    for s in states[1:]:
        for a, c in zip(states[0], s):
            assert np.array_equal(a, c)
