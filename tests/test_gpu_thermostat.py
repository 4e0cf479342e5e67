"""Thermostat statistics of the FP64 engine over 10^6 steps (BAOAB and G-JF).

* trimer gas (closed-form Boltzmann densities, tests/thermo_stats.py): bond-length and angle
  histograms against r^2 exp(-k (r - r0)^2 / kB T) and sin(theta) exp(-U(theta) / kB T) by a
  chi-square at two standard deviations of its distribution; <KE> = 3/2 N kB T within two
  standard errors of the block average;
* LJ fluid, 512 particles: <KE> = 3/2 N kB T within two standard errors (block averages).

Frames are taken every 1000 steps, past the correlation times of these systems (friction 1).
"""
import numpy as np
import pytest

import thermo_stats as ts

pytestmark = pytest.mark.gpu

STEPS = 1_000_000
EVERY = 1000


def _run(spec, equilibrate=20000):
    from mjolnir_amd import capi
    print(f"[gpu] thermostat {spec.name} {spec.integrator.type} N={spec.n}, {STEPS} steps", flush=True)
    e = capi.Engine(spec)
    e.step(equilibrate)
    for k in range(STEPS // EVERY):
        e.step(EVERY)
        x, v, _ = e.state()
        if (k + 1) % 100 == 0:
            print(f"[gpu]   {(k + 1) * EVERY} steps", flush=True)
        yield x, v


@pytest.mark.parametrize("integ", ["BAOABLangevin", "GJFNVTLangevin"])
def test_trimer_gas_boltzmann_1e6_steps(integ):
    from mjolnir_amd import workloads
    spec = workloads.trimer_gas(300, integrator=integ, seed=9)
    bonds, angles, ke = [], [], []
    for x, v in _run(spec):
        r, a = ts.trimer_observables(spec, x)
        bonds.append(r)
        angles.append(a)
        ke.append(0.5 * np.sum(spec.mass[:, None] * v * v))
    kT = workloads.KB_KCAL * 300.0
    (cb, db, bb), (ca, da, ba) = ts.trimer_checks(spec, np.concatenate(bonds), np.concatenate(angles), kT)
    m, se = ts.block_mean(np.array(ke))
    target = 1.5 * spec.n * kT
    print(f"[gpu] {integ}: bond chi2 {cb:.1f}/{db} (bar {bb:.1f}), angle chi2 {ca:.1f}/{da} (bar {ba:.1f}), "
          f"<KE>/(3/2 N kT) = {m / target:.5f} +- {se / target:.5f}", flush=True)
    assert cb <= bb
    assert ca <= ba
    assert abs(m - target) <= 2 * se


@pytest.mark.parametrize("integ", ["BAOABLangevin", "GJFNVTLangevin"])
def test_lj512_kinetic_temperature_1e6_steps(integ):
    from mjolnir_amd import workloads
    spec = workloads.lj_fluid(n_side=8, seed=12, gamma=1.0)
    if integ == "GJFNVTLangevin":
        spec.integrator.type = integ
        spec.integrator.coefs = spec.integrator.coefs * spec.mass
    ke = [0.5 * np.sum(spec.mass[:, None] * v * v) for _, v in _run(spec)]
    m, se = ts.block_mean(np.array(ke))
    target = 1.5 * spec.n * workloads.KB_KCAL * 300.0
    print(f"[gpu] LJ512 {integ}: <KE>/(3/2 N kT) = {m / target:.5f} +- {se / target:.5f}", flush=True)
    assert abs(m - target) <= 2 * se


def test_config4_chains_mixed_kinetic_temperature_equals_fp64():
    """Kinetic temperature of config-4 chains (24 chains of 50-500 beads at 300 K, the ensemble's
    model: bonds, Gaussian angles, Clementi dihedrals, Go contacts, excluded volume and Debye-Hueckel,
    BAOAB gamma = 0.1, dt = 0.01) in the replica kernel's mixed mode against its FP64 mode.  Each
    mode runs 10^4 steps to equilibrate, then 6 x 10^4 steps sampled every 100 steps; the pooled
    <KE> / (3/2 N kB T) of each 5000-step block is taken and the blocks' median compared.  Both
    modes sit about 3 % above 1 with heating spikes of 10-20 % in single blocks (a bond angle of
    the shallow Gaussian well reaching 180 degrees releases dihedral energy), so the median is the
    statistic; the mixed mode's must equal the FP64 mode's within 2 %."""
    from mjolnir_amd import capi, workloads
    lens = np.geomspace(50, 500, 24).astype(int)
    specs = [workloads.cg_chain(n=int(n), seed=700 + k, T=300.0, max_bend_deg=150.0) for k, n in enumerate(lens)]
    target = workloads.KB_KCAL * 300.0
    med = {}
    for prec in (capi.FP32_FIXED, capi.FP64):
        print(f"[gpu] config-4 chains, precision {prec}", flush=True)
        b = capi.Batch(prec)
        ids = [b.add(s) for s in specs]
        assert b.initialize() == []
        failed = set(b.step(10000))
        blocks = []
        for _ in range(12):
            ratios = []
            for _ in range(50):
                failed.update(b.step(100))
                ke = n = 0.0
                for r, s in zip(ids, specs):
                    if r in failed:
                        continue
                    _, v, _ = b.state(r)
                    ke += 0.5 * float(np.sum(s.mass[:, None] * v * v))
                    n += s.n
                ratios.append(ke / (1.5 * n * target))
            blocks.append(float(np.mean(ratios)))
        med[prec] = float(np.median(blocks))
        print(f"[gpu] precision {prec}: block means {np.round(blocks, 4)}, median {med[prec]:.4f}, "
              f"stopped replicas {sorted(failed)}", flush=True)
        assert len(failed) <= 1
        b.close()
    assert abs(med[capi.FP32_FIXED] - med[capi.FP64]) <= 0.02
    assert 0.97 <= med[capi.FP64] <= 1.08
