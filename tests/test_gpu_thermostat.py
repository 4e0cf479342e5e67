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
