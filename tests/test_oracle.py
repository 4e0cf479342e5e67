"""Pins the CPU restatement (oracle/) before it is trusted as the checker.

* against the reference's own classes: golden vectors written by oracle/_ref/ref_driver
  (tests/golden/make_golden.py) must be reproduced bit for bit -- forces, energies, neighbour
  sets and 1000-step zero-temperature trajectories;
* against published known answers: Philox4x32-10 (Salmon et al., SC'11);
* against what the mathematics fixes: finite-difference forces (the reference's check_force
  method, test/util/check_force.hpp:52-166), Newton's third law, closed-form potential
  minima, brute-force neighbour sets, and the Debye-Hueckel constants quoted for 0.1 M/300 K.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_case
import make_golden
import oracle
from mjolnir_amd import workloads
from mjolnir_amd.spec import GlobalSpec, LocalSpec, SystemSpec, IntegratorSpec

CASES = list(make_golden.cases().keys())


# ----------------------------------------------------------------------------------------
# known answers
# ----------------------------------------------------------------------------------------
def test_philox_known_answers():
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, expect in kat:
        assert tuple(int(x) for x in oracle.philox4x32_10(ctr, key)) == expect


def test_noise_moments():
    z = np.array([oracle.noise3(12345, s, p) for s in range(40) for p in range(500)]).ravel()
    assert abs(z.mean()) < 4 / np.sqrt(len(z))
    assert abs(z.var() - 1.0) < 5 * np.sqrt(2.0 / len(z))
    # fourth moment of a normal is 3
    assert abs(np.mean(z ** 4) - 3.0) < 0.1
    # different steps / particles give different draws
    assert not np.array_equal(oracle.noise3(1, 0, 0), oracle.noise3(1, 1, 0))
    assert not np.array_equal(oracle.noise3(1, 0, 0), oracle.noise3(1, 0, 1))


def test_units_and_debye_constants():
    # kB in kcal/mol/K from the CODATA 2014 values the reference uses (Unit.hpp:23-24):
    # 1.38064852e-23 J/K * 6.022140857e23 /mol / 4184 J/kcal = 0.0019872036
    assert oracle.lib().orc_kB() == pytest.approx(0.0019872036, rel=1e-7)
    # 1/(4 pi eps0) in kcal A/(mol e^2): Coulomb's constant 332.0637...
    assert 1.0 / (4 * np.pi * oracle.lib().orc_eps0()) == pytest.approx(332.0637, rel=1e-5)


# ----------------------------------------------------------------------------------------
# against the reference's own classes (golden vectors)
# ----------------------------------------------------------------------------------------
@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_bitwise(name):
    spec, g = golden_case(name)
    o = oracle.OracleEngine(spec, with_integrator=False)
    names = [str(x) for x in g["names"]]
    assert names == spec.interaction_names()
    for k in range(len(names)):
        assert o.energy_of(k) == float(g[f"energy_{k}"]), names[k]
        assert np.array_equal(o.force_of(k), g[f"force_{k}"]), names[k]
        if f"npairs_{k}" in g:
            p = o.pairs(k - len(spec.local))
            assert len(p) == int(g[f"npairs_{k}"])
            assert make_golden.pair_key_hash(p, spec.n) == str(g[f"pairhash_{k}"])
            if f"pairs_{k}" in g:
                assert np.array_equal(p, g[f"pairs_{k}"])
    f, e = o.force_and_energy()
    assert np.array_equal(f, g["total_force"])


@pytest.mark.parametrize("name", [n for n in CASES if make_golden.cases()[n][2] > 0])
def test_oracle_zero_temperature_trajectory_bitwise(name):
    spec, g = golden_case(name)
    if spec.n > 2000:
        pytest.skip("large case: covered on the GPU against the same golden trajectory")
    o = oracle.OracleEngine(spec)
    o.step(int(g["zt_steps"]))
    x, v, _ = o.state()
    assert np.array_equal(x, g["zt_pos"])
    assert np.array_equal(v, g["zt_vel"])


# ----------------------------------------------------------------------------------------
# what the mathematics fixes
# ----------------------------------------------------------------------------------------
def _energy(spec, k):
    return oracle.OracleEngine(spec, with_integrator=False).energy_of(k)


@pytest.mark.parametrize("k", range(6))
def test_finite_difference_forces_chain(k):
    """check_force (test/util/check_force.hpp:52-166): central differences of calc_energy
    against calc_force, dr = 1e-5, on a perturbed chain; every interaction of the chain."""
    spec = make_golden.perturb(workloads.cg_chain(n=40, seed=21), 0.3, 22)
    base = oracle.OracleEngine(spec, with_integrator=False)
    f = base.force_of(k)
    rng = np.random.default_rng(k)
    dr = 1e-5
    for i in rng.choice(spec.n, size=8, replace=False):
        for d in range(3):
            sp, sm = make_golden.perturb(spec, 0, 0), make_golden.perturb(spec, 0, 0)
            sp.pos[i, d] += dr
            sm.pos[i, d] -= dr
            num = -(_energy(sp, k) - _energy(sm, k)) / (2 * dr)
            assert num == pytest.approx(f[i, d], abs=1e-4 * max(1.0, abs(f[i, d])))


@pytest.mark.parametrize("k", range(8))
def test_finite_difference_forces_dna(k):
    """check_force on a perturbed 3SPN2 duplex (8 bp): every interaction of the 3SPN2 model
    (3SPN2 bond, angle, periodic Gaussian dihedral, base stacking, base pairing, cross stacking,
    3SPN2 excluded volume, Debye-Hueckel).  The energy is the calc_energy path and the force
    the calc_force path, so the two restated paths check each other."""
    spec = make_golden.perturb(workloads.dna_duplex(8, seed=31, rise=3.6), 0.25, 32)
    base = oracle.OracleEngine(spec, with_integrator=False)
    f = base.force_of(k)
    assert np.abs(f).max() > 0 or k in (0, 1), "the term is inactive in this geometry"
    rng = np.random.default_rng(100 + k)
    dr = 1e-5
    for i in rng.choice(spec.n, size=10, replace=False):
        for d in range(3):
            sp, sm = make_golden.perturb(spec, 0, 0), make_golden.perturb(spec, 0, 0)
            sp.pos[i, d] += dr
            sm.pos[i, d] -= dr
            num = -(_energy(sp, k) - _energy(sm, k)) / (2 * dr)
            assert num == pytest.approx(f[i, d], abs=1e-4 * max(1.0, abs(f[i, d])))


def test_newton_third_law_dna():
    spec, _ = golden_case("dna12")
    o = oracle.OracleEngine(spec, with_integrator=False)
    for k in range(len(spec.interaction_names())):
        f = o.force_of(k)
        assert np.abs(f.sum(axis=0)).max() < 1e-9 * max(1.0, np.abs(f).sum())


def test_newton_third_law():
    spec, _ = golden_case("chain200")
    o = oracle.OracleEngine(spec, with_integrator=False)
    for k in range(len(spec.interaction_names())):
        f = o.force_of(k)
        assert np.abs(f.sum(axis=0)).max() < 1e-9 * max(1.0, np.abs(f).sum())


def _two_lj(r, sigma=4.0, eps=0.2):
    pos = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    lj = GlobalSpec("LennardJones", np.arange(2), np.array([[sigma, eps], [sigma, eps]]), cutoff=2.5)
    return SystemSpec("unlimited", np.ones(2), pos, np.zeros((2, 3)), attributes={"temperature": 300.0}, global_=[lj])


def test_lj_minimum_closed_form():
    """LJ with Lorentz-Berthelot: force vanishes at r = 2^(1/6) sigma where E = -eps - 4 eps c,
    c = 2.5^-12 - 2.5^-6 (the shift, LennardJonesPotential.hpp:34-38); zero beyond 2.5 sigma."""
    sigma, eps = 4.0, 0.2
    rmin = 2 ** (1 / 6) * sigma
    o = oracle.OracleEngine(_two_lj(rmin), with_integrator=False)
    c = 2.5 ** -12 - 2.5 ** -6
    assert o.energy_of(0) == pytest.approx(-eps - 4 * eps * c, rel=1e-12)
    assert abs(o.force_of(0)[0, 0]) < 1e-12
    o2 = oracle.OracleEngine(_two_lj(2.5 * sigma * (1 + 1e-9)), with_integrator=False)
    assert o2.energy_of(0) == 0.0 and np.all(o2.force_of(0) == 0.0)


def test_harmonic_bond_closed_form():
    """E = k (r - r0)^2 without the 1/2 (HarmonicPotential.hpp:23-32); |F| = 2 k (r - r0)."""
    k, r0, r = 110.0, 3.8, 4.1
    pos = np.array([[0.0, 0.0, 0.0], [0.0, r, 0.0]])
    spec = SystemSpec("unlimited", np.ones(2), pos, np.zeros((2, 3)), attributes={"temperature": 300.0},
                      local=[LocalSpec("BondLength", "Harmonic", "bond", [[0, 1]], [[k, r0]])])
    o = oracle.OracleEngine(spec, with_integrator=False)
    assert o.energy_of(0) == pytest.approx(k * (r - r0) ** 2, rel=1e-12)
    assert o.force_of(0)[1, 1] == pytest.approx(-2 * k * (r - r0), rel=1e-12)


def _brute_pairs(spec, g):
    """All pairs i<j of participants with |r_ij|^2 < (r_c (1+m))^2 (minimum image), computed
    with numpy in the reference's operation order, with no exclusions."""
    part = g.participants.astype(np.int64)
    if g.potential == "LennardJones":
        rc = g.params[:, 0].max() * g.cutoff
    else:
        raise NotImplementedError
    rl2 = (rc * (1 + g.margin)) ** 2
    x = spec.pos[part]
    i, j = np.triu_indices(len(part), k=1)
    d = x[j] - x[i]
    if spec.boundary == "periodic":
        w = spec.upper - spec.lower
        hw = w * 0.5
        d = np.where(d < -hw, d + w, np.where(d >= hw, d - w, d))
    r2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
    keep = r2 < rl2
    return np.stack([part[i[keep]], part[j[keep]]], axis=1)


@pytest.mark.parametrize("name", ["lj512", "lj_boundary"])
def test_neighbor_set_brute_force(name):
    spec, _ = golden_case(name)
    o = oracle.OracleEngine(spec, with_integrator=False)
    got = o.pairs(0).astype(np.int64)
    want = _brute_pairs(spec, spec.global_[0])
    k = lambda p: set(map(tuple, p.tolist()))
    assert k(got) == k(want)


def test_velocity_verlet_harmonic_energy_drift():
    """gamma = 0: BAOAB is velocity Verlet (BAOABLangevinIntegrator.hpp:79-82); a harmonic
    dumbbell keeps its energy to O(dt^2) and oscillates with the analytic period."""
    k, r0 = 110.0, 3.8
    pos = np.array([[0.0, 0.0, 0.0], [4.0, 0.0, 0.0]])
    spec = SystemSpec("unlimited", np.ones(2), pos, np.zeros((2, 3)), attributes={"temperature": 300.0},
                      local=[LocalSpec("BondLength", "Harmonic", "bond", [[0, 1]], [[k, r0]])],
                      integrator=IntegratorSpec("BAOABLangevin", 0.001, np.zeros(2)))
    o = oracle.OracleEngine(spec)
    omega = np.sqrt(2 * k / 0.5)   # reduced mass 1/2, spring constant 2k
    xs = []
    for _ in range(2000):
        o.step(1)
        x, v, _ = o.state()
        xs.append(x[1, 0] - x[0, 0])
    xs = np.array(xs)
    t = 0.001 * np.arange(1, 2001)
    analytic = r0 + 0.2 * np.cos(omega * t)
    assert np.abs(xs - analytic).max() < 2e-3


def _parse_ene(text):
    rows = [l.split() for l in text.splitlines() if l and not l.startswith("#")]
    return np.array([[float(x) for x in r] for r in rows])


def test_energy_observer_golden_matches_oracle():
    """The reference's .ene lines (EnergyObserver.hpp:87-150) hold the oracle's per-interaction
    energies (calc_energy path), kinetic energy and attributes at steps 0, 100, 200, 300."""
    import make_golden
    import oracle
    from mjolnir_amd import workloads
    g = np.load(os.path.join(GOLDEN, "sim_chain200.npz"))
    spec = workloads.zero_temperature(make_golden.cases()["chain200"][0]())
    assert str(g["spec_hash"]) == make_golden.spec_hash(spec)
    rows = _parse_ene(str(g["ene"]))
    assert rows.shape[0] == 4
    o = oracle.OracleEngine(spec)
    ni = len(spec.interaction_names())
    for r, step in enumerate((0, 100, 200, 300)):
        if step:
            o.step(100)
        _, v, _ = o.state()
        e = [o.energy_of(k) for k in range(ni)]
        ek = 0.5 * np.sum(spec.mass * np.sum(v * v, axis=1))
        want = np.array([step] + e + [ek, spec.attributes["ionic_strength"], spec.attributes["temperature"]])
        assert np.allclose(rows[r], want, rtol=0, atol=1.5e-6), (step, rows[r] - want)


# ----------------------------------------------------------------------------------------
# g-BAOAB with bond constraints (gBAOABLangevinIntegrator.hpp)
# ----------------------------------------------------------------------------------------
def test_gbaoab_constraints_hold_at_finite_temperature():
    """After every step the constrained bonds sit at their lengths to the position tolerance
    (|r^2 - v0^2| <= tol, what correct_coordinate enforces) and the relative velocity along each
    bond is below the velocity tolerance (|lambda| <= tol, correct_velocity)."""
    spec = workloads.constrained_chain(80, seed=7, gamma=0.5)
    c = spec.constraint
    o = oracle.OracleEngine(spec)
    redm = 1.0 / (1.0 / spec.mass[c.indices[:, 0]] + 1.0 / spec.mass[c.indices[:, 1]])
    for _ in range(5):
        o.step(40)
        x, v, _ = o.state()
        d = x[c.indices[:, 1]] - x[c.indices[:, 0]]
        assert np.abs(np.sum(d * d, axis=1) - c.v0 ** 2).max() <= c.tolerance
        dv = v[c.indices[:, 1]] - v[c.indices[:, 0]]
        lam = np.sum(d * dv, axis=1) * redm / c.v0 ** 2
        assert np.abs(lam).max() <= c.tolerance * 1.0001


def test_gbaoab_without_constraints_is_baoab():
    """With no constraint every correction is empty and g-BAOAB is BAOAB: the two split the
    A step as x + h v + h v' against x + (h v + h v'), so they agree to rounding."""
    import copy
    spec = workloads.zero_temperature(make_golden.perturb(workloads.cg_chain(n=120, seed=4), 0.1, 3))
    a = oracle.OracleEngine(spec)
    s2 = copy.deepcopy(spec)
    s2.integrator.type = "gBAOABLangevin"
    b = oracle.OracleEngine(s2)
    a.step(200)
    b.step(200)
    xa, va, _ = a.state()
    xb, vb, _ = b.state()
    assert np.abs(xa - xb).max() < 1e-9
    assert np.abs(va - vb).max() < 1e-9
    assert not np.array_equal(xa, xb)     # the step really is the g-BAOAB one


# ----------------------------------------------------------------------------------------
# UnderdampedLangevin (UnderdampedLangevinIntegrator.hpp)
# ----------------------------------------------------------------------------------------
def test_underdamped_free_particles_velocity_variance():
    """Free particles: a = noise with variance s2 = 2 kB T gamma / (m dt) per component, and one
    step is v' = q v + c a + c a' (q = o (o^2 + g), o = 1 - g, g = gamma dt/2, c = (dt/2) o),
    an ARMA(1,1) process whose stationary variance is 2 c^2 s2 / (1 - q).  A noise entering once,
    or with a wrong amplitude, misses it by far more than the 1.5 % sampling bar."""
    n, m, gamma, dt, T = 20000, 2.0, 0.3, 0.05, 300.0
    spec = SystemSpec("unlimited", np.full(n, m), np.random.default_rng(1).normal(size=(n, 3)) * 100,
                      np.zeros((n, 3)), attributes={"temperature": T},
                      integrator=IntegratorSpec("UnderdampedLangevin", dt, np.full(n, gamma), seed=9))
    o = oracle.OracleEngine(spec)
    g = gamma * dt / 2
    om = 1 - g
    q = om * (om * om + g)
    c = dt / 2 * om
    s2 = 2 * oracle.kB() * T * gamma / (m * dt)
    want = 2 * c * c * s2 / (1 - q)
    o.step(400)
    got = []
    for _ in range(5):
        o.step(60)
        got.append(np.mean(o.state()[1] ** 2))
    assert np.mean(got) == pytest.approx(want, rel=0.015)
    assert want == pytest.approx(oracle.kB() * T / m, rel=0.02)   # equipartition to O(gamma dt)


def test_underdamped_without_friction_is_velocity_verlet():
    """gamma = 0, T = 0: UnderdampedLangevin moves x by dt v + dt^2/2 a and kicks v by dt/2 a
    twice -- velocity Verlet, as BAOAB with gamma = 0; the two agree to rounding."""
    import copy
    spec = workloads.zero_temperature(make_golden.perturb(workloads.cg_chain(n=120, seed=4), 0.1, 3))
    spec.integrator.coefs[:] = 0.0
    a = oracle.OracleEngine(spec)
    s2 = copy.deepcopy(spec)
    s2.integrator.type = "UnderdampedLangevin"
    b = oracle.OracleEngine(s2)
    a.step(200)
    b.step(200)
    xa, va, _ = a.state()
    xb, vb, _ = b.state()
    assert np.abs(xa - xb).max() < 1e-9
    assert np.abs(va - vb).max() < 1e-9
    assert not np.array_equal(xa, xb)
