"""Seeded synthetic workloads: the LJ fluid, the CG protein-like chain and replica ensembles.

Units are the reference's ``[units] length="angstrom", energy="kcal/mol"`` with masses
in amu.  The Boltzmann constant is formed exactly as ``read_units`` forms it
(``mjolnir/input/read_units.hpp:32-36`` from ``mjolnir/core/Unit.hpp:23-44``).

Builder choices that the workload definitions leave open (all applied identically in
every backend, because every backend receives the same ``SystemSpec``):

* chain masses 1.0 amu and ``dt = 0.01``;
* Gaussian-angle width 0.15 rad (the AICG2+ example uses 0.15);
* Clementi dihedral ``k1 = 1.0``, ``k3 = 0.5`` kcal/mol;
* neighbour-list margin 0.5 for ExcludedVolume and Debye-Hueckel;
* excluded-volume and Debye-Hueckel pairs ignore ``bond`` within 3 and ``contact``
  within 1 (the AICG2+ example's ``ignore.particles_within``);
* initial velocities are Maxwell-Boltzmann draws from numpy (not the reference's
  mt19937 stream), passed in as given velocities.
"""
from __future__ import annotations

import math
import os
from typing import List, Optional, Tuple

import numpy as np

from .spec import NIL, ConstraintSpec, GlobalSpec, IntegratorSpec, LocalSpec, SystemSpec

# kB [J/K] -> [kcal/mol/K]: kB * (J_to_cal / 1000) * N_A  (read_units.hpp:35-36)
_KB_SI = 1.38064852e-23
_NA = 6.022140857e23
KB_KCAL = _KB_SI * ((1.0 / 4.184) / 1000.0) * _NA


def maxwell_boltzmann(rng: np.random.Generator, mass: np.ndarray, T: float) -> np.ndarray:
    sd = np.sqrt(KB_KCAL * T / mass)
    return rng.standard_normal((len(mass), 3)) * sd[:, None]


# ----------------------------------------------------------------------------------------
# LJ fluid
# ----------------------------------------------------------------------------------------

def lj_fluid(n_side: int = 16, rho_reduced: float = 0.6, sigma: float = 4.0, epsilon: float = 0.2,
             cutoff: float = 2.5, margin: float = 0.5, T: float = 300.0, gamma: float = 0.1,
             dt: float = 0.01, seed: int = 1, jitter: float = 0.1, n: Optional[int] = None) -> SystemSpec:
    """N = n_side**3 identical LJ particles on a jittered simple-cubic lattice in a periodic cube.

    ``n_side=16`` is the 4096-particle LJ fluid (rho sigma^3 = 0.6, L = 75.880 A).  With ``n``
    the fluid has exactly n particles: a seeded random subset of the sites of the smallest
    lattice with at least n sites, in the cube that holds n particles at the same density.
    """
    rng = np.random.default_rng(seed)
    if n is None:
        n = n_side ** 3
    else:
        n_side = int(np.ceil(round(n ** (1.0 / 3.0), 9)))
        while n_side ** 3 < n:
            n_side += 1
    rho = rho_reduced / sigma ** 3
    L = (n / rho) ** (1.0 / 3.0)
    a = L / n_side
    g = np.arange(n_side, dtype=np.float64)
    gx, gy, gz = np.meshgrid(g, g, g, indexing="ij")
    lattice = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1) * a + 0.5 * a
    if len(lattice) > n:
        lattice = lattice[np.sort(rng.choice(len(lattice), size=n, replace=False))]
    pos = lattice + rng.uniform(-jitter * sigma, jitter * sigma, size=(n, 3))
    pos = np.mod(pos, L)
    mass = np.ones(n)
    vel = maxwell_boltzmann(rng, mass, T)
    lj = GlobalSpec(potential="LennardJones", participants=np.arange(n),
                    params=np.tile([sigma, epsilon], (n, 1)), cutoff=cutoff,
                    partition="CellList", margin=margin, ignore_molecule="Nothing",
                    ignore_within={})
    return SystemSpec(boundary="periodic", mass=mass, pos=pos, vel=vel,
                      lower=np.zeros(3), upper=np.full(3, L),
                      attributes={"temperature": T}, global_=[lj],
                      integrator=IntegratorSpec("BAOABLangevin", dt, np.full(n, gamma), seed=seed),
                      name=f"lj_fluid_N{n}")


# ----------------------------------------------------------------------------------------
# CG chain (AICG2+-style functional forms)
# ----------------------------------------------------------------------------------------

def self_avoiding_walk(n: int, rng: np.random.Generator, bond: float = 3.8,
                       min_sep: float = 4.0, max_tries: int = 200,
                       max_bend_deg: Optional[float] = None) -> np.ndarray:
    """Seeded self-avoiding random walk: consecutive beads ``bond`` apart, every other
    pair (|i-j| >= 2) at least ``min_sep`` apart.  Backtracks when a bead cannot be placed.

    ``max_bend_deg`` bounds the bond angle at each bead: angles near 180 degrees make the
    angle force (1/sin theta) and the dihedral of the next quadruplet singular, and a chain
    built with them can blow up within a few hundred steps, on the reference as here."""
    pos = np.zeros((n, 3))
    k = 1
    fails = np.zeros(n, dtype=np.int64)
    min_sep2 = min_sep * min_sep
    cos_max = None if max_bend_deg is None else math.cos(math.radians(max_bend_deg))
    while k < n:
        placed = False
        for _ in range(max_tries):
            d = rng.standard_normal(3)
            d *= bond / np.linalg.norm(d)
            cand = pos[k - 1] + d
            if k >= 2 and cos_max is not None:
                a = pos[k - 2] - pos[k - 1]
                if np.dot(a, d) / (np.linalg.norm(a) * bond) < cos_max:
                    continue
            if k >= 2:
                dd = pos[:k - 1] - cand
                if np.min(np.einsum("ij,ij->i", dd, dd)) < min_sep2:
                    continue
            pos[k] = cand
            placed = True
            break
        if placed:
            k += 1
        else:
            fails[k] += 1
            back = min(k - 1, 2 + int(fails[k]))
            k = max(1, k - back)
    return pos


def _angles(pos: np.ndarray, idx: np.ndarray) -> np.ndarray:
    a = pos[idx[:, 0]] - pos[idx[:, 1]]
    b = pos[idx[:, 2]] - pos[idx[:, 1]]
    c = np.einsum("ij,ij->i", a, b) / (np.linalg.norm(a, axis=1) * np.linalg.norm(b, axis=1))
    return np.arccos(np.clip(c, -1.0, 1.0))


def _dihedrals(pos: np.ndarray, idx: np.ndarray) -> np.ndarray:
    # same convention as DihedralAngleInteraction.hpp:176-195
    r_ij = pos[idx[:, 0]] - pos[idx[:, 1]]
    r_kj = pos[idx[:, 2]] - pos[idx[:, 1]]
    r_lk = pos[idx[:, 3]] - pos[idx[:, 2]]
    m = np.cross(r_ij, r_kj)
    nn = np.cross(r_kj, -r_lk)
    c = np.einsum("ij,ij->i", m, nn) / np.sqrt(np.einsum("ij,ij->i", m, m) * np.einsum("ij,ij->i", nn, nn))
    phi = np.arccos(np.clip(c, -1.0, 1.0))
    return np.copysign(phi, np.einsum("ij,ij->i", r_ij, nn))


def cg_chain(n: int = 1000, seed: int = 2, T: float = 300.0, ionic_strength: float = 0.1,
             gamma: float = 0.1, dt: float = 0.01, charged_fraction: float = 0.2,
             bond_k: float = 110.0, angle_k: float = -0.5, angle_sigma: float = 0.15,
             dihedral_k1: float = 1.0, dihedral_k3: float = 0.5, go_eps: float = 0.3,
             go_cutoff: float = 6.5, go_min_seq: int = 4, ev_radius: float = 2.0,
             ev_eps: float = 0.2, margin: float = 0.5, max_bend_deg: Optional[float] = None) -> SystemSpec:
    """Protein-like CG chain in an unbounded box (UnlimitedGridCellList)."""
    rng = np.random.default_rng(seed)
    pos = self_avoiding_walk(n, rng, max_bend_deg=max_bend_deg)
    mass = np.ones(n)
    vel = maxwell_boltzmann(rng, mass, T)

    bonds = np.stack([np.arange(n - 1), np.arange(1, n)], axis=1)
    r0 = np.linalg.norm(pos[bonds[:, 1]] - pos[bonds[:, 0]], axis=1)
    local = [LocalSpec("BondLength", "Harmonic", "bond", bonds,
                       np.stack([np.full(n - 1, bond_k), r0], axis=1))]
    if n >= 3:
        ang = np.stack([np.arange(n - 2), np.arange(1, n - 1), np.arange(2, n)], axis=1)
        th0 = _angles(pos, ang)
        local.append(LocalSpec("BondAngle", "Gaussian", "none", ang,
                               np.stack([np.full(len(ang), angle_k), np.full(len(ang), angle_sigma), th0], axis=1)))
    if n >= 4:
        dih = np.stack([np.arange(n - 3), np.arange(1, n - 2), np.arange(2, n - 1), np.arange(3, n)], axis=1)
        ph0 = _dihedrals(pos, dih)
        local.append(LocalSpec("DihedralAngle", "ClementiDihedral", "none", dih,
                               np.stack([np.full(len(dih), dihedral_k1), np.full(len(dih), dihedral_k3), ph0], axis=1)))
    # native contacts: |i-j| >= go_min_seq and within go_cutoff in the initial structure
    d2 = np.sum((pos[:, None, :] - pos[None, :, :]) ** 2, axis=-1) if n <= 6000 else None
    ii, jj = np.nonzero(np.triu(d2 < go_cutoff ** 2, k=go_min_seq))
    if len(ii):
        cidx = np.stack([ii, jj], axis=1)
        cr0 = np.sqrt(d2[ii, jj])
        local.append(LocalSpec("BondLength", "GoContact", "contact", cidx,
                               np.stack([np.full(len(ii), go_eps), cr0], axis=1)))

    within = {"bond": 3, "contact": 1}
    ev = GlobalSpec(potential="ExcludedVolume", participants=np.arange(n),
                    params=np.full((n, 1), ev_radius), cutoff=2.0, epsilon=ev_eps,
                    partition="CellList", margin=margin, ignore_molecule="Nothing", ignore_within=within)
    n_q = max(2, int(round(charged_fraction * n))) if n >= 2 else 0
    charged = np.sort(rng.choice(n, size=n_q, replace=False))
    q = rng.choice([-1.0, 1.0], size=n_q)
    dh = GlobalSpec(potential="DebyeHuckel", participants=charged, params=q[:, None], cutoff=5.5,
                    partition="CellList", margin=margin, ignore_molecule="Nothing", ignore_within=within)
    return SystemSpec(boundary="unlimited", mass=mass, pos=pos, vel=vel,
                      attributes={"temperature": T, "ionic_strength": ionic_strength},
                      local=local, global_=[ev, dh],
                      integrator=IntegratorSpec("BAOABLangevin", dt, np.full(n, gamma), seed=seed),
                      name=f"cg_chain_N{n}")


def constrained_chain(n: int = 100, seed: int = 5, gamma: float = 0.1, dt: float = 0.01,
                      max_iteration: int = 500, tolerance: float = 1e-6) -> SystemSpec:
    """The CG chain with its bonds held rigid instead of harmonic: bond constraints at the
    initial lengths (connection kind "bond", so exclusions and molecules are unchanged) and the
    g-BAOAB integrator (gBAOABLangevinIntegrator.hpp, read_constraint.hpp)."""
    s = cg_chain(n=n, seed=seed, gamma=gamma, dt=dt, max_bend_deg=150.0)
    bonds = s.local[0]
    assert (bonds.interaction, bonds.potential) == ("BondLength", "Harmonic")
    s.local = s.local[1:]
    s.constraint = ConstraintSpec(indices=bonds.indices, v0=bonds.params[:, 1], topology="bond",
                                  max_iteration=max_iteration, tolerance=tolerance)
    s.integrator = IntegratorSpec("gBAOABLangevin", dt, np.full(n, gamma), seed=seed)
    s.name = f"constrained_chain_N{n}"
    return s


def replica_ensemble(n_replicas: int = 512, n_min: int = 50, n_max: int = 2000,
                     T_min: float = 280.0, T_max: float = 400.0, seed: int = 4,
                     max_bend_deg: float = 150.0) -> List[SystemSpec]:
    """Independent CG chains with log-uniform lengths and uniform temperatures (bond angles
    of the initial walks at most ``max_bend_deg``)."""
    rng = np.random.default_rng(seed)
    lens = np.exp(rng.uniform(math.log(n_min), math.log(n_max), size=n_replicas)).astype(np.int64)
    temps = rng.uniform(T_min, T_max, size=n_replicas)
    seeds = rng.integers(1, 2 ** 31 - 1, size=n_replicas)
    jobs = [(int(n), int(s), float(t), max_bend_deg) for n, t, s in zip(lens, temps, seeds)]
    # MJG_WORKLOAD_PROCS=1 builds them in this process (e.g. under a profiler that preloads itself)
    workers = int(os.environ.get("MJG_WORKLOAD_PROCS", min(16, os.cpu_count() or 1)))
    if workers > 1 and n_replicas >= 32:
        # every replica has its own seed, so the result does not depend on the pool.  The children
        # are forked; a forked child of a process that holds engines or batches must never call
        # into the HIP runtime: the handles skip their destroy call in any process but their
        # creator's, and the parent's pending garbage is collected here first
        import gc
        import multiprocessing as mp
        gc.collect()
        with mp.get_context("fork").Pool(workers) as pool:
            return pool.map(_ensemble_member, jobs, chunksize=4)
    return [_ensemble_member(j) for j in jobs]


def _ensemble_member(job) -> SystemSpec:
    n, seed, T, max_bend_deg = job
    return cg_chain(n, seed=seed, T=T, max_bend_deg=max_bend_deg)


def trimer_gas(n_trimers: int = 300, integrator: str = "BAOABLangevin", T: float = 300.0, friction: float = 1.0,
               dt: float = 0.01, bond_k: float = 10.0, bond_r0: float = 3.8, angle_k: float = -2.0,
               angle_sigma: float = 0.3, angle_theta0: float = 1.9, spacing: float = 40.0,
               seed: int = 6) -> SystemSpec:
    """Independent trimers i-j-k in an unbounded box: harmonic bonds i-j and j-k
    (E = k (r - r0)^2, HarmonicPotential.hpp:23-32) and a Gaussian well on the angle i-j-k
    (GaussianPotential.hpp:20-38), no pair interaction.  The Boltzmann density factorizes, so each
    bond length has the density r^2 exp(-k (r - r0)^2 / kB T) and each angle
    sin(theta) exp(-U(theta) / kB T): closed forms a thermostat must sample.  Friction `friction`
    per particle (alpha = friction * mass for G-JF); unit masses."""
    rng = np.random.default_rng(seed)
    n = 3 * n_trimers
    side = int(np.ceil(n_trimers ** (1.0 / 3.0)))
    pos = np.zeros((n, 3))
    for t in range(n_trimers):
        c = np.array([t % side, (t // side) % side, t // (side * side)], dtype=np.float64) * spacing
        a = rng.uniform(1.6, 2.2)
        pos[3 * t] = c + [bond_r0, 0.0, 0.0]
        pos[3 * t + 1] = c
        pos[3 * t + 2] = c + bond_r0 * np.array([math.cos(a), math.sin(a), 0.0])
    mass = np.ones(n)
    vel = maxwell_boltzmann(rng, mass, T)
    bonds = np.concatenate([np.stack([3 * np.arange(n_trimers), 3 * np.arange(n_trimers) + 1], axis=1),
                            np.stack([3 * np.arange(n_trimers) + 1, 3 * np.arange(n_trimers) + 2], axis=1)])
    ang = np.stack([3 * np.arange(n_trimers), 3 * np.arange(n_trimers) + 1, 3 * np.arange(n_trimers) + 2], axis=1)
    local = [LocalSpec("BondLength", "Harmonic", "bond", bonds, np.tile([bond_k, bond_r0], (len(bonds), 1))),
             LocalSpec("BondAngle", "Gaussian", "none", ang,
                       np.tile([angle_k, angle_sigma, angle_theta0], (len(ang), 1)))]
    coefs = np.full(n, friction) * (mass if integrator == "GJFNVTLangevin" else 1.0)
    return SystemSpec(boundary="unlimited", mass=mass, pos=pos, vel=vel, attributes={"temperature": T},
                      local=local, integrator=IntegratorSpec(integrator, dt, coefs, seed=seed),
                      name=f"trimer_gas_{n_trimers}")


def zero_temperature(spec: SystemSpec) -> SystemSpec:
    """Copy of ``spec`` with gamma (or alpha) = 0, which zeroes the friction and the noise
    amplitude: BAOAB and G-JF then reduce to velocity Verlet
    (BAOABLangevinIntegrator.hpp:79-82, GJFNVTLangevinIntegrator.hpp:116-118).  The
    temperature attribute is kept because Debye-Hueckel parameters depend on it.

    A spec whose temperature attribute is already 0 is returned unchanged: its noise amplitudes
    are zero at any friction (sqrt(kB T ...) in every integrator's reset_parameters), so its
    damped trajectory is deterministic and the friction coefficients stay in it."""
    import copy
    s = copy.deepcopy(spec)
    if s.attributes.get("temperature") == 0.0:
        return s
    if s.integrator is not None:
        s.integrator.coefs = np.zeros_like(s.integrator.coefs)
    return s


# ------------------------------------------------------------------------------------------
# 3SPN2 double-stranded DNA
# ------------------------------------------------------------------------------------------
DNA_MASS = {"P": 94.97, "S": 83.11, "A": 134.1, "T": 125.1, "G": 150.1, "C": 110.1}
_COMP = {"A": "T", "T": "A", "G": "C", "C": "G"}
_BASE_CODE = {"A": 0, "T": 1, "G": 2, "C": 3}
_BEAD_CODE = {"P": 0, "S": 1, "A": 2, "T": 3, "G": 4, "C": 5}
# the base-pairing angle S-B-B' of each base in the reference's parameter set
# (ThreeSPN2BasePairPotential.hpp:530-560: theta1 of AT/TA/GC/CG is the angle at the first base)
_SBB_ANGLE = {"A": 156.54, "T": 135.78, "G": 159.81, "C": 141.16}


def dna_duplex(n_bp: int = 200, seed: int = 3, T: float = 300.0, ionic_strength: float = 0.15,
               dt: float = 0.02, gamma: float = 0.1, rise: float = 3.38, twist_deg: float = 36.0,
               r_base: float = 3.0, bp_dist: float = 6.0, r_phos: float = 8.9, sb_bond: float = 4.4,
               margin: float = 0.4) -> SystemSpec:
    """Ideal right-handed B-form duplex of the 3SPN2 model (3 beads per nucleotide, no 5' terminal
    phosphate), random sequence, with the interactions of the reference's dna_3SPN2 example:
    3SPN2 bonds, harmonic angles, Gaussian backbone dihedrals, base stacking, base pairing,
    cross stacking, 3SPN2 excluded volume and Debye-Hueckel on the phosphates (charge -0.6).

    Bases sit at radius r_base; the two bases of a pair lie in one plane, bp_dist apart (the
    partner turned by 2 asin(bp_dist / 2 r_base) about the axis; 180 degrees by default).  The
    rise is B-DNA's 3.38 A; at that rise consecutive bases of a strand sit inside the 3SPN2
    base-stacking distances, so the duplex starts with strained stacking (about +4 kcal/mol per
    step) that relaxes within a few hundred steps.  A rise of 3.6 A starts it near those distances
    instead, with pairing, stacking and cross stacking formed and no strained contacts.
    Each sugar lies in its base's plane, sb_bond from the base, at the angle S-B-B' of its base in
    the reference's pairing parameters, on the side away from the axis; phosphates sit at radius
    r_phos midway (angle and height) between consecutive sugars of their strand.  Bond, angle
    and dihedral minima are the initial values.  G-JF with alpha = gamma * mass."""
    rng = np.random.default_rng(seed)
    seq = [("A", "T", "G", "C")[k] for k in rng.integers(0, 4, size=n_bp)]
    tw = math.radians(twist_deg)
    groove = 2.0 * math.asin(min(1.0, bp_dist / (2.0 * r_base)))

    beads = []        # (kind, position)
    nucs = [[], []]   # per strand: dicts {P, S, B, base}

    def base_pos(strand: int, k: int):
        th = k * tw + (0.0 if strand == 0 else groove)
        return np.array([r_base * math.cos(th), r_base * math.sin(th), k * rise])

    def place(strand: int, k: int, base: str):
        rB = base_pos(strand, k)
        rP = base_pos(1 - strand, k)
        v = (rP - rB) / np.linalg.norm(rP - rB)
        a = math.radians(_SBB_ANGLE[base])
        best = None
        for sgn in (1.0, -1.0):
            # rotate v by +-a about the helix axis (z): the sugar stays in the base plane
            c, s_ = math.cos(sgn * a), math.sin(sgn * a)
            u = np.array([c * v[0] - s_ * v[1], s_ * v[0] + c * v[1], 0.0])
            rS = rB + sb_bond * u
            if best is None or np.hypot(rS[0], rS[1]) > np.hypot(best[0], best[1]):
                best = rS
        return best, rB

    pairs = [(seq[k], _COMP[seq[k]]) for k in range(n_bp)]
    for strand in (0, 1):
        order = range(n_bp) if strand == 0 else range(n_bp - 1, -1, -1)
        prev_S = None
        for j, k in enumerate(order):
            base = pairs[k][strand]
            rS, rB = place(strand, k, base)
            nuc = {"base": base}
            if j > 0:
                # phosphate between the previous sugar and this one, half a rise towards the previous
                mid = 0.5 * (prev_S + rS)
                ang = math.atan2(mid[1], mid[0])
                zP = mid[2]
                nuc["P"] = len(beads)
                beads.append(("P", np.array([r_phos * math.cos(ang), r_phos * math.sin(ang), zP])))
            nuc["S"] = len(beads)
            beads.append(("S", rS))
            nuc["B"] = len(beads)
            beads.append((base, rB))
            nucs[strand].append(nuc)
            prev_S = rS
    n = len(beads)
    pos = np.array([b[1] for b in beads])
    pos -= pos.mean(axis=0)
    mass = np.array([DNA_MASS[b[0]] for b in beads])
    vel = maxwell_boltzmann(rng, mass, T)

    bonds, angles, dihedrals, stacks, stack_kind = [], [], [], [], []
    for strand in (0, 1):
        ns = nucs[strand]
        for j, nu in enumerate(ns):
            if j > 0:
                bonds.append((nu["P"], nu["S"]))
            bonds.append((nu["S"], nu["B"]))
            if j + 1 < len(ns):
                bonds.append((nu["S"], ns[j + 1]["P"]))
        for j, nu in enumerate(ns):
            if j > 0:
                angles.append((nu["P"], nu["S"], nu["B"]))
                if j + 1 < len(ns):
                    angles.append((nu["P"], nu["S"], ns[j + 1]["P"]))
            if j + 1 < len(ns):
                nx = ns[j + 1]
                angles.append((nu["B"], nu["S"], nx["P"]))
                angles.append((nu["S"], nx["P"], nx["S"]))
                dihedrals.append((nu["S"], nx["P"], nx["S"], ns[j + 2]["P"]) if j + 2 < len(ns) else None)
                if j > 0:
                    dihedrals.append((nu["P"], nu["S"], nx["P"], nx["S"]))
                stacks.append((nu["S"], nu["B"], nx["B"], nu.get("P", NIL), nx["S"], nx.get("P", NIL)))
                stack_kind.append(4 * _BASE_CODE[nu["base"]] + _BASE_CODE[nx["base"]])
    dihedrals = [d for d in dihedrals if d is not None]
    bonds = np.array(bonds); angles = np.array(angles); dihedrals = np.array(dihedrals)
    r0 = np.linalg.norm(pos[bonds[:, 1]] - pos[bonds[:, 0]], axis=1)
    local = [
        LocalSpec("BondLength", "3SPN2Bond", "bond", bonds, np.stack([np.full(len(bonds), 0.143), r0], axis=1)),
        LocalSpec("BondAngle", "Harmonic", "angle", angles,
                  np.stack([np.full(len(angles), 47.801), _angles(pos, angles)], axis=1)),
        LocalSpec("DihedralAngle", "Gaussian", "dihedral", dihedrals,
                  np.stack([np.full(len(dihedrals), -1.434), np.full(len(dihedrals), 0.3),
                            _dihedrals(pos, dihedrals)], axis=1)),
        LocalSpec("3SPN2BaseStacking", "3SPN2", "nucleotide", np.array(stacks, dtype=np.uint64).astype(np.uint32),
                  np.array(stack_kind, dtype=np.float64)[:, None]),
    ]
    all_nucs = nucs[0] + nucs[1]
    bases = np.array([nu["B"] for nu in all_nucs])
    bp_par = np.array([[_BASE_CODE[nu["base"]], nu["S"]] for nu in all_nucs], dtype=np.float64)
    cs_par = []
    for strand in (0, 1):
        ns = nucs[strand]
        for j, nu in enumerate(ns):
            b5 = ns[j - 1]["B"] if j > 0 else -1
            b3 = ns[j + 1]["B"] if j + 1 < len(ns) else -1
            cs_par.append([_BASE_CODE[nu["base"]], nu["S"], b5, b3])
    cs_par = np.array(cs_par, dtype=np.float64)
    phos = np.array([i for i, b in enumerate(beads) if b[0] == "P"])
    kinds = np.array([_BEAD_CODE[b[0]] for b in beads], dtype=np.float64)
    bp = GlobalSpec(potential="3SPN2BasePair", participants=bases, params=bp_par, partition="VerletList",
                    margin=margin, ignore_molecule="Nothing", ignore_within={"nucleotide": 3})
    cs = GlobalSpec(potential="3SPN2CrossStacking", participants=bases, params=cs_par, partition="VerletList",
                    margin=margin, ignore_molecule="Nothing", ignore_within={"nucleotide": 3})
    ev = GlobalSpec(potential="3SPN2ExcludedVolume", participants=np.arange(n), params=kinds[:, None],
                    partition="CellList", margin=margin, ignore_molecule="Nothing",
                    ignore_within={"bond": 1, "angle": 1, "dihedral": 1, "nucleotide": 1})
    dh = GlobalSpec(potential="DebyeHuckel", participants=phos, params=np.full((len(phos), 1), -0.6),
                    cutoff=5.5, partition="VerletList", margin=margin, ignore_molecule="Nothing",
                    ignore_within={"nucleotide": 1})
    return SystemSpec(boundary="unlimited", mass=mass, pos=pos, vel=vel,
                      attributes={"temperature": T, "ionic_strength": ionic_strength},
                      local=local, global_=[bp, cs, ev, dh],
                      integrator=IntegratorSpec("GJFNVTLangevin", dt, gamma * mass, seed=seed),
                      name=f"dna_duplex_{n_bp}bp")
