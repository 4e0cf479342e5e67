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

from .spec import GlobalSpec, IntegratorSpec, LocalSpec, SystemSpec

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
             dt: float = 0.01, seed: int = 1, jitter: float = 0.1) -> SystemSpec:
    """N = n_side**3 identical LJ particles on a jittered simple-cubic lattice in a periodic cube.

    ``n_side=16`` is the 4096-particle LJ fluid (rho sigma^3 = 0.6, L = 75.880 A).
    """
    rng = np.random.default_rng(seed)
    n = n_side ** 3
    rho = rho_reduced / sigma ** 3
    L = (n / rho) ** (1.0 / 3.0)
    a = L / n_side
    g = np.arange(n_side, dtype=np.float64)
    gx, gy, gz = np.meshgrid(g, g, g, indexing="ij")
    lattice = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1) * a + 0.5 * a
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
    workers = min(16, os.cpu_count() or 1)
    if workers > 1 and n_replicas >= 32:
        # every replica has its own seed, so the result does not depend on the pool
        import multiprocessing as mp
        with mp.get_context("fork").Pool(workers) as pool:
            return pool.map(_ensemble_member, jobs, chunksize=4)
    return [_ensemble_member(j) for j in jobs]


def _ensemble_member(job) -> SystemSpec:
    n, seed, T, max_bend_deg = job
    return cg_chain(n, seed=seed, T=T, max_bend_deg=max_bend_deg)


def zero_temperature(spec: SystemSpec) -> SystemSpec:
    """Copy of ``spec`` with gamma (or alpha) = 0, which zeroes the friction and the noise
    amplitude: BAOAB and G-JF then reduce to velocity Verlet
    (BAOABLangevinIntegrator.hpp:79-82, GJFNVTLangevinIntegrator.hpp:116-118).  The
    temperature attribute is kept because Debye-Hueckel parameters depend on it."""
    import copy
    s = copy.deepcopy(spec)
    if s.integrator is not None:
        s.integrator.coefs = np.zeros_like(s.integrator.coefs)
    return s
