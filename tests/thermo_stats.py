"""Statistics of thermostat tests: closed-form Boltzmann densities of the trimer gas
(``mjolnir_amd.workloads.trimer_gas``), a binned chi-square against them, and block averages.

Method of the reference's integration tests (test/integration/test_BAOABIntegrator/,
test_G-JF_Integrator/: a sampled density against exp(-U / kB T) normalized numerically), with the
comparison made quantitative: a chi-square over bins of at least `min_expected` expected counts,
passed at two standard deviations of its own distribution, dof + 2 sqrt(2 dof).
"""
from __future__ import annotations

import numpy as np


def bond_density(r: np.ndarray, k: float, r0: float, kT: float) -> np.ndarray:
    """Unnormalized density of a harmonic bond length in 3D: r^2 exp(-k (r - r0)^2 / kT)
    (the reference's harmonic bond has no factor 1/2, HarmonicPotential.hpp:23-32)."""
    return r * r * np.exp(-k * (r - r0) ** 2 / kT)


def angle_density(th: np.ndarray, k: float, sigma: float, th0: float, kT: float) -> np.ndarray:
    """Unnormalized density of a bond angle under U = k exp(-(theta - theta0)^2 / 2 sigma^2)
    (GaussianPotential.hpp:20-38): sin(theta) exp(-U / kT)."""
    u = k * np.exp(-((th - th0) ** 2) / (2.0 * sigma * sigma))
    return np.sin(th) * np.exp(-u / kT)


def bin_probabilities(density, lo: float, hi: float, edges: np.ndarray, fine: int = 400000) -> np.ndarray:
    """Probability of each bin [edges[b], edges[b+1]) under `density` normalized on [lo, hi]
    (midpoint rule on `fine` points; the mass outside [lo, hi] is negligible by construction)."""
    x = lo + (np.arange(fine) + 0.5) * (hi - lo) / fine
    w = density(x)
    w = w / w.sum()
    idx = np.clip(np.searchsorted(edges, x, side="right") - 1, -1, len(edges) - 1)
    p = np.zeros(len(edges) - 1)
    inside = (idx >= 0) & (idx < len(edges) - 1)
    np.add.at(p, idx[inside], w[inside])
    return p


def chi_square(samples: np.ndarray, density, lo: float, hi: float, n_bins: int = 60, min_expected: float = 25.0):
    """Chi-square of the samples' histogram against the density.  Bins are equal-width on
    [lo, hi]; neighbouring bins are merged from the tails inwards until each expects at least
    `min_expected` counts, and everything outside [lo, hi] joins the outer bins.  Returns
    (chi2, dof, bar) with bar = dof + 2 sqrt(2 dof)."""
    edges = np.linspace(lo, hi, n_bins + 1)
    p = bin_probabilities(density, lo, hi, edges)
    n = len(samples)
    counts = np.histogram(np.clip(samples, lo, np.nextafter(hi, lo)), bins=edges)[0].astype(np.float64)
    # merge bins with small expectation into their neighbours (left half leftwards, right half rightwards)
    groups = []
    cur_p, cur_c = 0.0, 0.0
    for b in range(n_bins):
        cur_p += p[b]
        cur_c += counts[b]
        if cur_p * n >= min_expected:
            groups.append([cur_p, cur_c])
            cur_p, cur_c = 0.0, 0.0
    if cur_p > 0 or cur_c > 0:
        groups[-1][0] += cur_p
        groups[-1][1] += cur_c
    g = np.array(groups)
    e = g[:, 0] * n
    chi2 = float(np.sum((g[:, 1] - e) ** 2 / e))
    dof = len(g) - 1
    return chi2, dof, dof + 2.0 * np.sqrt(2.0 * dof)


def block_mean(x: np.ndarray, n_blocks: int = 20):
    """Mean and its standard error from `n_blocks` block means (correlated samples)."""
    m = len(x) // n_blocks * n_blocks
    b = np.asarray(x[:m], dtype=np.float64).reshape(n_blocks, -1).mean(axis=1)
    return float(b.mean()), float(b.std(ddof=1) / np.sqrt(n_blocks))


def trimer_observables(spec, x: np.ndarray):
    """Bond lengths (both bonds of every trimer) and angles of one frame."""
    b = spec.local[0].indices
    a = spec.local[1].indices
    r = np.linalg.norm(x[b[:, 1]] - x[b[:, 0]], axis=1)
    u = x[a[:, 0]] - x[a[:, 1]]
    w = x[a[:, 2]] - x[a[:, 1]]
    c = np.einsum("ij,ij->i", u, w) / (np.linalg.norm(u, axis=1) * np.linalg.norm(w, axis=1))
    return r, np.arccos(np.clip(c, -1.0, 1.0))


def trimer_checks(spec, bonds: np.ndarray, angles: np.ndarray, kT: float):
    """Chi-square of the sampled bond lengths and angles against the closed forms at kT."""
    k, r0 = spec.local[0].params[0]
    ka, sg, th0 = spec.local[1].params[0]
    width = 6.0 * np.sqrt(kT / (2.0 * k))
    cb = chi_square(bonds, lambda r: bond_density(r, k, r0, kT), max(1e-3, r0 - width), r0 + width)
    ca = chi_square(angles, lambda t: angle_density(t, ka, sg, th0, kT), 0.0, np.pi)
    return cb, ca
