import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine through the C ABI)")


def load_golden(name: str):
    return dict(np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False))


def golden_case(name: str):
    """(spec as the golden vectors were made, golden dict); checks the input hash."""
    import make_golden
    from mjolnir_amd import workloads
    factory, _, _ = make_golden.cases()[name]
    spec = workloads.zero_temperature(factory())
    g = load_golden(name)
    assert make_golden.spec_hash(spec) == str(g["spec_hash"]), f"{name}: regenerated inputs differ from the golden inputs"
    return spec, g


def force_tolerance(ref: np.ndarray, rel: float, scale: float = 0.0) -> np.ndarray:
    """Per-particle bound rel * (|F_ref,i| + F_rms + scale): relative to the particle's own force,
    with the RMS force as the floor for particles whose force nearly cancels.  `scale` adds a
    floor from outside the array: one interaction's forces are compared on the scale of the total
    force, so an interaction whose forces all vanish (every bonded term of an ideal structure at
    its minimum) is held to rounding of the forces that act, not to zero."""
    mag = np.linalg.norm(ref, axis=1)
    rms = np.sqrt(np.mean(mag ** 2)) if len(mag) else 0.0
    return rel * (mag + rms + scale)


def rms_force(f: np.ndarray) -> float:
    return float(np.sqrt(np.mean(np.sum(f * f, axis=1)))) if len(f) else 0.0


def base_pair_energy_paths(names, per_interaction, total):
    """The reference computes the 3SPN2 base-pairing energy along two paths:
    ThreeSPN2BasePairInteraction::calc_energy (R/S vectors, ThreeSPN2BasePairInteraction.hpp:133-284;
    what EnergyObserver prints, and the golden per-interaction energy) and calc_force_and_energy
    (m/n vectors, :289-514; part of the golden total).  The oracle, which follows both paths bit for
    bit, gives the same value along both to about 4e-15 relative on the 200-bp duplex built at a 3.6 A
    rise and on perturbed duplexes, but values 1.1e-9 relative apart on the ideal B-form duplex at
    3.38 A (its planar base pairs put the S-B-B'-S' dihedral where acos amplifies rounding).  Returns
    {index: (calc_energy value, calc_force_and_energy value)} for the base-pairing terms; the second
    is the total less every other interaction's energy."""
    out = {}
    for k, n in enumerate(names):
        if str(n) == "3SPN2BasePair":
            rest = sum(float(per_interaction[q]) for q in range(len(names)) if q != k)
            out[k] = (float(per_interaction[k]), float(total) - rest)
    return out


@pytest.fixture(scope="session")
def gpu_lib():
    from mjolnir_amd import capi
    return capi.lib()
