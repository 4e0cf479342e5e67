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


def force_tolerance(ref: np.ndarray, rel: float) -> np.ndarray:
    """Per-particle bound rel * (|F_ref,i| + F_rms): relative to the particle's own force,
    with the RMS force as the floor for particles whose force nearly cancels."""
    mag = np.linalg.norm(ref, axis=1)
    rms = np.sqrt(np.mean(mag ** 2)) if len(mag) else 0.0
    return rel * (mag + rms)


@pytest.fixture(scope="session")
def gpu_lib():
    from mjolnir_amd import capi
    return capi.lib()
