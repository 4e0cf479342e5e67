"""The drop-in boundary in C++: include/mjg_forcefield.hpp puts libmjg.so behind the reference's own
ForceFieldBase and integrator shapes, and oracle/_ref/dropin_check (built by oracle/Makefile from the
reference sources and libmjg.so) drives it with the reference's own classes:

* forces: every ForceFieldBase call through mjg::GpuForceField against the reference ForceField on
  the same input (calc_force, calc_energy, calc_force_and_energy at 1e-10; format_energy_name equal;
  format_energy columns of the same widths and values; calc_force_and_virial: NaN virial and
  MJG_UNSUPPORTED, the declared choice);
* run: the reference's BAOAB / G-JF / UnderdampedLangevin integrator stepping either force field for
  1000 steps (zero temperature, friction on): positions within 1e-8 A, or ten times the reference's
  own sensitivity where that is larger;
* sim: the reference's MolecularDynamicsSimulator with mjg::GpuForceField and mjg::GpuBAOAB writes,
  through the reference's EnergyObserver, the .ene the reference writes.

The checker needs the reference sources at build time; without the binary these tests skip.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, golden_case, load_golden

CHECK = os.path.join(ROOT, "oracle", "_ref", "dropin_check")
needs_check = pytest.mark.skipif(not os.path.exists(CHECK), reason="oracle/_ref/dropin_check not built (reference absent)")


def _spec_file(tmp_path, spec):
    from mjolnir_amd.spec import write_mjspec
    p = str(tmp_path / "in.mjspec")
    write_mjspec(spec, p)
    return p


def _run(args, timeout=600):
    print(f"[dropin] {' '.join(args[:2])}", flush=True)
    r = subprocess.run([CHECK] + args, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr
    return json.loads([l for l in r.stdout.splitlines() if l.startswith("{")][-1])


@needs_check
def test_checker_links_and_reports_a_missing_device(tmp_path):
    """CPU: the checker loads libmjg.so and the reference build together; without a GPU the engine's
    creation fails with a status, not a crash."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    spec, _ = golden_case("lj512_damped")
    r = subprocess.run([CHECK, "forces", _spec_file(tmp_path, spec)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "mjg_engine_create" in r.stderr


FORCE_CASES = ["lj512_damped", "chain200", "dna12", "lj_boundary", "two_chains_self", "chain100_gbaoab"]
RUN_CASES = ["lj512_damped", "chain200_damped", "lj512_gjf_damped", "lj512_ul_damped", "dna12_damped", "chain200"]


@needs_check
@pytest.mark.gpu
@pytest.mark.parametrize("name", FORCE_CASES)
def test_forcefieldbase_calls_vs_reference(tmp_path, name):
    spec, _ = golden_case(name)
    r = _run(["forces", _spec_file(tmp_path, spec)])
    assert r["force_rel_err"] <= 1e-10 and r["force_and_energy_rel_err"] <= 1e-10, r
    for a, b in (("energy_ref", "energy_gpu"), ("force_and_energy_ref", "force_and_energy_gpu"),
                 ("format_total_ref", "format_total_gpu")):
        assert abs(r[a] - r[b]) <= 1e-10 * max(1.0, abs(r[a])), (a, r)
    assert r["names_equal"], r
    assert r["format_line_equal_width"] and r["format_numbers"] == len(spec.interaction_names())
    assert r["format_rel_err"] <= 2e-6      # six printed decimals
    assert r["virial_nan"] and r["virial_status"] == 6


@needs_check
@pytest.mark.gpu
@pytest.mark.parametrize("name", RUN_CASES)
def test_reference_integrator_with_gpu_forcefield(tmp_path, name):
    spec, g = golden_case(name)
    r = _run(["run", _spec_file(tmp_path, spec), "1000"])
    bar = max(1e-8, 10 * float(np.max(g["zt_sens"])))
    print(f"[dropin] {name}: max |dx| {r['max_abs_dx']:.2e} (bar {bar:.1e})", flush=True)
    assert r["status"] == 0
    assert r["max_abs_dx"] <= bar, r


@needs_check
@pytest.mark.gpu
def test_reference_simulator_with_gpu_integrator_writes_reference_ene(tmp_path):
    import make_golden
    from mjolnir_amd import workloads
    g = np.load(os.path.join(GOLDEN, "sim_chain200.npz"))
    spec = workloads.zero_temperature(make_golden.cases()["chain200"][0]())
    assert str(g["spec_hash"]) == make_golden.spec_hash(spec)
    prefix = str(tmp_path / "run")
    _run(["sim", _spec_file(tmp_path, spec), str(int(g["total"])), str(int(g["save"])), prefix])
    got = open(prefix + ".ene").read().splitlines()
    ref = str(g["ene"]).splitlines()
    assert len(got) == len(ref) and got[:2] == ref[:2]
    for a, b in zip(got[2:], ref[2:]):
        x = np.array([float(t) for t in a.split()])
        y = np.array([float(t) for t in b.split()])
        assert np.allclose(x, y, rtol=1e-9, atol=2e-6), (a, b)
        assert len(a) == len(b)
