"""TOML front end (the reference's input schema, read_input_file.hpp and the read_* readers).

CPU: specs written with toml_input.dump and read back with toml_input.load come back bit for bit
(every array, every interaction setting), for the CG chain, an LJ fluid, the 3SPN2 duplex and the
constrained chain; ``include``, ``env`` and ``offset`` resolve as the reference's readers do.
GPU: an input file run through toml_input.run writes the reference's .ene file for the same
system (tests/golden/sim_chain200.npz, written by the reference's EnergyObserver)."""
import os

import numpy as np
import pytest

import make_golden
from mjolnir_amd import toml_input, workloads


def _same(a, b):
    assert np.array_equal(a.mass, b.mass) and np.array_equal(a.pos, b.pos) and np.array_equal(a.vel, b.vel)
    assert a.boundary == b.boundary and np.array_equal(a.lower, b.lower) and np.array_equal(a.upper, b.upper)
    assert a.attributes == b.attributes
    assert len(a.local) == len(b.local) and len(a.global_) == len(b.global_)
    for x, y in zip(a.local, b.local):
        assert (x.interaction, x.potential, x.topology) == (y.interaction, y.potential, y.topology)
        assert np.array_equal(x.indices, y.indices) and np.array_equal(x.params, y.params)
    for x, y in zip(a.global_, b.global_):
        assert (x.potential, x.cutoff, x.epsilon, x.partition, x.margin, x.ignore_molecule, x.ignore_within) == \
               (y.potential, y.cutoff, y.epsilon, y.partition, y.margin, y.ignore_molecule, y.ignore_within)
        assert np.array_equal(x.participants, y.participants) and np.array_equal(x.params, y.params)
    ia, ib = a.integrator, b.integrator
    assert (ia.type, ia.dt, ia.seed, ia.remove) == (ib.type, ib.dt, ib.seed, ib.remove)
    assert np.array_equal(ia.coefs, ib.coefs)
    assert (a.constraint is None) == (b.constraint is None)
    if a.constraint is not None:
        assert np.array_equal(a.constraint.indices, b.constraint.indices)
        assert np.array_equal(a.constraint.v0, b.constraint.v0)
        assert (a.constraint.topology, a.constraint.max_iteration, a.constraint.tolerance) == \
               (b.constraint.topology, b.constraint.max_iteration, b.constraint.tolerance)


@pytest.mark.parametrize("make", [
    lambda: workloads.cg_chain(n=60, seed=3, max_bend_deg=150.0),
    lambda: workloads.lj_fluid(n=64, seed=2),
    lambda: workloads.dna_duplex(12, seed=5, rise=3.6),
    lambda: workloads.constrained_chain(30, seed=3),
    lambda: make_golden.chain_with_remover(),
    lambda: make_golden.cases()["lj512_ul"][0](),
], ids=["chain", "lj", "dna", "constrained", "remover", "underdamped"])
def test_dump_load_round_trip(tmp_path, make):
    spec = make()
    p = str(tmp_path / "in.toml")
    toml_input.dump(spec, 1234, 100, p, prefix="run")
    inp = toml_input.load(p)
    _same(spec, inp.spec)
    assert (inp.total_step, inp.save_step) == (1234, 100)
    assert inp.output_prefix == os.path.join("./", "run")


def test_include_env_offset(tmp_path):
    """The reference's file layout: [simulator], [[systems]] and [[forcefields]] each pulled in
    with `include` from input.path; an `env` name and an index `offset` in a parameter table."""
    spec = workloads.cg_chain(n=20, seed=4, max_bend_deg=150.0)
    full = str(tmp_path / "full.toml")
    toml_input.dump(spec, 500, 50, full)
    text = open(full).read()
    # split the single file into the four files of the reference's examples
    head, rest = text.split("[simulator]\n", 1)
    sim, rest = rest.split("[[systems]]\n", 1)
    system, ff = rest.split("[[forcefields]]\n", 1)
    sub = tmp_path / "input"
    sub.mkdir()
    (sub / "sim.toml").write_text(sim)
    (sub / "system.toml").write_text(system)
    ff = ff.replace("[[forcefields.local]]", "[[local]]").replace("[[forcefields.global]]", "[[global]]")
    # the first bond: index offset and the spring constant through env
    bond = "{indices = [0, 1], k = 110.0,"
    assert bond in ff
    ff = ff.replace(bond, "{offset = 1, indices = [-1, 0], k = \"K\",", 1)
    ff = ff.replace('topology = "bond"\n', 'topology = "bond"\nenv.K = 110.0\n', 1)
    (sub / "ff.toml").write_text(ff)
    main = head.replace("[files]\n", '[files]\ninput.path = "input/"\n')
    main += '[simulator]\ninclude = "sim.toml"\n\n[[systems]]\ninclude = "system.toml"\n\n' \
            '[[forcefields]]\ninclude = "ff.toml"\n'
    (tmp_path / "main.toml").write_text(main)
    _same(spec, toml_input.load(str(tmp_path / "main.toml")).spec)


REF_LJ = "/root/reference/input/lennard-jones/lennard-jones.toml"


@pytest.mark.skipif(not os.path.exists(REF_LJ), reason="the reference's example inputs are not present")
def test_reference_lennard_jones_example(monkeypatch):
    """The reference's own example input (four files joined by `include`, input.path relative to
    the directory it is run from) reads as written: 4096 particles in a 32 A periodic box at
    300 K, LJ sigma 2 / epsilon 1 on a cell list with margin 0.25, UnderdampedLangevin with
    gamma 1, dt 0.01 and seed 2374."""
    monkeypatch.chdir(os.path.dirname(os.path.dirname(os.path.dirname(REF_LJ))))
    inp = toml_input.load("input/lennard-jones/lennard-jones.toml")
    s = inp.spec
    assert (inp.total_step, inp.save_step) == (10000, 100)
    assert inp.output_prefix == os.path.join("./", "lennard-jones")
    assert s.n == 4096 and s.boundary == "periodic"
    assert np.array_equal(s.lower, [0, 0, 0]) and np.array_equal(s.upper, [32, 32, 32])
    assert s.attributes["temperature"] == 300.0
    assert np.all(s.mass == 1.0)
    assert np.array_equal(s.pos[0], [1.0, 1.0, 1.0]) and np.array_equal(s.vel[0], [0.5101, 0.7734, -2.0929])
    assert np.array_equal(s.pos[1], [3.0, 1.0, 1.0])
    it = s.integrator
    assert (it.type, it.dt, it.seed) == ("UnderdampedLangevin", 0.01, 2374)
    assert np.all(it.coefs == 1.0) and len(it.coefs) == 4096
    assert len(s.local) == 0 and len(s.global_) == 1
    g = s.global_[0]
    assert (g.potential, g.partition, g.margin, g.ignore_molecule) == ("LennardJones", "CellList", 0.25, "Nothing")
    assert np.array_equal(g.participants, np.arange(4096))
    assert np.all(g.params == [2.0, 1.0])


def test_rejects_what_it_does_not_run(tmp_path):
    spec = workloads.lj_fluid(n=27, seed=2)
    p = tmp_path / "in.toml"
    toml_input.dump(spec, 10, 10, str(p))
    p.write_text(p.read_text().replace('energy = "kcal/mol"', 'energy = "kJ/mol"'))
    with pytest.raises(ValueError):
        toml_input.load(str(p))


@pytest.mark.gpu
def test_toml_run_writes_the_reference_energy_file(tmp_path):
    from conftest import GOLDEN
    g = np.load(os.path.join(GOLDEN, "sim_chain200.npz"))
    spec = workloads.zero_temperature(make_golden.cases()["chain200"][0]())
    assert str(g["spec_hash"]) == make_golden.spec_hash(spec)
    p = str(tmp_path / "chain200.toml")
    toml_input.dump(spec, int(g["total"]), int(g["save"]), p, prefix=str(tmp_path / "run"))
    print("[gpu] toml run chain200", flush=True)
    ene = toml_input.run(p)
    got, ref = open(ene).read().splitlines(), str(g["ene"]).splitlines()
    assert len(got) == len(ref) and got[:2] == ref[:2]
    for a, b in zip(got[2:], ref[2:]):
        x = np.array([float(t) for t in a.split()])
        y = np.array([float(t) for t in b.split()])
        assert np.allclose(x, y, rtol=1e-9, atol=2e-6), (a, b)
