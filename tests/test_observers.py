"""Trajectory observers of the simulator loop: XYZObserver and DCDObserver (XYZObserver.hpp,
DCDObserver.hpp), with the EnergyObserver beside them.

The fixtures (tests/golden/sim_*.npz) hold the files the reference's own observers wrote over a
300-step zero-temperature run, output every 100 steps: a 200-bead chain in an unlimited box
(BAOAB) and 512 LJ particles in a periodic box (UnderdampedLangevin).

CPU: the reference's DCD and XYZ bytes are parsed by the independent reader below, which pins
the layout the engine writes -- the Fortran record marks, the control record (frame count, save
interval, total steps, molecules, dt, unit-cell flag, version 24), the unit-cell record of a
periodic box, float32 coordinates -- and the first frame is the system's initial state.
GPU: the engine's files for the same run: identical XYZ text structure and DCD header bytes
(except the 80-byte title record, whose text is the writer's own), coordinates within the
zero-temperature trajectory bar."""
import os
import struct

import numpy as np
import pytest

from conftest import GOLDEN
import make_golden
from mjolnir_amd import toml_input, workloads


def read_dcd(data: bytes):
    """Header dict and frames [(unitcell or None, (3, N) float32)] of a DCD file."""
    def rec(off):
        (n,) = struct.unpack_from("<i", data, off)
        body = data[off + 4: off + 4 + n]
        (m,) = struct.unpack_from("<i", data, off + 4 + n)
        assert m == n, "record marks differ"
        return body, off + 8 + n
    ctl, off = rec(0)
    assert ctl[:4] == b"CORD" and len(ctl) == 84
    ints = struct.unpack_from("<9i", ctl, 4)
    (dt,) = struct.unpack_from("<f", ctl, 40)
    tail = struct.unpack_from("<10i", ctl, 44)
    hdr = {"frames": ints[0], "first": ints[1], "interval": ints[2], "total": ints[3], "molecules": ints[4],
           "zeros": ints[5:9], "dt": dt, "unitcell": tail[0], "zeros8": tail[1:9], "version": tail[9]}
    title, off = rec(off)
    assert len(title) == 84 and struct.unpack_from("<i", title, 0)[0] == 1
    natom, off = rec(off)
    (hdr["n"],) = struct.unpack("<i", natom)
    frames = []
    while off < len(data):
        cell = None
        (n,) = struct.unpack_from("<i", data, off)
        if n == 48:
            body, off = rec(off)
            cell = struct.unpack("<6d", body)
        xyz = []
        for _ in range(3):
            body, off = rec(off)
            xyz.append(np.frombuffer(body, dtype="<f4"))
        frames.append((cell, np.stack(xyz)))
    return hdr, frames


def read_xyz(text: str):
    """[(step, names, (N, 3) array)] of an XYZ file."""
    lines = text.splitlines()
    out, k = [], 0
    while k < len(lines):
        n = int(lines[k])
        assert lines[k + 1].startswith("step = ")
        step = int(lines[k + 1][7:])
        rows = [l.split() for l in lines[k + 2: k + 2 + n]]
        out.append((step, [r[0] for r in rows], np.array([[float(x) for x in r[1:]] for r in rows])))
        k += 2 + n
    return out


def _case(name):
    g = np.load(os.path.join(GOLDEN, name + ".npz"))
    case, total, save = make_golden.sim_cases()[name]
    spec = workloads.zero_temperature(make_golden.cases()[case][0]())
    assert str(g["spec_hash"]) == make_golden.spec_hash(spec)
    return spec, g, int(total), int(save)


@pytest.mark.parametrize("name", list(make_golden.sim_cases().keys()))
def test_reference_dcd_layout(name):
    spec, g, total, save = _case(name)
    periodic = spec.boundary == "periodic"
    for key, state in (("dcd_pos", spec.pos), ("dcd_vel", spec.vel)):
        hdr, frames = read_dcd(g[key].tobytes())
        assert hdr["frames"] == total // save + 1 == len(frames)
        assert (hdr["first"], hdr["interval"], hdr["total"], hdr["version"]) == (0, save, total, 24)
        assert hdr["n"] == spec.n and hdr["dt"] == np.float32(spec.integrator.dt)
        assert hdr["unitcell"] == int(periodic) and not any(hdr["zeros"]) and not any(hdr["zeros8"])
        # chains: one molecule per bonded chain (Topology::construct_molecules); LJ: one per particle
        assert hdr["molecules"] == (1 if spec.local else spec.n)
        for cell, _ in frames:
            if periodic and key == "dcd_pos":
                w = spec.upper - spec.lower
                assert cell == (w[0], 90.0, w[1], 90.0, 90.0, w[2])
            else:
                assert cell is None
        assert np.array_equal(frames[0][1], state.T.astype(np.float32))


@pytest.mark.parametrize("name", list(make_golden.sim_cases().keys()))
def test_reference_xyz_layout(name):
    spec, g, total, save = _case(name)
    for key, state in (("xyz_pos", spec.pos), ("xyz_vel", spec.vel)):
        frames = read_xyz(str(g[key]))
        assert [f[0] for f in frames] == list(range(0, total + 1, save))
        assert all(f[1] == ["X"] * spec.n for f in frames)
        assert np.array_equal(frames[0][2], np.round(state, 8))
    # the DCD frames and the XYZ frames are the same states
    x = read_xyz(str(g["xyz_pos"]))
    _, d = read_dcd(g["dcd_pos"].tobytes())
    for (_, _, a), (_, b) in zip(x, d):
        assert np.allclose(a, b.T, atol=1e-5 * max(1.0, np.abs(a).max()))


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(make_golden.sim_cases().keys()))
def test_engine_observers_vs_reference(tmp_path, name):
    spec, g, total, save = _case(name)
    p = str(tmp_path / (name + ".toml"))
    prefix = str(tmp_path / "run")
    toml_input.dump(spec, total, save, p, prefix=prefix, formats=("xyz", "dcd"))
    print(f"[gpu] observers {name}", flush=True)
    toml_input.run(p)
    # DCD: header bytes identical but for the title text (bytes 100-180); frames close
    for key, suffix in (("dcd_pos", "_position.dcd"), ("dcd_vel", "_velocity.dcd")):
        ref = g[key].tobytes()
        got = open(prefix + suffix, "rb").read()
        assert len(got) == len(ref)
        hdr_end = 92 + 92 + 12
        assert got[:100] == ref[:100] and got[180:hdr_end] == ref[180:hdr_end]
        _, fr = read_dcd(ref)
        _, fg = read_dcd(got)
        for (cr, ar), (cg, ag) in zip(fr, fg):
            assert cr == cg
            assert np.allclose(ag, ar, rtol=1e-6, atol=1e-6)
    # XYZ: same lines but the last printed digits
    for key, suffix in (("xyz_pos", "_position.xyz"), ("xyz_vel", "_velocity.xyz")):
        ref = read_xyz(str(g[key]))
        got = read_xyz(open(prefix + suffix).read())
        assert [(s, n) for s, n, _ in got] == [(s, n) for s, n, _ in ref]
        for (_, _, a), (_, _, b) in zip(got, ref):
            assert np.abs(a - b).max() <= 1e-7
    # energy file beside them
    got, ref = open(prefix + ".ene").read().splitlines(), str(g["ene"]).splitlines()
    assert len(got) == len(ref) and got[:2] == ref[:2]


# ----------------------------------------------------------------------------------------
# host-side writers (batch replicas): byte layout against the reference's files
# ----------------------------------------------------------------------------------------
@pytest.mark.parametrize("name", list(make_golden.sim_cases().keys()))
def test_host_writers_reproduce_reference_files(tmp_path, name):
    """Fed the reference's own frames (the DCD's float32 values; the initial state for the first
    XYZ frame), the host writers give the reference's DCD bytes but the title record, and its
    XYZ lines; the molecule count comes from the spec's bonds as construct_molecules counts."""
    from mjolnir_amd import trajectory
    spec, g, total, save = _case(name)
    _, fp = read_dcd(g["dcd_pos"].tobytes())
    _, fv = read_dcd(g["dcd_vel"].tobytes())
    width = (spec.upper - spec.lower) if spec.boundary == "periodic" else None
    prefix = str(tmp_path / "w")
    w = trajectory.DCDWriter(prefix, spec.n, total, save, spec.integrator.dt, trajectory.n_molecules(spec), width)
    for (_, p), (_, v) in zip(fp, fv):
        w.write(p.T, v.T)
    w.close()
    for key, suffix in (("dcd_pos", "_position.dcd"), ("dcd_vel", "_velocity.dcd")):
        ref = g[key].tobytes()
        got = open(prefix + suffix, "rb").read()
        assert len(got) == len(ref)
        assert got[:100] == ref[:100] and got[180:] == ref[180:]
    x = trajectory.XYZWriter(prefix)
    x.write(0, spec.pos, spec.vel)
    for key, suffix in (("xyz_pos", "_position.xyz"), ("xyz_vel", "_velocity.xyz")):
        n_lines = spec.n + 2
        assert open(prefix + suffix).read().splitlines() == str(g[key]).splitlines()[:n_lines]


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(make_golden.sim_cases().keys()))
def test_batch_replica_frames_vs_reference(tmp_path, name):
    """The same run as one FP64 batch replica, its frames written by the host writers between
    batch steps: the reference's DCD header bytes but the title, its XYZ steps and names, and
    coordinates within the zero-temperature trajectory bar."""
    from mjolnir_amd import capi, trajectory
    spec, g, total, save = _case(name)
    b = capi.Batch(capi.FP64)
    r = b.add(spec)
    b.initialize()
    prefix = str(tmp_path / "rep")
    print(f"[gpu] batch replica frames {name}", flush=True)
    trajectory.batch_replica_frames(b, r, total, save, prefix)
    assert b.info(r)["step"] == total
    for key, suffix in (("dcd_pos", "_position.dcd"), ("dcd_vel", "_velocity.dcd")):
        ref = g[key].tobytes()
        got = open(prefix + suffix, "rb").read()
        assert len(got) == len(ref)
        hdr_end = 92 + 92 + 12
        assert got[:100] == ref[:100] and got[180:hdr_end] == ref[180:hdr_end]
        for (cr, ar), (cg, ag) in zip(read_dcd(ref)[1], read_dcd(got)[1]):
            assert cr == cg
            assert np.allclose(ag, ar, rtol=1e-6, atol=1e-6)
    # the replica kernel rounds each force contribution to 2^-32 kcal/mol/A (exact int64 sums);
    # over 300 steps that moves positions by ~1e-8 and velocities, which follow the force
    # directly, by up to a few 1e-7: the velocity file is held to the DCD frames' 1e-6
    for key, suffix, bar in (("xyz_pos", "_position.xyz", 1e-7), ("xyz_vel", "_velocity.xyz", 1e-6)):
        ref = read_xyz(str(g[key]))
        got = read_xyz(open(prefix + suffix).read())
        assert [(s, n) for s, n, _ in got] == [(s, n) for s, n, _ in ref]
        for (_, _, a), (_, _, c) in zip(got, ref):
            assert np.abs(a - c).max() <= bar
    b.close()


def test_molecule_count_of_mixed_topology():
    """construct_molecules on two bonded chains (10 and 6 beads), an angle written with its
    indices out of order whose "bond" edges join two more beads, a "contact" edge (not a bond)
    and free particles.  On this topology the reference's greedy pass (next test) and the
    connected components of the bond graph, found here by breadth-first search, agree; the
    check is that contact edges join nothing and angle edges do."""
    from collections import deque
    from mjolnir_amd import trajectory
    from mjolnir_amd.spec import LocalSpec, SystemSpec
    n = 24
    bonds = [(i, i + 1) for i in range(9)] + [(i, i + 1) for i in range(10, 15)]
    spec = SystemSpec("unlimited", np.ones(n), np.zeros((n, 3)), np.zeros((n, 3)))
    spec.local.append(LocalSpec("BondLength", "Harmonic", "bond", np.array(bonds, dtype=np.uint32),
                                np.ones((len(bonds), 2))))
    spec.local.append(LocalSpec("BondAngle", "Gaussian", "bond", np.array([[18, 16, 17]], dtype=np.uint32),
                                np.ones((1, 3))))
    spec.local.append(LocalSpec("Contact", "GoContact", "contact", np.array([[20, 21]], dtype=np.uint32),
                                np.ones((1, 2))))
    adj = {i: set() for i in range(n)}
    for a, b in bonds + [(18, 16), (18, 17), (16, 17)]:
        adj[a].add(b)
        adj[b].add(a)
    seen, components = set(), 0
    for s in range(n):
        if s in seen:
            continue
        components += 1
        todo = deque([s])
        seen.add(s)
        while todo:
            for t in adj[todo.popleft()]:
                if t not in seen:
                    seen.add(t)
                    todo.append(t)
    assert components == 2 + 1 + 5      # two chains, the angle's three beads, 19..23 free
    assert trajectory.n_molecules(spec) == components


def test_molecule_count_follows_the_single_greedy_pass():
    """construct_molecules (Topology.hpp:280-317) visits particles once in index order and gives
    a particle the molecule of its first bonded neighbour that already has one, else a new id.
    It is not a connected-components count: bonds (0, 2) and (1, 2) form one component, but
    particles 0 and 1 are visited before 2 has an id, so the pass opens two molecules and 2 joins
    the first.  Bonds (0, 1), (1, 2) are one molecule either way; a chain bonded from its end,
    (2, 1), (1, 0), is one as well (0 opens it, 1 and 2 join in turn)."""
    from mjolnir_amd import trajectory
    from mjolnir_amd.spec import LocalSpec, SystemSpec

    def count(n, bonds):
        spec = SystemSpec("unlimited", np.ones(n), np.zeros((n, 3)), np.zeros((n, 3)))
        spec.local.append(LocalSpec("BondLength", "Harmonic", "bond", np.array(bonds, dtype=np.uint32),
                                    np.ones((len(bonds), 2))))
        return trajectory.n_molecules(spec)

    assert count(3, [(0, 2), (1, 2)]) == 2
    assert count(3, [(0, 1), (1, 2)]) == 1
    assert count(3, [(2, 1), (1, 0)]) == 1
    # a star whose centre has the highest index: every leaf opens its own molecule
    assert count(5, [(0, 4), (1, 4), (2, 4), (3, 4)]) == 4
    # four beads 0-3-1-2: 0 opens, 1 is visited before 3 has an id and opens a second,
    # 2 joins 1's, 3 joins 0's
    assert count(4, [(0, 3), (3, 1), (1, 2)]) == 2
