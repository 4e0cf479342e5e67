"""Trajectory files in the reference's layouts for states that live on the host: the frames of
a batch replica (or any state downloaded from an engine), written the way XYZObserver
(XYZObserver.hpp:24-107) and DCDObserver (DCDObserver.hpp:27-341) write them.

The FP64 engine's simulator loop writes the same files on its own (mjg_observer_xyz_open,
mjg_observer_dcd_open); these writers serve the batch engine, whose replicas are advanced in
segments and read back with ``Batch.state``.
"""
from __future__ import annotations

import struct
from typing import List, Optional, Sequence

import numpy as np

# 80-byte title record of the DCD header (any text; readers show it as a remark)
DCD_TITLE = b"libmjg (MI355X) trajectory, Mjolnir DCDObserver layout".ljust(80, b"\0")


def n_molecules(spec) -> int:
    """Topology::construct_molecules (Topology.hpp:280-317) over the "bond" connections of the
    spec's local interactions and constraints: a particle joins the molecule of the first bonded
    neighbour that already has one, in the order the connections were added."""
    n = spec.n
    adj: List[List[int]] = [[] for _ in range(n)]

    def add(i, j):
        if j not in adj[i]:
            adj[i].append(j)
        if i not in adj[j]:
            adj[j].append(i)

    for loc in spec.local:
        if loc.topology != "bond":
            continue
        for idx in np.asarray(loc.indices).reshape(len(loc.indices), -1):
            ids = [int(x) for x in idx if int(x) != 0xFFFFFFFF]
            for a in range(len(ids)):
                for b in range(a + 1, len(ids)):
                    add(ids[a], ids[b])
    if spec.constraint is not None and spec.constraint.topology == "bond":
        for i, j in np.asarray(spec.constraint.indices):
            add(int(i), int(j))
    mol = [-1] * n
    nxt = 0
    for i in range(n):
        for j in adj[i]:
            if mol[j] != -1:
                mol[i] = mol[j]
                break
        if mol[i] == -1:
            mol[i] = nxt
            nxt += 1
    return nxt


class XYZWriter:
    """prefix_position.xyz and prefix_velocity.xyz: per frame "N", "step = S", then
    "name x y z" per particle in fixed notation with 8 decimals (XYZObserver::output)."""

    def __init__(self, prefix: str, names: Optional[Sequence[str]] = None):
        self.paths = (prefix + "_position.xyz", prefix + "_velocity.xyz")
        self.names = None if names is None else [str(x) for x in names]
        for p in self.paths:
            open(p, "w").close()

    def write(self, step: int, pos: np.ndarray, vel: np.ndarray):
        for path, a in zip(self.paths, (pos, vel)):
            a = np.asarray(a, dtype=np.float64).reshape(-1, 3)
            names = self.names or ["X"] * len(a)
            lines = [f"{len(a)}", f"step = {int(step)}"]
            lines += [f"{nm} {x:.8f} {y:.8f} {z:.8f}" for nm, (x, y, z) in zip(names, a)]
            with open(path, "a") as f:
                f.write("\n".join(lines) + "\n")


class DCDWriter:
    """prefix_position.dcd and prefix_velocity.dcd: the header (DCDObserver::write_header),
    float32 frames with the unit-cell record in the position file of a periodic box, and the
    frame count written by close() (DCDObserver::finalize)."""

    def __init__(self, prefix: str, n: int, total_step: int, save_step: int, dt: float, molecules: int,
                 box_width: Optional[Sequence[float]] = None):
        self.paths = (prefix + "_position.dcd", prefix + "_velocity.dcd")
        self.n = int(n)
        self.width = None if box_width is None else [float(w) for w in box_width]
        self.frames = 0
        head = struct.pack("<i", 84) + b"CORD" + struct.pack("<9i", 0, 0, int(save_step), int(total_step),
                                                               int(molecules), 0, 0, 0, 0)
        head += struct.pack("<f", float(dt)) + struct.pack("<i", 1 if self.width is not None else 0)
        head += struct.pack("<8i", *([0] * 8)) + struct.pack("<i", 24) + struct.pack("<i", 84)
        head += struct.pack("<ii", 84, 1) + DCD_TITLE + struct.pack("<i", 84)
        head += struct.pack("<iii", 4, self.n, 4)
        for p in self.paths:
            with open(p, "wb") as f:
                f.write(head)

    def write(self, pos: np.ndarray, vel: np.ndarray):
        self.frames += 1
        nb = 4 * self.n
        for k, (path, a) in enumerate(zip(self.paths, (pos, vel))):
            a = np.asarray(a).reshape(-1, 3)
            out = bytearray()
            if k == 0 and self.width is not None:
                w = self.width
                out += struct.pack("<i6di", 48, w[0], 90.0, w[1], 90.0, 90.0, w[2], 48)
            for c in range(3):
                out += struct.pack("<i", nb) + a[:, c].astype("<f4").tobytes() + struct.pack("<i", nb)
            with open(path, "ab") as f:
                f.write(out)

    def close(self):
        for p in self.paths:
            with open(p, "r+b") as f:
                f.seek(8)
                f.write(struct.pack("<i", self.frames))


def batch_replica_frames(batch, r: int, total_step: int, save_step: int, prefix: str, formats=("xyz", "dcd")):
    """Advance replica ``r`` of ``batch`` (the whole batch steps) by ``total_step`` steps, writing
    a frame of it at every multiple of ``save_step`` and after the last step, as the simulator
    loop's observers do."""
    spec = batch.specs[r]
    xyz = XYZWriter(prefix, spec.particle_names) if "xyz" in formats else None
    dcd = None
    if "dcd" in formats:
        width = (spec.upper - spec.lower) if spec.boundary == "periodic" else None
        dcd = DCDWriter(prefix, spec.n, total_step, save_step, spec.integrator.dt, n_molecules(spec), width)

    def frame(step):
        x, v, _ = batch.state(r)
        if xyz:
            xyz.write(step, x, v)
        if dcd:
            dcd.write(x, v)

    done = 0
    while done < total_step:
        if done % save_step == 0:
            frame(done)
        nxt = min(total_step, (done // save_step + 1) * save_step)
        batch.step(nxt - done)
        done = nxt
    frame(total_step)
    if dcd:
        dcd.close()
