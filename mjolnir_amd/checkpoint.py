"""MsgPack checkpoints in the reference's format (MsgPackSaver.hpp:18-160, loaded by
MsgPackLoader.hpp:45-627 through read_system.hpp:84-124).

``<prefix>_system.msg`` holds the system: a fixmap of five entries in this order -- "boundary"
(nil for an unlimited box, else {"lower": [x, y, z], "upper": [x, y, z]}), "particles" (one
fixmap of six per particle: "mass", "position", "velocity", "force", "name", "group"), "virial"
(nine reals), "attributes" (string -> real, in key order) and "dynamic_variables" (empty here).
Reals are big-endian float64; arrays and maps use the shortest MsgPack header, as the
reference's saver does, so a file written here is byte-identical to the reference's for the
same state and either side can load the other's.

``<prefix>_rng.msg`` holds the random stream as {"internal_state": string}.  The reference
stores its mt19937 state there; the engines here draw counter-based Philox noise, whose whole
state is the seed and the step count, written as "philox4x32-10 <seed> <step>".
"""
from __future__ import annotations

import struct
from typing import Dict, Optional

import numpy as np

_F64 = struct.Struct(">d")


def _str(out: bytearray, s: str):
    b = s.encode()
    n = len(b)
    if n < 32:
        out.append(0xA0 + n)
    elif n <= 0xFF:
        out += bytes((0xD9, n))
    elif n <= 0xFFFF:
        out.append(0xDA)
        out += struct.pack(">H", n)
    else:
        out.append(0xDB)
        out += struct.pack(">I", n)
    out += b


def _real(out: bytearray, x: float):
    out.append(0xCB)
    out += _F64.pack(float(x))


def _vec3(out: bytearray, v):
    out.append(0x93)
    for x in v:
        _real(out, x)


def _container_len(out: bytearray, n: int, fix: int, c16: int, c32: int, fix_max: int = 16):
    if n < fix_max:
        out.append(fix + n)
    elif n < 65536:
        out.append(c16)
        out += struct.pack(">H", n)
    else:
        out.append(c32)
        out += struct.pack(">I", n)


def encode_system(mass, pos, vel, force, lower=None, upper=None, attributes: Optional[Dict[str, float]] = None,
                  names=None, groups=None, virial=None) -> bytes:
    """The bytes of a ``_system.msg`` file.  ``lower``/``upper`` None: unlimited boundary."""
    n = len(mass)
    out = bytearray([0x85])
    _str(out, "boundary")
    if lower is None:
        out.append(0xC0)
    else:
        out.append(0x82)
        _str(out, "lower")
        _vec3(out, lower)
        _str(out, "upper")
        _vec3(out, upper)
    _str(out, "particles")
    _container_len(out, n, 0x90, 0xDC, 0xDD)
    for i in range(n):
        out.append(0x86)
        _str(out, "mass"); _real(out, mass[i])
        _str(out, "position"); _vec3(out, pos[i])
        _str(out, "velocity"); _vec3(out, vel[i])
        _str(out, "force"); _vec3(out, force[i])
        _str(out, "name"); _str(out, names[i] if names is not None else "X")
        _str(out, "group"); _str(out, groups[i] if groups is not None else "NONE")
    _str(out, "virial")
    out.append(0x99)
    for x in (np.zeros(9) if virial is None else np.asarray(virial, dtype=np.float64).reshape(9)):
        _real(out, x)
    _str(out, "attributes")
    attrs = dict(sorted((attributes or {}).items()))
    _container_len(out, len(attrs), 0x80, 0xDE, 0xDF)
    for k, v in attrs.items():
        _str(out, k)
        _real(out, v)
    _str(out, "dynamic_variables")
    out.append(0x80)
    return bytes(out)


def decode_system(data: bytes) -> dict:
    """Arrays of a ``_system.msg`` file (either side's): mass (N,), pos/vel/force (N, 3),
    names, groups, lower/upper (None if unlimited), virial (9,), attributes."""
    import msgpack
    d = msgpack.unpackb(data, raw=False)
    parts = d["particles"]
    b = d["boundary"]
    return {
        "mass": np.array([p["mass"] for p in parts], dtype=np.float64),
        "pos": np.array([p["position"] for p in parts], dtype=np.float64).reshape(-1, 3),
        "vel": np.array([p["velocity"] for p in parts], dtype=np.float64).reshape(-1, 3),
        "force": np.array([p["force"] for p in parts], dtype=np.float64).reshape(-1, 3),
        "names": [p["name"] for p in parts],
        "groups": [p["group"] for p in parts],
        "lower": None if b is None else np.array(b["lower"], dtype=np.float64),
        "upper": None if b is None else np.array(b["upper"], dtype=np.float64),
        "virial": np.array(d["virial"], dtype=np.float64),
        "attributes": dict(d["attributes"]),
    }


def encode_rng(seed: int, step: int) -> bytes:
    out = bytearray([0x81])
    _str(out, "internal_state")
    _str(out, f"philox4x32-10 {int(seed)} {int(step)}")
    return bytes(out)


def decode_rng(data: bytes):
    import msgpack
    kind, seed, step = msgpack.unpackb(data, raw=False)["internal_state"].split()
    if kind != "philox4x32-10":
        raise ValueError(f"not a counter-based noise state: {kind}")
    return int(seed), int(step)


def save_engine(engine, prefix: str):
    """Write ``<prefix>_system.msg`` and ``<prefix>_rng.msg`` from an FP64 engine's state
    (MolecularDynamicsSimulator's checkpoint, MolecularDynamicsSimulator.hpp:97-108)."""
    spec = engine.spec
    x, v, f = engine.state()
    periodic = spec.boundary == "periodic"
    with open(prefix + "_system.msg", "wb") as fh:
        fh.write(encode_system(spec.mass, x, v, f, spec.lower if periodic else None, spec.upper if periodic else None,
                               spec.attributes, names=spec.particle_names))
    steps, _ = engine.counters()
    with open(prefix + "_rng.msg", "wb") as fh:
        fh.write(encode_rng(spec.integrator.seed, steps))


def load_engine(spec, prefix: str, precision: int = 0, device: int = 0):
    """An FP64 engine for ``spec`` (interactions, integrator) continuing from a checkpoint: state
    from ``<prefix>_system.msg``, noise position from ``<prefix>_rng.msg``
    (read_system.hpp:107-124 loads the system the same way)."""
    import copy
    from . import capi
    with open(prefix + "_system.msg", "rb") as fh:
        s = decode_system(fh.read())
    with open(prefix + "_rng.msg", "rb") as fh:
        seed, step = decode_rng(fh.read())
    sp = copy.deepcopy(spec)
    sp.mass, sp.pos, sp.vel = s["mass"], s["pos"], s["vel"]
    if s["lower"] is not None:
        sp.lower, sp.upper = s["lower"], s["upper"]
    sp.attributes.update(s["attributes"])
    sp.integrator.seed = seed
    e = capi.Engine(sp, precision, device=device)
    e.set_step_count(step)
    return e


def save_batch_replica(batch, r: int, prefix: str):
    """The same two files for one replica of a batch (its state, its step count)."""
    spec = batch.specs[r]
    x, v, f = batch.state(r)
    periodic = spec.boundary == "periodic"
    with open(prefix + "_system.msg", "wb") as fh:
        fh.write(encode_system(spec.mass, x, v, f, spec.lower if periodic else None, spec.upper if periodic else None,
                               spec.attributes, names=spec.particle_names))
    with open(prefix + "_rng.msg", "wb") as fh:
        fh.write(encode_rng(spec.integrator.seed, batch.info(r)["step"]))


def load_batch_replica(batch, spec, prefix: str) -> int:
    """Add a replica for ``spec`` (interactions, integrator) to ``batch`` from a checkpoint, make it
    resident and continue its noise stream; returns the replica id."""
    import copy
    with open(prefix + "_system.msg", "rb") as fh:
        s = decode_system(fh.read())
    with open(prefix + "_rng.msg", "rb") as fh:
        seed, step = decode_rng(fh.read())
    sp = copy.deepcopy(spec)
    sp.mass, sp.pos, sp.vel = s["mass"], s["pos"], s["vel"]
    if s["lower"] is not None:
        sp.lower, sp.upper = s["lower"], s["upper"]
    sp.attributes.update(s["attributes"])
    sp.integrator.seed = seed
    r = batch.add(sp)
    batch.initialize()
    batch.set_step_count(r, step)
    return r
