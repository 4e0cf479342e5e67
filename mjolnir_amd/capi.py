"""ctypes binding of libmjg.so (the C ABI in include/mjg.h) and a System/ForceField/Integrator
facade over it.

The facade mirrors the reference's calling sequence for the force-and-step path
(MolecularDynamicsSimulator::initialize, mjolnir/core/MolecularDynamicsSimulator.hpp:72-81):
System state -> ForceField::initialize -> Integrator::initialize -> Integrator::step.
There is no fallback: if the HIP library is missing or no GPU is present, construction
raises ``MjgError``.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional

import numpy as np

from .spec import SystemSpec

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmjg.so")

OK, INVALID_ARG, OOM, DEVICE_ERROR, NAN_DETECTED, NOT_INITIALIZED, UNSUPPORTED = range(7)
FP64, FP32_FIXED = 0, 1
UNLIMITED, PERIODIC = 0, 1

LOCAL_CODE = {
    ("BondLength", "Harmonic"): 1,
    ("BondLength", "GoContact"): 2,
    ("Contact", "GoContact"): 3,
    ("BondAngle", "Harmonic"): 4,
    ("BondAngle", "Gaussian"): 5,
    ("DihedralAngle", "ClementiDihedral"): 6,
    ("BondLength", "3SPN2Bond"): 7,
    ("DihedralAngle", "Gaussian"): 8,
    ("3SPN2BaseStacking", "3SPN2"): 9,
}
GLOBAL_CODE = {"LennardJones": 1, "ExcludedVolume": 2, "DebyeHuckel": 3, "3SPN2ExcludedVolume": 4,
               "3SPN2BasePair": 5, "3SPN2CrossStacking": 6}
PARTITION_CODE = {"CellList": 1, "VerletList": 2, "Naive": 3}
IGNORE_CODE = {"Nothing": 0, "Self": 1, "Intra": 1, "Others": 2, "Inter": 2}
INTEG_CODE = {"BAOABLangevin": 1, "GJFNVTLangevin": 2, "gBAOABLangevin": 3, "UnderdampedLangevin": 4}


class MjgError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"mjg status {status}: {msg}")
        self.status = status


_LIB = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")

EXPORTED = [
    "mjg_version", "mjg_engine_create", "mjg_engine_destroy", "mjg_last_error",
    "mjg_system_set_state", "mjg_system_set_positions", "mjg_system_get_state", "mjg_system_set_attribute",
    "mjg_add_local", "mjg_add_global", "mjg_add_constraint", "mjg_set_step_count", "mjg_forcefield_initialize", "mjg_num_interactions",
    "mjg_calc_force", "mjg_calc_force_and_energy", "mjg_calc_energies", "mjg_calc_force_and_virial",
    "mjg_interaction_name", "mjg_calc_force_of",
    "mjg_reduce_margin", "mjg_neighbor_pairs", "mjg_integrator_set", "mjg_integrator_set_remover", "mjg_observer_energy_open",
    "mjg_observer_energy_output", "mjg_observer_xyz_open", "mjg_observer_dcd_open", "mjg_simulate", "mjg_integrator_initialize",
    "mjg_step", "mjg_counters", "mjg_kinetic_energy", "mjg_kinetic_energy_static", "mjg_profile_interaction",
    "mjg_batch_create", "mjg_batch_destroy", "mjg_batch_last_error", "mjg_batch_add_replica",
    "mjg_batch_set_attribute", "mjg_batch_add_local", "mjg_batch_add_global", "mjg_batch_set_integrator", "mjg_batch_set_step_count",
    "mjg_batch_initialize", "mjg_batch_step", "mjg_batch_set_steps", "mjg_batch_advance", "mjg_batch_retire",
    "mjg_batch_get_state", "mjg_batch_set_autotune", "mjg_batch_set_tuning", "mjg_batch_replica_tuning",
    "mjg_batch_replica_status", "mjg_batch_replica_info", "mjg_batch_list_sizes", "mjg_batch_evaluate", "mjg_batch_counters",
    "mjg_batch_observe", "mjg_batch_energies",
    "mjg_batch_profile",
]


def lib():
    """Load libmjg.so (built in-tree by ``mjolnir_amd.build``); raises if it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise MjgError(NOT_INITIALIZED, f"{LIB_PATH} is missing: build it with `python -m mjolnir_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, sz, i32, u64, dbl = C.c_void_p, C.c_size_t, C.c_int, C.c_uint64, C.c_double
    L.mjg_version.restype = C.c_char_p
    L.mjg_engine_create.argtypes = [i32, i32, _dp, _dp, sz, i32, C.POINTER(vp)]
    L.mjg_engine_destroy.argtypes = [vp]
    L.mjg_last_error.restype = C.c_char_p
    L.mjg_last_error.argtypes = [vp]
    L.mjg_system_set_state.argtypes = [vp, _dp, _dp, _dp]
    L.mjg_system_get_state.argtypes = [vp, vp, vp, vp]
    L.mjg_system_set_attribute.argtypes = [vp, C.c_char_p, dbl]
    L.mjg_add_local.argtypes = [vp, i32, C.c_char_p, sz, _up, _dp, dbl, C.POINTER(i32)]
    L.mjg_add_constraint.argtypes = [vp, C.c_char_p, sz, _up, _dp, sz, dbl]
    L.mjg_set_step_count.argtypes = [vp, C.c_uint64]
    L.mjg_add_global.argtypes = [vp, i32, dbl, dbl, i32, dbl, i32, sz, C.POINTER(C.c_char_p), _up, sz, _up, _dp,
                                 C.POINTER(i32)]
    L.mjg_forcefield_initialize.argtypes = [vp]
    L.mjg_num_interactions.restype = sz
    L.mjg_num_interactions.argtypes = [vp]
    L.mjg_calc_force.argtypes = [vp]
    L.mjg_calc_force_and_energy.argtypes = [vp, C.POINTER(dbl)]
    L.mjg_calc_energies.argtypes = [vp, vp, C.POINTER(dbl)]
    L.mjg_calc_force_of.argtypes = [vp, i32, _dp]
    L.mjg_reduce_margin.argtypes = [vp, dbl]
    L.mjg_neighbor_pairs.argtypes = [vp, i32, vp, sz, C.POINTER(sz)]
    L.mjg_integrator_set.argtypes = [vp, i32, dbl, _dp, u64]
    L.mjg_integrator_set_remover.argtypes = [vp, i32, i32, i32]
    L.mjg_observer_energy_open.argtypes = [vp, C.c_char_p]
    L.mjg_observer_energy_output.argtypes = [vp, u64]
    L.mjg_observer_xyz_open.argtypes = [vp, C.c_char_p, C.POINTER(C.c_char_p)]
    L.mjg_observer_dcd_open.argtypes = [vp, C.c_char_p]
    L.mjg_simulate.argtypes = [vp, u64, u64]
    L.mjg_integrator_initialize.argtypes = [vp]
    L.mjg_step.argtypes = [vp, sz]
    L.mjg_counters.argtypes = [vp, C.POINTER(u64), C.POINTER(u64)]
    L.mjg_kinetic_energy.argtypes = [vp, C.POINTER(dbl)]
    L.mjg_kinetic_energy_static.argtypes = [vp, C.POINTER(dbl)]
    L.mjg_profile_interaction.argtypes = [vp, i32, i32, C.POINTER(dbl), C.POINTER(u64)]
    L.mjg_batch_create.argtypes = [i32, i32, C.POINTER(vp)]
    L.mjg_batch_destroy.argtypes = [vp]
    L.mjg_batch_last_error.restype = C.c_char_p
    L.mjg_batch_last_error.argtypes = [vp]
    L.mjg_batch_add_replica.argtypes = [vp, i32, _dp, _dp, sz, _dp, _dp, _dp, C.POINTER(i32)]
    L.mjg_batch_set_attribute.argtypes = [vp, i32, C.c_char_p, dbl]
    L.mjg_batch_add_local.argtypes = [vp, i32, i32, C.c_char_p, sz, _up, _dp]
    L.mjg_batch_add_global.argtypes = [vp, i32, i32, dbl, dbl, i32, dbl, i32, sz, C.POINTER(C.c_char_p), _up, sz,
                                       _up, _dp]
    L.mjg_batch_set_integrator.argtypes = [vp, i32, i32, dbl, _dp, u64]
    L.mjg_batch_set_step_count.argtypes = [vp, i32, u64]
    L.mjg_batch_initialize.argtypes = [vp]
    L.mjg_batch_step.argtypes = [vp, sz]
    L.mjg_batch_set_steps.argtypes = [vp, i32, u64]
    L.mjg_batch_advance.argtypes = [vp, i32]
    L.mjg_batch_retire.argtypes = [vp, i32]
    L.mjg_batch_get_state.argtypes = [vp, i32, vp, vp, vp]
    L.mjg_batch_replica_info.argtypes = [vp, i32, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64), C.POINTER(sz)]
    L.mjg_batch_replica_status.argtypes = [vp, i32, C.POINTER(i32)]
    L.mjg_batch_set_autotune.argtypes = [vp, i32]
    L.mjg_batch_set_tuning.argtypes = [vp, i32, dbl, dbl, i32, i32]
    L.mjg_batch_replica_tuning.argtypes = [vp, i32, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(i32), C.POINTER(i32)]
    L.mjg_batch_evaluate.argtypes = [vp, i32, vp, C.POINTER(dbl), vp]
    L.mjg_batch_observe.argtypes = [vp]
    L.mjg_batch_energies.argtypes = [vp, i32, vp, C.POINTER(dbl)]
    L.mjg_batch_counters.argtypes = [vp, C.POINTER(u64), C.POINTER(u64)]
    L.mjg_batch_list_sizes.argtypes = [vp, i32, vp, i32]
    L.mjg_batch_profile.argtypes = [vp, i32, i32, C.POINTER(dbl), C.POINTER(u64)]
    _LIB = L
    return L


class Engine:
    """One system resident on the GPU, built from a ``SystemSpec``.

    Interaction ids follow the reference's ``format_energy`` order: locals in insertion
    order, then globals (mjolnir/core/ForceField.hpp:154-161).
    """

    def __init__(self, spec: SystemSpec, precision: int = FP64, device: int = 0, with_integrator: bool = True):
        L = lib()
        self._pid = os.getpid()
        self.spec = spec
        self.n = spec.n
        h = C.c_void_p()
        st = L.mjg_engine_create(precision, PERIODIC if spec.boundary == "periodic" else UNLIMITED,
                                 np.ascontiguousarray(spec.lower), np.ascontiguousarray(spec.upper),
                                 spec.n, device, C.byref(h))
        if st != OK:
            raise MjgError(st, "mjg_engine_create failed (is a GPU present?)")
        self.h = h
        self._check(L.mjg_system_set_state(self.h, spec.mass, spec.pos, spec.vel))
        for k, v in spec.attributes.items():
            self._check(L.mjg_system_set_attribute(self.h, k.encode(), float(v)))
        iid = C.c_int()
        for l in spec.local:
            self._check(L.mjg_add_local(self.h, LOCAL_CODE[(l.interaction, l.potential)], l.topology.encode(),
                                        len(l.indices), np.ascontiguousarray(l.indices.ravel()),
                                        np.ascontiguousarray(l.params.ravel()), float(l.margin), C.byref(iid)))
        for g in spec.global_:
            kinds = list(g.ignore_within.keys())
            arr = (C.c_char_p * max(1, len(kinds)))(*[k.encode() for k in kinds])
            dists = np.array([g.ignore_within[k] for k in kinds] or [0], dtype=np.uint32)
            self._check(L.mjg_add_global(self.h, GLOBAL_CODE[g.potential], float(g.cutoff), float(g.epsilon),
                                         PARTITION_CODE[g.partition], float(g.margin),
                                         IGNORE_CODE[g.ignore_molecule], len(kinds), arr, dists,
                                         len(g.participants), np.ascontiguousarray(g.participants),
                                         np.ascontiguousarray(g.params.ravel()), C.byref(iid)))
        if spec.constraint is not None:
            c = spec.constraint
            self._check(L.mjg_add_constraint(self.h, c.topology.encode(), len(c.v0),
                                             np.ascontiguousarray(c.indices.ravel()), c.v0, int(c.max_iteration),
                                             float(c.tolerance)))
        self._check(L.mjg_forcefield_initialize(self.h))
        self.n_local = len(spec.local)
        if with_integrator and spec.integrator is not None:
            it = spec.integrator
            self._check(L.mjg_integrator_set(self.h, INTEG_CODE[it.type], float(it.dt), it.coefs, int(it.seed)))
            self._check(L.mjg_integrator_set_remover(self.h, *[int(bool(x)) for x in it.remove]))
            self._check(L.mjg_integrator_initialize(self.h))

    def _check(self, st: int):
        if st != OK:
            msg = lib().mjg_last_error(self.h)
            raise MjgError(st, msg.decode() if msg else "")

    def close(self):
        # a handle belongs to the process that created it: a forked child that collects the
        # object must not call into the HIP runtime with it
        if getattr(self, "h", None):
            if getattr(self, "_pid", os.getpid()) == os.getpid():
                lib().mjg_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- ForceFieldBase ------------------------------------------------------------------
    @property
    def num_interactions(self) -> int:
        return lib().mjg_num_interactions(self.h)

    def calc_force(self):
        self._check(lib().mjg_calc_force(self.h))

    def calc_force_and_energy(self) -> float:
        e = C.c_double()
        self._check(lib().mjg_calc_force_and_energy(self.h, C.byref(e)))
        return e.value

    def energies(self):
        per = np.zeros(max(1, self.num_interactions))
        tot = C.c_double()
        self._check(lib().mjg_calc_energies(self.h, per.ctypes.data, C.byref(tot)))
        return per[: self.num_interactions], tot.value

    def force_of(self, k: int) -> np.ndarray:
        out = np.zeros(3 * self.n)
        self._check(lib().mjg_calc_force_of(self.h, k, out))
        return out.reshape(-1, 3)

    def reduce_margin(self, dmargin: float):
        self._check(lib().mjg_reduce_margin(self.h, float(dmargin)))

    def pairs(self, k: int) -> np.ndarray:
        n = C.c_size_t()
        self._check(lib().mjg_neighbor_pairs(self.h, k, None, 0, C.byref(n)))
        out = np.zeros(2 * max(1, n.value), dtype=np.uint32)
        self._check(lib().mjg_neighbor_pairs(self.h, k, out.ctypes.data, n.value, C.byref(n)))
        return out[: 2 * n.value].reshape(-1, 2)

    # -- Integrator ----------------------------------------------------------------------
    def step(self, n: int):
        self._check(lib().mjg_step(self.h, int(n)))

    def counters(self):
        s, r = C.c_uint64(), C.c_uint64()
        self._check(lib().mjg_counters(self.h, C.byref(s), C.byref(r)))
        return s.value, r.value

    def set_step_count(self, step: int):
        """Continue the noise stream at `step` (checkpoint restore)."""
        self._check(lib().mjg_set_step_count(self.h, int(step)))

    def kinetic_energy(self) -> float:
        e = C.c_double()
        self._check(lib().mjg_kinetic_energy(self.h, C.byref(e)))
        return e.value

    def kinetic_energy_static(self) -> float:
        """mjg_kinetic_energy_static: the same sum in the library's static device scratch (no
        device allocation; at most 1048576 particles)."""
        e = C.c_double()
        self._check(lib().mjg_kinetic_energy_static(self.h, C.byref(e)))
        return e.value

    def profile_interaction(self, k: int, reps: int = 50):
        """(average kernel ms, half-list pairs) of interaction k's force kernel."""
        ms, hp = C.c_double(), C.c_uint64()
        self._check(lib().mjg_profile_interaction(self.h, k, reps, C.byref(ms), C.byref(hp)))
        return ms.value, hp.value

    # -- System --------------------------------------------------------------------------
    def open_energy_observer(self, prefix: str):
        self._check(lib().mjg_observer_energy_open(self.h, prefix.encode()))

    def output_energy(self, step: int):
        self._check(lib().mjg_observer_energy_output(self.h, int(step)))

    def open_xyz_observer(self, prefix: str, names=None):
        arr = None
        if names is not None:
            if len(names) != self.n:
                raise ValueError("one name per particle")
            arr = (C.c_char_p * self.n)(*[str(x).encode() for x in names])
        self._check(lib().mjg_observer_xyz_open(self.h, prefix.encode(), arr))

    def open_dcd_observer(self, prefix: str):
        self._check(lib().mjg_observer_dcd_open(self.h, prefix.encode()))

    def simulate(self, total_step: int, save_step: int):
        self._check(lib().mjg_simulate(self.h, int(total_step), int(save_step)))

    def state(self):
        x = np.zeros(3 * self.n); v = np.zeros(3 * self.n); f = np.zeros(3 * self.n)
        self._check(lib().mjg_system_get_state(self.h, x.ctypes.data, v.ctypes.data, f.ctypes.data))
        return x.reshape(-1, 3), v.reshape(-1, 3), f.reshape(-1, 3)


class Batch:
    """Independent replicas advanced together on one GPU (one persistent workgroup each).

    ``add(spec)`` uploads one ``SystemSpec`` as a replica and returns its id; ``initialize()``
    makes every new replica resident (lists, first forces); ``step(n)`` advances all of them.

    A replica whose state stops being finite is stopped by the library while the others go on:
    ``initialize``, ``step``, ``advance`` and ``profile`` then record its id in ``failed`` (and
    ``step`` / ``advance`` / ``initialize`` return the ids that failed during that call).
    """

    def __init__(self, precision: int = FP32_FIXED, device: int = 0):
        L = lib()
        self._pid = os.getpid()
        h = C.c_void_p()
        st = L.mjg_batch_create(precision, device, C.byref(h))
        if st != OK:
            raise MjgError(st, "mjg_batch_create failed (is a GPU present?)")
        self.h = h
        self.specs = {}
        self.failed = set()

    def _check(self, st: int):
        if st != OK:
            msg = lib().mjg_batch_last_error(self.h)
            raise MjgError(st, msg.decode() if msg else "")

    def _check_run(self, st: int):
        """Status of a call that advances replicas: MJG_NAN_DETECTED names replicas that were stopped
        (the others completed the call); returns their ids."""
        if st != NAN_DETECTED:
            self._check(st)
            return []
        new = [r for r in self.specs if r not in self.failed and self.status(r) != OK]
        self.failed.update(new)
        return new

    def set_autotune(self, on: bool):
        self._check(lib().mjg_batch_set_autotune(self.h, int(bool(on))))

    def set_tuning(self, r: int, skin_short: float, skin_dh: float, grid_bits: int, grid_lanes: int):
        """Fix replica r's launch parameters (before initialize()); they change no result."""
        self._check(lib().mjg_batch_set_tuning(self.h, r, float(skin_short), float(skin_dh), int(grid_bits),
                                               int(grid_lanes)))

    def tuning(self, r: int) -> dict:
        a, b, c, d = C.c_double(), C.c_double(), C.c_int(), C.c_int()
        self._check(lib().mjg_batch_replica_tuning(self.h, r, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return {"skin_short": a.value, "skin_dh": b.value, "grid_bits": c.value, "grid_lanes": d.value}

    def status(self, r: int) -> int:
        """OK, or NAN_DETECTED for a replica stopped with non-finite state."""
        out = C.c_int()
        self._check(lib().mjg_batch_replica_status(self.h, r, C.byref(out)))
        return out.value

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_pid", os.getpid()) == os.getpid():
                lib().mjg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, spec: SystemSpec) -> int:
        L = lib()
        rid = C.c_int()
        self._check(L.mjg_batch_add_replica(self.h, PERIODIC if spec.boundary == "periodic" else UNLIMITED,
                                            np.ascontiguousarray(spec.lower), np.ascontiguousarray(spec.upper),
                                            spec.n, spec.mass, spec.pos, spec.vel, C.byref(rid)))
        r = rid.value
        for k, v in spec.attributes.items():
            self._check(L.mjg_batch_set_attribute(self.h, r, k.encode(), float(v)))
        for l in spec.local:
            self._check(L.mjg_batch_add_local(self.h, r, LOCAL_CODE[(l.interaction, l.potential)], l.topology.encode(),
                                              len(l.indices), np.ascontiguousarray(l.indices.ravel()),
                                              np.ascontiguousarray(l.params.ravel())))
        for g in spec.global_:
            kinds = list(g.ignore_within.keys())
            arr = (C.c_char_p * max(1, len(kinds)))(*[k.encode() for k in kinds])
            dists = np.array([g.ignore_within[k] for k in kinds] or [0], dtype=np.uint32)
            self._check(L.mjg_batch_add_global(self.h, r, GLOBAL_CODE[g.potential], float(g.cutoff), float(g.epsilon),
                                               PARTITION_CODE[g.partition], float(g.margin),
                                               IGNORE_CODE[g.ignore_molecule], len(kinds), arr, dists,
                                               len(g.participants), np.ascontiguousarray(g.participants),
                                               np.ascontiguousarray(g.params.ravel())))
        it = spec.integrator
        self._check(L.mjg_batch_set_integrator(self.h, r, INTEG_CODE[it.type], float(it.dt), it.coefs, int(it.seed)))
        self.specs[r] = spec
        return r

    def initialize(self):
        return self._check_run(lib().mjg_batch_initialize(self.h))

    def step(self, n: int):
        return self._check_run(lib().mjg_batch_step(self.h, int(n)))

    def set_steps(self, r: int, n: int):
        self._check(lib().mjg_batch_set_steps(self.h, r, int(n)))

    def advance(self, segment: int):
        return self._check_run(lib().mjg_batch_advance(self.h, int(segment)))

    def retire(self, r: int):
        self._check(lib().mjg_batch_retire(self.h, r))
        self.specs.pop(r, None)
        self.failed.discard(r)

    def state(self, r: int):
        n = self.specs[r].n
        x = np.zeros(3 * n); v = np.zeros(3 * n); f = np.zeros(3 * n)
        self._check(lib().mjg_batch_get_state(self.h, r, x.ctypes.data, v.ctypes.data, f.ctypes.data))
        return x.reshape(-1, 3), v.reshape(-1, 3), f.reshape(-1, 3)

    def set_step_count(self, r: int, step: int):
        self._check(lib().mjg_batch_set_step_count(self.h, r, int(step)))

    def info(self, r: int):
        s, rem, rb, fp = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_size_t()
        self._check(lib().mjg_batch_replica_info(self.h, r, C.byref(s), C.byref(rem), C.byref(rb), C.byref(fp)))
        return {"step": s.value, "remaining": rem.value, "rebuilds": rb.value, "footprint": fp.value}

    def list_sizes(self, r: int):
        ng = len(self.specs[r].global_)
        out = np.zeros(max(1, ng), dtype=np.uint64)
        self._check(lib().mjg_batch_list_sizes(self.h, r, out.ctypes.data, ng))
        return out[:ng]

    def evaluate(self, r: int):
        spec = self.specs[r]
        ne = len(spec.local) + len(spec.global_)
        per = np.zeros(max(1, ne))
        tot = C.c_double()
        f = np.zeros(3 * spec.n)
        self._check(lib().mjg_batch_evaluate(self.h, r, per.ctypes.data, C.byref(tot), f.ctypes.data))
        return per[:ne], tot.value, f.reshape(-1, 3)

    def observe(self):
        """Energies of every resident replica at the current state (one launch)."""
        self._check(lib().mjg_batch_observe(self.h))

    def energies(self, r: int):
        """(per-interaction energies, kinetic energy) of replica r at the last observe()."""
        spec = self.specs[r]
        per = np.zeros(max(1, len(spec.local) + len(spec.global_)))
        ke = C.c_double()
        self._check(lib().mjg_batch_energies(self.h, r, per.ctypes.data, C.byref(ke)))
        return per[:len(spec.local) + len(spec.global_)], ke.value

    def counters(self):
        ps, rb = C.c_uint64(), C.c_uint64()
        self._check(lib().mjg_batch_counters(self.h, C.byref(ps), C.byref(rb)))
        return ps.value, rb.value

    def profile(self, segment_steps: int, reps: int = 3):
        ms, ps = C.c_double(), C.c_uint64()
        self._check_run(lib().mjg_batch_profile(self.h, segment_steps, reps, C.byref(ms), C.byref(ps)))
        return ms.value, ps.value
