"""ctypes wrapper of the CPU restatement (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only the tests, the smoke entry point and the CPU-baseline leg of the benchmark use this
module, and only as the checker.  The product path (``mjolnir_amd``) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

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

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run `make -C oracle oracle` (or __graft_entry__.build())")
        L = C.CDLL(path)
        L.orc_kB.restype = C.c_double
        L.orc_eps0.restype = C.c_double
        L.orc_L_to_volume.restype = C.c_double
        L.orc_philox4x32_10.argtypes = [_up, _up, _up]
        L.orc_noise3.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, _dp]
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, _dp, _dp, C.c_size_t]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_state.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_get_state.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_set_attribute.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
        L.orc_add_local.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_size_t, _up, _dp, C.c_double]
        L.orc_add_global.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int,
                                     C.c_size_t, C.POINTER(C.c_char_p), _up, C.c_size_t, _up, _dp]
        L.orc_initialize.argtypes = [C.c_void_p]
        L.orc_add_constraint.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t, _up, _dp, C.c_size_t, C.c_double]
        L.orc_num_interactions.restype = C.c_size_t
        L.orc_num_interactions.argtypes = [C.c_void_p]
        L.orc_calc_energy_of.restype = C.c_double
        L.orc_calc_energy_of.argtypes = [C.c_void_p, C.c_int]
        L.orc_calc_force_of.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_calc_force_and_energy.restype = C.c_double
        L.orc_calc_force_and_energy.argtypes = [C.c_void_p]
        L.orc_pairs_of.restype = C.c_size_t
        L.orc_pairs_of.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]
        L.orc_reduce_margin.argtypes = [C.c_void_p, C.c_double]
        L.orc_set_integrator.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp, C.c_uint64]
        L.orc_integrator_initialize.argtypes = [C.c_void_p]
        L.orc_set_remover.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.orc_step.argtypes = [C.c_void_p, C.c_size_t]
        L.orc_step_count.restype = C.c_uint64
        L.orc_step_count.argtypes = [C.c_void_p]
        L.orc_rebuild_count.restype = C.c_size_t
        L.orc_rebuild_count.argtypes = [C.c_void_p]
        _LIB = L
    return _LIB


def kB() -> float:
    """Boltzmann's constant in kcal/mol/K after read_units (read_units.hpp:32-92)."""
    return float(lib().orc_kB())


def philox4x32_10(ctr, key):
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(np.asarray(ctr, dtype=np.uint32), np.asarray(key, dtype=np.uint32), out)
    return out


def noise3(seed: int, step: int, particle: int):
    out = np.zeros(3)
    lib().orc_noise3(seed, step, particle, out)
    return out


class OracleEngine:
    """The restated reference engine, built from a ``mjolnir_amd.spec.SystemSpec``."""

    def __init__(self, spec, with_integrator: bool = True):
        L = lib()
        self.spec = spec
        self.n = spec.n
        self.h = L.orc_create(1 if spec.boundary == "periodic" else 0,
                              np.ascontiguousarray(spec.lower), np.ascontiguousarray(spec.upper), spec.n)
        L.orc_set_state(self.h, spec.mass, spec.pos, spec.vel)
        for k, v in spec.attributes.items():
            L.orc_set_attribute(self.h, k.encode(), float(v))
        for l in spec.local:
            rc = L.orc_add_local(self.h, LOCAL_CODE[(l.interaction, l.potential)], l.topology.encode(),
                                 len(l.indices), np.ascontiguousarray(l.indices.ravel()),
                                 np.ascontiguousarray(l.params.ravel()), float(l.margin))
            if rc < 0:
                raise ValueError(f"oracle rejected local {l.name}")
        for g in spec.global_:
            kinds = list(g.ignore_within.keys())
            arr = (C.c_char_p * max(1, len(kinds)))(*[k.encode() for k in kinds])
            dists = np.array([g.ignore_within[k] for k in kinds] or [0], dtype=np.uint32)
            rc = L.orc_add_global(self.h, GLOBAL_CODE[g.potential], float(g.cutoff), float(g.epsilon),
                                  PARTITION_CODE[g.partition], float(g.margin), IGNORE_CODE[g.ignore_molecule],
                                  len(kinds), arr, dists, len(g.participants),
                                  np.ascontiguousarray(g.participants), np.ascontiguousarray(g.params.ravel()))
            if rc < 0:
                raise ValueError(f"oracle rejected global {g.name}")
        c = getattr(spec, "constraint", None)
        if c is not None:
            rc = L.orc_add_constraint(self.h, c.topology.encode(), len(c.v0), np.ascontiguousarray(c.indices.ravel()),
                                      c.v0, int(c.max_iteration), float(c.tolerance))
            if rc < 0:
                raise ValueError("oracle rejected the constraints")
        if L.orc_initialize(self.h) != 0:
            raise ValueError("oracle initialize failed")
        self.n_local = len(spec.local)
        if with_integrator and spec.integrator is not None:
            it = spec.integrator
            L.orc_set_integrator(self.h, INTEG_CODE[it.type], float(it.dt), it.coefs, int(it.seed))
            L.orc_set_remover(self.h, *[int(bool(x)) for x in getattr(it, "remove", (False, False, False))])
            if L.orc_integrator_initialize(self.h) != 0:
                raise ValueError("oracle integrator initialize failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def energy_of(self, k: int) -> float:
        return lib().orc_calc_energy_of(self.h, k)

    def force_of(self, k: int) -> np.ndarray:
        f = np.zeros(3 * self.n)
        lib().orc_calc_force_of(self.h, k, f)
        return f.reshape(-1, 3)

    def force_and_energy(self):
        e = lib().orc_calc_force_and_energy(self.h)
        return self.state()[2], e

    def pairs(self, gid: int) -> np.ndarray:
        L = lib()
        n = L.orc_pairs_of(self.h, gid, None, 0)
        out = np.zeros(2 * max(n, 1), dtype=np.uint32)
        L.orc_pairs_of(self.h, gid, out.ctypes.data, n)
        return out[: 2 * n].reshape(-1, 2)

    def step(self, n: int):
        if lib().orc_step(self.h, n) != 0:
            raise RuntimeError("oracle step failed")

    def state(self):
        x = np.zeros(3 * self.n); v = np.zeros(3 * self.n); f = np.zeros(3 * self.n)
        lib().orc_get_state(self.h, x, v, f)
        return x.reshape(-1, 3), v.reshape(-1, 3), f.reshape(-1, 3)

    @property
    def rebuilds(self) -> int:
        return lib().orc_rebuild_count(self.h)
