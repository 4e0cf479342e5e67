"""System/force-field/integrator description shared by every backend.

A ``SystemSpec`` holds what a Mjolnir input file describes for the force-and-step path:
the ``[[systems]]`` table (boundary, attributes, particles), the ``[[forcefields]]``
local and global interaction tables, and the ``[simulator]`` integrator table.  The
field names follow the reference's TOML keys:

* boundary / particles / attributes: ``mjolnir/input/read_system.hpp:100-299``
* local interactions (``interaction``, ``potential``, ``topology``, ``parameters``):
  ``mjolnir/input/read_local_interaction.hpp:31-715``
* global pair interactions (``potential``, ``cutoff``, ``epsilon``, ``ignore.molecule``,
  ``ignore.particles_within``, ``spatial_partition.{type,margin}``, ``parameters``):
  ``mjolnir/input/read_global_interaction.hpp``, ``read_global_potential.hpp:262-705``,
  ``read_spatial_partition.hpp:23-116``
* integrators (``delta_t``, ``integrator.gammas`` / ``integrator.alphas``, ``seed``):
  ``mjolnir/input/read_integrator.hpp:18-307``

The same spec is uploaded to the HIP engine through the C ABI (``include/mjg.h``), to
the CPU restatement in ``oracle/``, and serialised (``write_mjspec``) for the driver
that runs the reference's own classes.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

# Number of index columns and parameter columns of each local (interaction, potential).
# Parameter order follows the potential constructors:
#   HarmonicPotential(k, v0)                 mjolnir/forcefield/local/HarmonicPotential.hpp:18
#   GoContactPotential(k, v0)                mjolnir/forcefield/local/GoContactPotential.hpp:21
#   GaussianPotential(k, sigma, v0)          mjolnir/forcefield/local/GaussianPotential.hpp:20
#   ClementiDihedralPotential(k1, k3, v0)    mjolnir/forcefield/local/ClementiDihedralPotential.hpp:21
#   ThreeSPN2BondPotential(k, v0)            mjolnir/forcefield/3SPN2/ThreeSPN2BondPotential.hpp
LOCAL_KINDS = {
    ("BondLength", "Harmonic"): (2, 2),
    ("BondLength", "GoContact"): (2, 2),
    ("Contact", "GoContact"): (2, 2),
    ("BondAngle", "Harmonic"): (3, 2),
    ("BondAngle", "Gaussian"): (3, 3),
    ("DihedralAngle", "ClementiDihedral"): (4, 3),
    ("DihedralAngle", "Gaussian"): (4, 3),
    ("BondLength", "3SPN2Bond"): (2, 2),
    # 3SPN2BaseStacking: (S5, B5, B3, P5, S3, P3), absent phosphate NIL; {stack kind 4*b5 + b3}
    ("3SPN2BaseStacking", "3SPN2"): (6, 1),
}

NIL = 0xFFFFFFFF   # absent index (the 5' phosphate of a strand's first nucleotide)

# Per-participant parameter columns of each global pair potential.
#   LennardJones: {sigma, epsilon}           LennardJonesPotential.hpp:21-25
#   ExcludedVolume: {radius}                 ExcludedVolumePotential.hpp:32-35
#   DebyeHuckel: {charge}                    DebyeHuckelPotential.hpp:212
#   3SPN2ExcludedVolume: {bead kind P 0, S 1, A 2, T 3, G 4, C 5}   ThreeSPN2ExcludedVolumePotential.hpp
#   3SPN2BasePair: {base A 0, T 1, G 2, C 3, sugar index}        ThreeSPN2BasePairPotential.hpp:207-211
#   3SPN2CrossStacking: {base, sugar, 5' base, 3' base (-1: none)} ThreeSPN2CrossStackingPotential.hpp:203-207
GLOBAL_KINDS = {
    "LennardJones": 2,
    "ExcludedVolume": 1,
    "DebyeHuckel": 1,
    "3SPN2ExcludedVolume": 1,
    "3SPN2BasePair": 2,
    "3SPN2CrossStacking": 4,
}

# relative cutoffs; the 3SPN2 interactions carry absolute cutoffs in their parameter sets
DEFAULT_CUTOFF = {"LennardJones": 2.5, "ExcludedVolume": 2.0, "DebyeHuckel": 5.5, "3SPN2ExcludedVolume": 1.0,
                  "3SPN2BasePair": 1.0, "3SPN2CrossStacking": 1.0}


@dataclass
class LocalSpec:
    interaction: str            # "BondLength", "BondAngle", "DihedralAngle", "Contact"
    potential: str              # "Harmonic", "Gaussian", "ClementiDihedral", "GoContact"
    topology: str               # connection kind written to the topology ("bond", "contact", "none")
    indices: np.ndarray         # (n_terms, arity) uint32
    params: np.ndarray          # (n_terms, n_params) float64
    margin: float = 0.5         # only Contact uses it (ContactInteraction.hpp:50-53)

    def __post_init__(self):
        arity, npar = LOCAL_KINDS[(self.interaction, self.potential)]
        self.indices = np.ascontiguousarray(self.indices, dtype=np.uint32).reshape(-1, arity)
        self.params = np.ascontiguousarray(self.params, dtype=np.float64).reshape(-1, npar)
        assert len(self.indices) == len(self.params)

    @property
    def name(self) -> str:
        """Energy-column name as the reference prints it (e.g. ``BondLength:Harmonic``;
        the sqrt-free Go specialisation calls itself ``BondLengthGoContact``,
        BondLengthGoContactInteraction.hpp:102)."""
        if (self.interaction, self.potential) == ("BondLength", "GoContact"):
            return "BondLengthGoContact"
        # BondLengthInteraction.hpp:88 with ThreeSPN2BondPotential::name(), a Gaussian dihedral
        # is a PeriodicGaussianPotential (read_local_interaction.hpp:282-290),
        # ThreeSPN2BaseStackingInteraction.hpp:106
        if (self.interaction, self.potential) == ("BondLength", "3SPN2Bond"):
            return "BondLength:ThreeSPN2BondPotential"
        if (self.interaction, self.potential) == ("DihedralAngle", "Gaussian"):
            return "DihedralAngle:PeriodicGaussian"
        if self.interaction == "3SPN2BaseStacking":
            return "3SPN2BaseStacking"
        return f"{self.interaction}:{self.potential}"


@dataclass
class GlobalSpec:
    potential: str                      # "LennardJones", "ExcludedVolume", "DebyeHuckel"
    participants: np.ndarray            # (n,) uint32, in input order
    params: np.ndarray                  # (n, k) float64 per participant
    cutoff: Optional[float] = None      # relative cutoff (None -> reference default)
    epsilon: float = 0.0                # ExcludedVolume only
    partition: str = "CellList"         # "CellList", "VerletList", "Naive"
    margin: float = 0.5
    ignore_molecule: str = "Nothing"    # "Nothing", "Self"/"Intra", "Others"/"Inter"
    ignore_within: Dict[str, int] = field(default_factory=dict)

    def __post_init__(self):
        k = GLOBAL_KINDS[self.potential]
        self.participants = np.ascontiguousarray(self.participants, dtype=np.uint32).reshape(-1)
        self.params = np.ascontiguousarray(self.params, dtype=np.float64).reshape(-1, k)
        if self.cutoff is None:
            self.cutoff = DEFAULT_CUTOFF[self.potential]

    @property
    def name(self) -> str:
        # specialisations: GlobalPairLennardJonesInteraction.hpp:134,
        # GlobalPairExcludedVolumeInteraction.hpp:134; generic: GlobalPairInteraction.hpp:112
        if self.potential in ("LennardJones", "ExcludedVolume"):
            return "GlobalPair" + self.potential
        if self.potential in ("3SPN2BasePair", "3SPN2CrossStacking"):
            return self.potential       # ThreeSPN2BasePairInteraction.hpp:106, ...CrossStacking...:106
        return "Pair:" + self.potential


@dataclass
class ConstraintSpec:
    """Bond constraints for gBAOABLangevin (``[[forcefields.constraint]]``,
    read_constraint.hpp:12-46): pairs (i, j) held at distance v0, the connection kind they
    write into the topology, the Gauss-Seidel sweep cap and the tolerance."""
    indices: np.ndarray                 # (m, 2) uint32
    v0: np.ndarray                      # (m,)
    topology: str = "bond"
    max_iteration: int = 500
    tolerance: float = 1e-6

    def __post_init__(self):
        self.indices = np.ascontiguousarray(self.indices, dtype=np.uint32).reshape(-1, 2)
        self.v0 = np.ascontiguousarray(self.v0, dtype=np.float64).reshape(-1)


@dataclass
class IntegratorSpec:
    type: str                  # "BAOABLangevin", "GJFNVTLangevin", "gBAOABLangevin", "UnderdampedLangevin"
    dt: float
    coefs: np.ndarray          # gammas (BAOAB, g-BAOAB, Underdamped) or alphas (G-JF), one per particle
    seed: int = 12345
    # SystemMotionRemover flags remove.{translation, rotation, rescale} (read_integrator.hpp:18-35)
    remove: Tuple[bool, bool, bool] = (False, False, False)

    def __post_init__(self):
        self.coefs = np.ascontiguousarray(self.coefs, dtype=np.float64).reshape(-1)


@dataclass
class SystemSpec:
    boundary: str                          # "unlimited" or "periodic"
    mass: np.ndarray                       # (N,)
    pos: np.ndarray                        # (N, 3)
    vel: np.ndarray                        # (N, 3)
    lower: np.ndarray = field(default_factory=lambda: np.zeros(3))
    upper: np.ndarray = field(default_factory=lambda: np.zeros(3))
    attributes: Dict[str, float] = field(default_factory=dict)
    local: List[LocalSpec] = field(default_factory=list)
    global_: List[GlobalSpec] = field(default_factory=list)
    integrator: Optional[IntegratorSpec] = None
    constraint: Optional[ConstraintSpec] = None
    name: str = "system"
    particle_names: Optional[List[str]] = None   # System::name(i); None: "X" for every particle

    def __post_init__(self):
        self.mass = np.ascontiguousarray(self.mass, dtype=np.float64).reshape(-1)
        self.pos = np.ascontiguousarray(self.pos, dtype=np.float64).reshape(-1, 3)
        self.vel = np.ascontiguousarray(self.vel, dtype=np.float64).reshape(-1, 3)
        self.lower = np.asarray(self.lower, dtype=np.float64).reshape(3)
        self.upper = np.asarray(self.upper, dtype=np.float64).reshape(3)
        assert self.boundary in ("unlimited", "periodic")

    @property
    def n(self) -> int:
        return len(self.mass)

    def interaction_names(self) -> List[str]:
        return [l.name for l in self.local] + [g.name for g in self.global_]


def _fmt(x: float) -> str:
    return repr(float(x))


def write_mjspec(spec: SystemSpec, path: str) -> None:
    """Serialise ``spec`` into the plain-text format read by ``oracle/ref_driver.cpp``.

    Every float is written with ``repr`` (shortest round-trip), so the reference driver
    reconstructs bit-identical doubles.
    """
    out = ["mjspec 1"]
    if spec.boundary == "periodic":
        out.append("boundary periodic " + " ".join(_fmt(v) for v in list(spec.lower) + list(spec.upper)))
    else:
        out.append("boundary unlimited")
    for k, v in spec.attributes.items():
        out.append(f"attribute {k} {_fmt(v)}")
    out.append(f"particles {spec.n}")
    for i in range(spec.n):
        p, v = spec.pos[i], spec.vel[i]
        out.append(" ".join(_fmt(x) for x in (spec.mass[i], *p, *v)))
    for l in spec.local:
        out.append(f"local {l.interaction} {l.potential} {l.topology} {len(l.indices)} {_fmt(l.margin)}")
        for idx, par in zip(l.indices, l.params):
            out.append(" ".join(str(int(i)) for i in idx) + " " + " ".join(_fmt(p) for p in par))
    for g in spec.global_:
        within = ",".join(f"{k}:{v}" for k, v in g.ignore_within.items()) or "-"
        out.append(f"global {g.potential} {len(g.participants)} {_fmt(g.cutoff)} {_fmt(g.epsilon)} "
                   f"{g.partition} {_fmt(g.margin)} {g.ignore_molecule} {within}")
        for idx, par in zip(g.participants, g.params):
            out.append(str(int(idx)) + " " + " ".join(_fmt(p) for p in par))
    if spec.constraint is not None:
        c = spec.constraint
        out.append(f"constraint {c.topology} {len(c.v0)} {int(c.max_iteration)} {_fmt(c.tolerance)}")
        for (i, j), v0 in zip(c.indices, c.v0):
            out.append(f"{int(i)} {int(j)} {_fmt(v0)}")
    if spec.integrator is not None:
        it = spec.integrator
        out.append(f"integrator {it.type} {_fmt(it.dt)} {int(it.seed)}")
        out.append(" ".join(_fmt(c) for c in it.coefs))
        if any(it.remove):
            out.append("remover " + " ".join(str(int(bool(x))) for x in it.remove))
    out.append("end")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
