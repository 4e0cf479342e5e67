"""TOML front end: the reference's input files (read_input_file.hpp:147-193) for the subset this
engine runs, turned into a ``SystemSpec`` and run on the FP64 engine with the `.ene` energy
observer.

Supported:

* ``[files]`` ``output.prefix`` / ``output.path``, ``input.path``; ``include`` in ``[simulator]``,
  ``[[systems]]`` and ``[[forcefields]]`` (utility.hpp:27-290: the included file's keys merge into
  the table, paths relative to ``input.path``);
* ``[units]`` length "angstrom" / "Å", energy "kcal/mol" (read_units.hpp:32-104; other units are
  rejected, not converted);
* ``[simulator]`` type "MolecularDynamics", boundary_type "Unlimited" / "Periodic" /
  "PeriodicCuboid", delta_t, total_step, save_step, seed, integrator.type "BAOABLangevin" /
  "g-BAOABLangevin" / "UnderdampedLangevin" (gammas or parameters) and "G-JFLangevin" (alphas or
  parameters),
  integrator.remove.{translation, rotation, rescale} (read_simulator.hpp, read_integrator.hpp);
* ``[[systems]]``: attributes, boundary_shape.{lower, upper}, particles with mass/m,
  position/pos, velocity/vel (read_system.hpp:100-299);
* ``[[forcefields.local]]``: BondLength (Harmonic, GoContact, 3SPN2Bond), BondAngle (Harmonic,
  Gaussian), DihedralAngle (ClementiDihedral, Gaussian), Contact (GoContact), 3SPN2BaseStacking
  (read_local_interaction.hpp);
* ``[[forcefields.global]]``: Pair (LennardJones, ExcludedVolume, DebyeHuckel,
  3SPN2ExcludedVolume), 3SPN2BasePair, 3SPN2CrossStacking with ignore.molecule,
  ignore.particles_within, spatial_partition.{type, margin}, cutoff (read_global_interaction.hpp);
* ``[[forcefields.constraint]]`` (read_constraint.hpp:12-46).

``env`` tables substitute string parameter values and ``offset`` shifts indices, as the
reference's find_parameter does (utility.hpp).  A particle without a velocity gets one drawn from
the Maxwell-Boltzmann distribution at the system temperature with the simulator's seed; the
reference draws these from its own generator, so such a start is not the reference's.
"""
from __future__ import annotations

import argparse
import copy
import os
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional

import numpy as np

from .spec import NIL, ConstraintSpec, GlobalSpec, IntegratorSpec, LocalSpec, SystemSpec

_BASE = {"A": 0, "T": 1, "G": 2, "C": 3}
_BEAD = {"P": 0, "S": 1, "A": 2, "T": 3, "G": 4, "C": 5}


@dataclass
class SimulatorInput:
    spec: SystemSpec
    total_step: int
    save_step: int
    output_prefix: str           # path + prefix of the output files
    output_formats: List[str] = field(default_factory=list)   # trajectory observers: "xyz", "dcd"


def _load_toml(path: str) -> Dict[str, Any]:
    import tomli
    with open(path, "rb") as f:
        return tomli.load(f)


def _merge_include(table: Dict[str, Any], input_path: str, main_dir: str) -> Dict[str, Any]:
    """Merge the keys of ``include = "file"`` (or a list of files) into ``table``.  The file is
    ``input.path + name`` relative to the working directory, as the reference opens it
    (utility.hpp:286-307); a relative path that does not exist there is tried relative to the
    main input file's directory."""
    if "include" not in table:
        return table
    out = {k: v for k, v in table.items() if k != "include"}
    inc = table["include"]
    for name in ([inc] if isinstance(inc, str) else inc):
        fname = os.path.join(input_path, name)
        if not os.path.isabs(fname) and not os.path.exists(fname):
            fname = os.path.join(main_dir, fname)
        sub = _load_toml(fname)
        for k, v in sub.items():
            if k in out:
                raise ValueError(f"key {k!r} defined both in a table and in its include {name}")
            out[k] = v
    return out


def _param(item: Dict[str, Any], env: Dict[str, Any], *names, default=None):
    """find_parameter: the first of ``names`` present; a string value is looked up in env."""
    for n in names:
        if n in item:
            v = item[n]
            if isinstance(v, str) and n not in ("Base", "kind", "name", "group"):
                if v not in env:
                    raise KeyError(f"parameter {n} = {v!r} is not defined in env")
                v = env[v]
            return v
    if default is not None:
        return default
    raise KeyError(f"none of {names} in {item}")


def _per_particle(items: List[Dict[str, Any]], env: Dict[str, Any], n: int, *names) -> np.ndarray:
    out = np.zeros(n)
    seen = set()
    for it in items:
        idx = int(_param(it, env, "index")) + int(_param(it, env, "offset", default=0))
        if idx in seen:
            raise ValueError(f"index {idx} appears twice")
        seen.add(idx)
        out[idx] = float(_param(it, env, *names))
    return out


# ------------------------------------------------------------------------------------------
# systems
# ------------------------------------------------------------------------------------------
def _read_system(sys_t: Dict[str, Any], boundary: str, seed: int):
    parts = sys_t["particles"]
    n = len(parts)
    mass = np.array([float(p["mass"] if "mass" in p else p["m"]) for p in parts])
    pos = np.array([p["position"] if "position" in p else p["pos"] for p in parts], dtype=np.float64)
    attrs = {k: float(v) for k, v in sys_t.get("attributes", {}).items()}
    vel = np.zeros((n, 3))
    missing = []
    for i, p in enumerate(parts):
        v = p.get("velocity", p.get("vel"))
        if v is None:
            missing.append(i)
        else:
            vel[i] = v
    if missing:
        from .workloads import maxwell_boltzmann
        if "temperature" not in attrs:
            raise ValueError("particles without velocity need attributes.temperature")
        mb = maxwell_boltzmann(np.random.default_rng(seed), mass, attrs["temperature"])
        vel[missing] = mb[missing]
    lower = upper = np.zeros(3)
    if boundary == "periodic":
        b = sys_t["boundary_shape"]
        lower, upper = np.array(b["lower"], dtype=np.float64), np.array(b["upper"], dtype=np.float64)
    # name defaults to "X" (read_system.hpp:277)
    names = [str(p.get("name", "X")) for p in parts] if any("name" in p for p in parts) else None
    return mass, pos, vel, attrs, lower, upper, names


# ------------------------------------------------------------------------------------------
# force fields
# ------------------------------------------------------------------------------------------
_LOCAL_KEYS = {
    ("BondLength", "Harmonic"): ("k", "v0"),
    ("BondLength", "GoContact"): ("k", "v0"),
    ("BondLength", "3SPN2Bond"): ("k", "v0"),
    ("Contact", "GoContact"): ("k", "v0"),
    ("BondAngle", "Harmonic"): ("k", "v0"),
    ("BondAngle", "Gaussian"): ("k", "sigma", "v0"),
    ("DihedralAngle", "ClementiDihedral"): ("k1", "k3", "v0"),
    ("DihedralAngle", "Gaussian"): ("k", "sigma", "v0"),
}


def _nucleotides(items, env):
    out = []
    for it in items:
        ofs = int(_param(it, env, "offset", default=0))
        out.append({"strand": int(_param(it, env, "strand")),
                    "P": int(_param(it, env, "P")) + ofs if "P" in it else None,
                    "S": int(_param(it, env, "S")) + ofs, "B": int(_param(it, env, "B")) + ofs,
                    "base": _BASE[it["Base"]]})
    return out


def _read_local(t: Dict[str, Any]) -> LocalSpec:
    inter, pot = t["interaction"], t["potential"]
    topo = t.get("topology", "none")
    env = t.get("env", {})
    items = t["parameters"]
    margin = float(t.get("margin", 0.5))
    if inter == "3SPN2BaseStacking":
        # read_local_interaction.hpp:395-508: a term for each consecutive pair of one strand
        if pot != "3SPN2":
            raise ValueError(f"3SPN2BaseStacking potential {pot!r} is not supported")
        nuc = _nucleotides(items, env)
        idx, par = [], []
        for a, b in zip(nuc[:-1], nuc[1:]):
            if a["strand"] != b["strand"]:
                continue
            idx.append([a["S"], a["B"], b["B"], NIL if a["P"] is None else a["P"], b["S"],
                        NIL if b["P"] is None else b["P"]])
            par.append([4 * a["base"] + b["base"]])
        return LocalSpec(inter, "3SPN2", topo, np.array(idx, dtype=np.uint32).reshape(-1, 6),
                         np.array(par, dtype=np.float64).reshape(-1, 1), margin=margin)
    keys = _LOCAL_KEYS.get((inter, pot))
    if keys is None:
        raise ValueError(f"local interaction {inter}:{pot} is not supported")
    idx, par = [], []
    for it in items:
        ofs = int(_param(it, env, "offset", default=0))
        idx.append([int(i) + ofs for i in _param(it, env, "indices")])
        par.append([float(_param(it, env, k)) for k in keys])
    return LocalSpec(inter, pot, topo, np.array(idx, dtype=np.uint32), np.array(par, dtype=np.float64), margin=margin)


def _read_global(t: Dict[str, Any], n: int) -> GlobalSpec:
    inter, pot = t["interaction"], t["potential"]
    env = t.get("env", {})
    items = t["parameters"]
    ign = t.get("ignore", {})
    within = {k: int(v) for k, v in ign.get("particles_within", {}).items()}
    mol = ign.get("molecule", "Nothing")
    sp = t.get("spatial_partition", {"type": "Naive", "margin": 0.0})
    part_type, margin = sp["type"], float(sp.get("margin", 0.0))
    if inter in ("3SPN2BasePair", "3SPN2CrossStacking"):
        # read_global_interaction.hpp:230-420: per base {base, sugar}, cross stacking also the
        # neighbouring bases of the strand
        nuc = _nucleotides(items, env)
        part, par = [], []
        for k, b in enumerate(nuc):
            row = [b["base"], b["S"]]
            if inter == "3SPN2CrossStacking":
                b5 = nuc[k - 1]["B"] if k > 0 and nuc[k - 1]["strand"] == b["strand"] else -1
                b3 = nuc[k + 1]["B"] if k + 1 < len(nuc) and nuc[k + 1]["strand"] == b["strand"] else -1
                row += [b5, b3]
            part.append(b["B"])
            par.append(row)
        return GlobalSpec(potential=inter, participants=np.array(part), params=np.array(par, dtype=np.float64),
                          partition=part_type, margin=margin, ignore_molecule=mol, ignore_within=within)
    if inter != "Pair":
        raise ValueError(f"global interaction {inter} is not supported")
    part = []
    for it in items:
        part.append(int(_param(it, env, "index")) + int(_param(it, env, "offset", default=0)))
    if pot == "LennardJones":
        par = [[float(_param(it, env, "sigma", "σ")), float(_param(it, env, "epsilon", "ε"))] for it in items]
    elif pot == "ExcludedVolume":
        par = [[float(_param(it, env, "radius"))] for it in items]
    elif pot == "DebyeHuckel":
        par = [[float(_param(it, env, "charge"))] for it in items]
    elif pot == "3SPN2ExcludedVolume":
        par = [[float(_BEAD[it["kind"]])] for it in items]
    else:
        raise ValueError(f"pair potential {pot} is not supported")
    g = GlobalSpec(potential=pot, participants=np.array(part), params=np.array(par, dtype=np.float64),
                   partition=part_type, margin=margin, ignore_molecule=mol, ignore_within=within)
    if "cutoff" in t:
        g.cutoff = float(t["cutoff"])
    if pot == "ExcludedVolume":
        g.epsilon = float(t["epsilon"])
    return g


def _read_constraint(t: Dict[str, Any]) -> ConstraintSpec:
    env = t.get("env", {})
    idx, v0 = [], []
    for it in t["parameters"]:
        ofs = int(_param(it, env, "offset", default=0))
        idx.append([int(i) + ofs for i in _param(it, env, "indices")])
        v0.append(float(_param(it, env, "v0")))
    return ConstraintSpec(indices=np.array(idx), v0=np.array(v0), topology=t["topology"],
                          max_iteration=int(t["max_iteration"]), tolerance=float(t["tolerance"]))


# ------------------------------------------------------------------------------------------
# the input file
# ------------------------------------------------------------------------------------------
def load(path: str) -> SimulatorInput:
    root = _load_toml(path)
    files = root.get("files", {})
    input_path = files.get("input", {}).get("path", "./")
    main_dir = os.path.dirname(os.path.abspath(path))
    units = root.get("units", {})
    if units.get("length", "angstrom") not in ("angstrom", "Å") or units.get("energy", "kcal/mol") != "kcal/mol":
        raise ValueError("only [units] length = \"angstrom\", energy = \"kcal/mol\" are supported")
    sim = _merge_include(root["simulator"], input_path, main_dir)
    if sim.get("type", "MolecularDynamics") != "MolecularDynamics":
        raise ValueError(f"simulator type {sim.get('type')!r} is not supported")
    btype = sim["boundary_type"]
    boundary = {"Unlimited": "unlimited", "Periodic": "periodic", "PeriodicCuboid": "periodic"}[btype]
    seed = int(sim.get("seed", sim.get("integrator", {}).get("seed", 0)))
    systems = root["systems"]
    if len(systems) != 1:
        raise ValueError("exactly one [[systems]] table is supported")
    mass, pos, vel, attrs, lower, upper, names = _read_system(_merge_include(systems[0], input_path, main_dir),
                                                              boundary, seed)
    n = len(mass)
    ffs = root["forcefields"]
    if len(ffs) != 1:
        raise ValueError("exactly one [[forcefields]] table is supported")
    ff = _merge_include(ffs[0], input_path, main_dir)
    local = [_read_local(t) for t in ff.get("local", [])]
    global_ = [_read_global(t, n) for t in ff.get("global", [])]
    constraints = ff.get("constraint", [])
    if len(constraints) > 1:
        raise ValueError("at most one [[forcefields.constraint]] table is supported")
    integ = sim["integrator"]
    ienv = integ.get("env", {})
    itype = integ["type"]
    name = {"BAOABLangevin": "BAOABLangevin", "g-BAOABLangevin": "gBAOABLangevin",
            "G-JFLangevin": "GJFNVTLangevin", "UnderdampedLangevin": "UnderdampedLangevin"}.get(itype)
    if name is None:
        raise ValueError(f"integrator {itype!r} is not supported")
    if name == "GJFNVTLangevin":
        coefs = _per_particle(integ.get("alphas", integ.get("parameters", [])), ienv, n, "alpha", "α")
    else:
        coefs = _per_particle(integ.get("gammas", integ.get("parameters", [])), ienv, n, "gamma", "γ")
    rm = integ.get("remove", {})
    it = IntegratorSpec(name, float(sim["delta_t"]), coefs, seed=seed,
                        remove=(bool(rm.get("translation", False)), bool(rm.get("rotation", False)),
                                bool(rm.get("rescale", False))))
    spec = SystemSpec(boundary=boundary, mass=mass, pos=pos, vel=vel, lower=lower, upper=upper, attributes=attrs,
                      local=local, global_=global_, integrator=it,
                      constraint=_read_constraint(constraints[0]) if constraints else None,
                      name=os.path.splitext(os.path.basename(path))[0], particle_names=names)
    out = files.get("output", {})
    prefix = os.path.join(out.get("path", "./"), out.get("prefix", spec.name))
    # output.format: one name or a list (read_observer.hpp:61-101); the energy file is always written
    fmt = out.get("format", [])
    formats = [fmt] if isinstance(fmt, str) else list(fmt)
    for f in formats:
        if f not in ("xyz", "dcd"):
            raise ValueError(f"output.format {f!r} is not supported (xyz, dcd)")
    return SimulatorInput(spec, int(sim["total_step"]), int(sim["save_step"]), prefix, formats)


# ------------------------------------------------------------------------------------------
# writing a spec as an input file
# ------------------------------------------------------------------------------------------
def _f(x: float) -> str:
    r = repr(float(x))
    return r if ("." in r or "e" in r or "n" in r) else r + ".0"


def _dna_nucleotides(spec: SystemSpec):
    """Nucleotide table (strand order, 5' to 3') of a spec with 3SPN2 terms: strands follow the
    cross-stacking neighbours, phosphates come from the base-stacking terms."""
    cs = next(g for g in spec.global_ if g.potential == "3SPN2CrossStacking")
    info = {int(b): (int(p[0]), int(p[1]), int(p[2]), int(p[3])) for b, p in zip(cs.participants, cs.params)}
    phos = {}
    for l in spec.local:
        if l.interaction == "3SPN2BaseStacking":
            for t in l.indices:
                if int(t[5]) != NIL:
                    phos[int(t[2])] = int(t[5])
    inv = {"A": 0, "T": 1, "G": 2, "C": 3}
    code = {v: k for k, v in inv.items()}
    out = []
    strand = 0
    for b in cs.participants:
        b = int(b)
        if info[b][2] != -1:
            continue
        k = b
        while k != -1:
            base, S, _, b3 = info[k]
            out.append({"strand": strand, "P": phos.get(k), "S": S, "B": k, "Base": code[base]})
            k = b3
        strand += 1
    return out


def _nuc_rows(nucs):
    rows = []
    for k, nu in enumerate(nucs):
        p = f"P = {nu['P']}, " if nu["P"] is not None else ""
        rows.append(f"{{strand = {nu['strand']}, nucleotide = {k}, {p}S = {nu['S']}, B = {nu['B']}, Base = \"{nu['Base']}\"}},")
    return rows


def dump(spec: SystemSpec, total_step: int, save_step: int, path: str, prefix: str = "out", formats=("xyz",)):
    """Write ``spec`` as one input file in the reference's schema (no includes)."""
    inv_bead = {v: k for k, v in _BEAD.items()}
    fmt = ", ".join(f'"{f}"' for f in formats)
    L = ["[files]", f'output.prefix = "{prefix}"', 'output.path = "./"', f"output.format = [{fmt}]", "",
         "[units]", 'length = "angstrom"', 'energy = "kcal/mol"', "", "[simulator]", 'type = "MolecularDynamics"',
         f'boundary_type = "{"PeriodicCuboid" if spec.boundary == "periodic" else "Unlimited"}"', 'precision = "double"',
         f"delta_t = {_f(spec.integrator.dt)}", f"total_step = {total_step}", f"save_step = {save_step}",
         f"seed = {int(spec.integrator.seed)}"]
    itype = {"BAOABLangevin": "BAOABLangevin", "gBAOABLangevin": "g-BAOABLangevin",
             "GJFNVTLangevin": "G-JFLangevin", "UnderdampedLangevin": "UnderdampedLangevin"}[spec.integrator.type]
    L.append(f'integrator.type = "{itype}"')
    if any(spec.integrator.remove):
        t, r, c = (str(bool(x)).lower() for x in spec.integrator.remove)
        L.append(f"integrator.remove = {{translation = {t}, rotation = {r}, rescale = {c}}}")
    key = "alpha" if itype == "G-JFLangevin" else "gamma"
    L.append(f"integrator.{key}s = [")
    L += [f"{{index = {i}, {key} = {_f(c)}}}," for i, c in enumerate(spec.integrator.coefs)]
    L += ["]", "", "[[systems]]"]
    for k, v in spec.attributes.items():
        L.append(f"attributes.{k} = {_f(v)}")
    if spec.boundary == "periodic":
        L.append("boundary_shape.lower = [" + ", ".join(_f(x) for x in spec.lower) + "]")
        L.append("boundary_shape.upper = [" + ", ".join(_f(x) for x in spec.upper) + "]")
    L.append("particles = [")
    for i, (m, x, v) in enumerate(zip(spec.mass, spec.pos, spec.vel)):
        nm = f', name = "{spec.particle_names[i]}"' if spec.particle_names is not None else ""
        L.append(f"{{mass = {_f(m)}, position = [{', '.join(_f(a) for a in x)}], "
                 f"velocity = [{', '.join(_f(a) for a in v)}]{nm}}},")
    L += ["]", "", "[[forcefields]]"]
    nucs = None
    for l in spec.local:
        L += ["[[forcefields.local]]", f'interaction = "{l.interaction}"', f'potential = "{l.potential}"',
              f'topology = "{l.topology}"']
        if (l.interaction, l.potential) == ("Contact", "GoContact"):
            L.append(f"margin = {_f(l.margin)}")
        L.append("parameters = [")
        if l.interaction == "3SPN2BaseStacking":
            nucs = nucs or _dna_nucleotides(spec)
            L += _nuc_rows(nucs)
        else:
            keys = _LOCAL_KEYS[(l.interaction, l.potential)]
            for idx, par in zip(l.indices, l.params):
                L.append(f"{{indices = [{', '.join(str(int(i)) for i in idx)}], "
                         + ", ".join(f"{k} = {_f(p)}" for k, p in zip(keys, par)) + "},")
        L.append("]")
    for g in spec.global_:
        inter = g.potential if g.potential in ("3SPN2BasePair", "3SPN2CrossStacking") else "Pair"
        pot = "3SPN2" if inter != "Pair" else g.potential
        L += ["[[forcefields.global]]", f'interaction = "{inter}"', f'potential = "{pot}"',
              f'ignore.molecule = "{g.ignore_molecule}"']
        for k, v in g.ignore_within.items():
            L.append(f"ignore.particles_within.{k} = {int(v)}")
        L.append(f'spatial_partition = {{type = "{g.partition}", margin = {_f(g.margin)}}}')
        if inter == "Pair" and g.potential != "3SPN2ExcludedVolume":
            L.append(f"cutoff = {_f(g.cutoff)}")
        if g.potential == "ExcludedVolume":
            L.append(f"epsilon = {_f(g.epsilon)}")
        L.append("parameters = [")
        if inter != "Pair":
            nucs = nucs or _dna_nucleotides(spec)
            L += _nuc_rows(nucs)
        else:
            for i, p in zip(g.participants, g.params):
                if g.potential == "LennardJones":
                    L.append(f"{{index = {int(i)}, sigma = {_f(p[0])}, epsilon = {_f(p[1])}}},")
                elif g.potential == "ExcludedVolume":
                    L.append(f"{{index = {int(i)}, radius = {_f(p[0])}}},")
                elif g.potential == "DebyeHuckel":
                    L.append(f"{{index = {int(i)}, charge = {_f(p[0])}}},")
                else:
                    L.append(f'{{index = {int(i)}, kind = "{inv_bead[int(p[0])]}"}},')
        L.append("]")
    if spec.constraint is not None:
        c = spec.constraint
        L += ["[[forcefields.constraint]]", f'topology = "{c.topology}"', f"max_iteration = {int(c.max_iteration)}",
              f"tolerance = {_f(c.tolerance)}", "parameters = ["]
        L += [f"{{indices = [{int(i)}, {int(j)}], v0 = {_f(v)}}}," for (i, j), v in zip(c.indices, c.v0)]
        L.append("]")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def run(path: str, total_step: Optional[int] = None) -> str:
    """Run an input file on the FP64 engine: MolecularDynamicsSimulator with the observers of
    output.format (prefix_{position,velocity}.xyz / .dcd) and the energy observer (prefix.ene);
    returns the path of the .ene file."""
    from . import capi
    inp = load(path)
    e = capi.Engine(inp.spec)
    os.makedirs(os.path.dirname(os.path.abspath(inp.output_prefix)), exist_ok=True)
    if "xyz" in inp.output_formats:
        e.open_xyz_observer(inp.output_prefix, inp.spec.particle_names)
    if "dcd" in inp.output_formats:
        e.open_dcd_observer(inp.output_prefix)
    e.open_energy_observer(inp.output_prefix)
    e.simulate(inp.total_step if total_step is None else total_step, inp.save_step)
    e.close()
    return inp.output_prefix + ".ene"


def main():
    ap = argparse.ArgumentParser(description="run a Mjolnir input file on the MI355X engine")
    ap.add_argument("input")
    ap.add_argument("--total-step", type=int, default=None, help="override [simulator] total_step")
    args = ap.parse_args()
    print(run(args.input, args.total_step))


if __name__ == "__main__":
    main()
