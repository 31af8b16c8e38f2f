"""Minimal SBML reader and writer for mass-action reaction networks.

Covers the subset of AMICI's SBML import (``SbmlImporter``,
python/sdist/amici/importers/sbml/__init__.py) that this backend compiles,
without libsbml:

* species with initial concentrations or amounts (``_process_species`` /
  ``_process_species_initial``, :1012-1100); an initial assignment
  ``species = parameter`` makes that parameter a fixed parameter (AMICI's
  ``fixed_parameters``, edata.h);
* reactions with reactant / product stoichiometries (``_process_reactions``,
  :1362-1428: S[species, reaction] += sign * stoichiometry) and mass-action
  kinetic laws ``k * prod(reactant ** stoichiometry)``, optionally multiplied by
  a compartment of size 1; a reversible law ``kf * prod(reactants) - kr *
  prod(products)`` becomes two reactions;
* global parameters and kinetic-law local parameters (renamed
  ``<reaction>_<parameter>``); every rate constant becomes a free parameter on
  the log10 scale;
* observables as linear combinations of species, given by the caller (as
  ``sbml2amici(observation_model=...)`` does) or taken from assignment rules
  whose variable starts with ``observable_`` (AMICI's legacy convention).

Anything outside the subset (events, rules other than observables, non
mass-action rate laws, compartments of size other than 1, boundary species)
raises :class:`SbmlError`.
"""
from __future__ import annotations

import math
import xml.etree.ElementTree as ET
from pathlib import Path
from typing import Mapping, Sequence

import numpy as np
import sympy as sp

from .networks import NOISE_NAMES, Reaction, ReactionNetwork

SBML_NS = "http://www.sbml.org/sbml/level3/version1/core"
MATHML_NS = "http://www.w3.org/1998/Math/MathML"


class SbmlError(ValueError):
    pass


def _local(tag: str) -> str:
    return tag.rsplit("}", 1)[-1]


def _children(el, name: str):
    return [c for c in el if _local(c.tag) == name]


def _child(el, name: str):
    for c in el:
        if _local(c.tag) == name:
            return c
    return None


def _list(model, name: str, item: str):
    lst = _child(model, name)
    return [] if lst is None else _children(lst, item)


def _mathml(el, rename: Mapping[str, str]) -> sp.Expr:
    """MathML content (apply / ci / cn with the arithmetic operators) -> sympy."""
    tag = _local(el.tag)
    if tag == "math":
        kids = list(el)
        if len(kids) != 1:
            raise SbmlError("<math> must hold one expression")
        return _mathml(kids[0], rename)
    if tag == "ci":
        name = el.text.strip()
        return sp.Symbol(rename.get(name, name))
    if tag == "cn":
        kind = el.get("type", "real")
        if kind == "e-notation":
            parts = [t.strip() for t in el.itertext() if t.strip()]
            return sp.Float(float(parts[0]) * 10.0 ** float(parts[1]))
        if kind == "rational":
            parts = [t.strip() for t in el.itertext() if t.strip()]
            return sp.Rational(int(parts[0]), int(parts[1]))
        text = el.text.strip()
        return sp.Integer(int(text)) if kind == "integer" else sp.Float(float(text))
    if tag == "apply":
        kids = list(el)
        op = _local(kids[0].tag)
        args = [_mathml(k, rename) for k in kids[1:]]
        if op == "times":
            return sp.Mul(*args)
        if op == "plus":
            return sp.Add(*args)
        if op == "minus":
            return -args[0] if len(args) == 1 else args[0] - args[1]
        if op == "divide":
            return args[0] / args[1]
        if op == "power":
            return args[0] ** args[1]
        raise SbmlError(f"MathML operator <{op}> is outside the supported subset")
    raise SbmlError(f"MathML element <{tag}> is outside the supported subset")


def _mass_action_term(term: sp.Expr, species: Sequence[str], params: set, unit: set):
    """Splits k * prod(x_i ** n_i) (times compartments of size 1) into
    (parameter name, {species: power}); None if the term has another form."""
    rate = None
    powers: dict = {}
    for f in sp.Mul.make_args(term):
        base, exp = f.as_base_exp()
        if base.is_Symbol and str(base) in unit and exp == 1:
            continue
        if base.is_Symbol and str(base) in params and exp == 1 and rate is None:
            rate = str(base)
            continue
        if base.is_Symbol and str(base) in species and exp.is_Integer and int(exp) > 0:
            powers[str(base)] = powers.get(str(base), 0) + int(exp)
            continue
        return None
    return None if rate is None else (rate, powers)


def read_sbml(source, observables: Mapping[str, str] | None = None,
              fixed_parameters: Sequence[str] = (), noise: str = "normal",
              name: str | None = None) -> ReactionNetwork:
    """Reads an SBML model (path or XML text) into a :class:`ReactionNetwork`.

    ``observables`` maps observable ids to linear expressions in the species
    ids; ``fixed_parameters`` names the parameters that are per-condition
    inputs (they may only appear as species initial assignments); ``noise`` is
    one of AMICI's continuous noise distributions (``"normal"``,
    ``"log-normal"``, ``"log10-normal"``, ``"laplace"``, ``"log-laplace"``,
    ``"log10-laplace"``; ``"lognormal"`` is accepted for ``"log-normal"``).
    """
    text = Path(source).read_text() if isinstance(source, Path) or (
        isinstance(source, str) and not source.lstrip().startswith("<")) else source
    root = ET.fromstring(text)
    model = _child(root, "model")
    if model is None:
        raise SbmlError("no <model> element")
    for unsupported in ("listOfEvents", "listOfConstraints", "listOfFunctionDefinitions"):
        lst = _child(model, unsupported)
        if lst is not None and len(list(lst)):
            raise SbmlError(f"{unsupported} is outside the supported subset")

    unit_compartments = set()
    for c in _list(model, "listOfCompartments", "compartment"):
        size = float(c.get("size", c.get("volume", "1")))
        if size != 1.0:
            raise SbmlError(f"compartment {c.get('id')} has size {size}; only size 1 is supported")
        unit_compartments.add(c.get("id"))

    params = {p.get("id"): float(p.get("value", "nan"))
              for p in _list(model, "listOfParameters", "parameter")}
    fixed = list(fixed_parameters)
    for f in fixed:
        if f not in params:
            raise SbmlError(f"fixed parameter {f} is not a model parameter")

    species, x0 = [], []
    for s in _list(model, "listOfSpecies", "species"):
        if s.get("boundaryCondition") == "true" or s.get("constant") == "true":
            raise SbmlError(f"species {s.get('id')}: boundary / constant species are not supported")
        species.append(s.get("id"))
        x0.append(float(s.get("initialConcentration", s.get("initialAmount", "0"))))
    sidx = {s: i for i, s in enumerate(species)}
    x0_fixed = np.full(len(species), -1, np.int32)
    for ia in _list(model, "listOfInitialAssignments", "initialAssignment"):
        sym = ia.get("symbol")
        expr = _mathml(_child(ia, "math"), {})
        if sym not in sidx:
            raise SbmlError(f"initial assignment to {sym}: only species targets are supported")
        if expr.is_Symbol and str(expr) in fixed:
            x0_fixed[sidx[sym]] = fixed.index(str(expr))
        elif expr.is_number:
            x0[sidx[sym]] = float(expr)
        elif expr.is_Symbol and str(expr) in params:
            x0[sidx[sym]] = params[str(expr)]
        else:
            raise SbmlError(f"initial assignment to {sym} is outside the supported subset")

    rules = {}
    for r in _list(model, "listOfRules", "assignmentRule"):
        rules[r.get("variable")] = _mathml(_child(r, "math"), {})
    other_rules = [r for r in (_child(model, "listOfRules") or []) if _local(r.tag) != "assignmentRule"]
    if other_rules or any(not v.startswith("observable_") for v in rules):
        raise SbmlError("rules other than observable_* assignment rules are not supported")

    rate_names: list = []
    reactions: list = []
    for rx in _list(model, "listOfReactions", "reaction"):
        rid = rx.get("id")

        def refs(listname):
            out = []
            for ref in _list(rx, listname, "speciesReference"):
                st = float(ref.get("stoichiometry", "1"))
                if st != int(st) or st < 0:
                    raise SbmlError(f"reaction {rid}: non-integer stoichiometry")
                out += [ref.get("species")] * int(st)
            return out

        reac, prod = refs("listOfReactants"), refs("listOfProducts")
        kl = _child(rx, "kineticLaw")
        if kl is None:
            raise SbmlError(f"reaction {rid} has no kinetic law")
        rename, local_params = {}, {}
        for lname in ("listOfLocalParameters", "listOfParameters"):
            for lp in _list(kl, lname, "localParameter" if lname == "listOfLocalParameters" else "parameter"):
                gname = f"{rid}_{lp.get('id')}"
                rename[lp.get("id")] = gname
                local_params[gname] = float(lp.get("value", "nan"))
        law = sp.expand(_mathml(_child(kl, "math"), rename))
        all_params = set(params) | set(local_params)
        terms = sp.Add.make_args(law)

        def as_reaction(term, left, right, sign):
            got = _mass_action_term(sign * term, species, all_params, unit_compartments)
            need = {s: left.count(s) for s in set(left)}
            if got is None or got[1] != need:
                return None
            return got[0]

        parts = []
        if len(terms) == 1:
            k = as_reaction(terms[0], reac, prod, 1)
            if k is None:
                raise SbmlError(f"reaction {rid}: kinetic law is not mass action in its reactants")
            parts.append((reac, prod, k))
        elif len(terms) == 2:
            for fwd, bwd in ((terms[0], terms[1]), (terms[1], terms[0])):
                kf = as_reaction(fwd, reac, prod, 1)
                kr = as_reaction(bwd, prod, reac, -1)
                if kf is not None and kr is not None:
                    parts = [(reac, prod, kf), (prod, reac, kr)]
                    break
            if not parts:
                raise SbmlError(f"reaction {rid}: kinetic law is not reversible mass action")
        else:
            raise SbmlError(f"reaction {rid}: kinetic law is not mass action")
        for left, right, k in parts:
            if len(left) > 2:
                raise SbmlError(f"reaction {rid}: more than two reactant molecules")
            if k in fixed:
                raise SbmlError(f"reaction {rid}: rate constant {k} is declared fixed")
            if k not in rate_names:
                rate_names.append(k)
            reactions.append(Reaction(tuple(sorted(sidx[s] for s in left)),
                                      tuple(sorted(sidx[s] for s in right)), rate_names.index(k)))
        params.update(local_params)

    values = []
    for k in rate_names:
        v = params[k]
        if not (v > 0 and math.isfinite(v)):
            raise SbmlError(f"rate constant {k} = {v} must be positive (log10 scale)")
        values.append(math.log10(v))

    if observables is None:
        observables = {v[len("observable_"):]: e for v, e in rules.items()}
    C = np.zeros((len(observables), len(species)))
    syms = {s: sp.Symbol(s) for s in species}
    for row, (oid, expr) in enumerate(observables.items()):
        e = sp.expand(expr if isinstance(expr, sp.Basic) else sp.sympify(expr, locals=syms))
        for s in species:
            d = sp.diff(e, syms[s])
            if not d.is_number:
                raise SbmlError(f"observable {oid} is not linear in the species")
            C[row, sidx[s]] = float(d)
        rest = sp.expand(e - sum(C[row, sidx[s]] * syms[s] for s in species))
        if rest != 0:
            raise SbmlError(f"observable {oid} is not linear in the species")
    net = ReactionNetwork(
        name=name or model.get("id") or "sbml_model",
        species=species, reactions=reactions, parameter_names=rate_names,
        theta_nominal=np.asarray(values), fixed_parameter_names=fixed,
        k_nominal=np.asarray([params[f] for f in fixed]),
        x0_const=np.where(x0_fixed >= 0, 0.0, np.asarray(x0)), x0_fixed=x0_fixed,
        observable_matrix=C,
        noise=_noise_id(noise))
    net.validate()
    return net


def _noise_id(noise: str) -> int:
    if noise not in NOISE_NAMES:
        raise SbmlError(f"noise distribution {noise!r} is not supported "
                        f"(one of {sorted(NOISE_NAMES)})")
    return NOISE_NAMES[noise]


def _ci(name):
    return f"<ci> {name} </ci>"


def write_sbml(net: ReactionNetwork) -> str:
    """SBML Level 3 Version 1 text of a network (mass-action kinetic laws,
    inputs as initial assignments from fixed parameters, observables as
    observable_* assignment rules)."""
    L = [f'<?xml version="1.0" encoding="UTF-8"?>',
         f'<sbml xmlns="{SBML_NS}" level="3" version="1">',
         f'  <model id="{net.name}">',
         '    <listOfCompartments>',
         '      <compartment id="cell" size="1" constant="true" spatialDimensions="3"/>',
         '    </listOfCompartments>',
         '    <listOfSpecies>']
    for i, s in enumerate(net.species):
        L.append(f'      <species id="{s}" compartment="cell" initialConcentration="{float(net.x0_const[i])!r}" '
                 'hasOnlySubstanceUnits="false" boundaryCondition="false" constant="false"/>')
    L.append('    </listOfSpecies>')
    L.append('    <listOfParameters>')
    for j, p in enumerate(net.parameter_names):
        L.append(f'      <parameter id="{p}" value="{float(10.0 ** net.theta_nominal[j])!r}" constant="true"/>')
    for j, p in enumerate(net.fixed_parameter_names):
        L.append(f'      <parameter id="{p}" value="{float(net.k_nominal[j])!r}" constant="true"/>')
    L.append('    </listOfParameters>')
    if (net.x0_fixed >= 0).any():
        L.append('    <listOfInitialAssignments>')
        for i, s in enumerate(net.species):
            if net.x0_fixed[i] >= 0:
                L.append(f'      <initialAssignment symbol="{s}"><math xmlns="{MATHML_NS}">'
                         f'{_ci(net.fixed_parameter_names[net.x0_fixed[i]])}</math></initialAssignment>')
        L.append('    </listOfInitialAssignments>')
    L.append('    <listOfRules>')
    for k in range(net.ny):
        terms = "".join(f"<apply><times/><cn> {float(net.observable_matrix[k, i])!r} </cn>{_ci(s)}</apply>"
                        for i, s in enumerate(net.species) if net.observable_matrix[k, i] != 0)
        L.append(f'      <assignmentRule variable="observable_y{k}"><math xmlns="{MATHML_NS}">'
                 f'<apply><plus/>{terms}<cn> 0 </cn></apply></math></assignmentRule>')
    L.append('    </listOfRules>')
    L.append('    <listOfReactions>')
    for r, rx in enumerate(net.reactions):
        def refs(idx):
            out = []
            for i in sorted(set(idx)):
                out.append(f'<speciesReference species="{net.species[i]}" '
                           f'stoichiometry="{idx.count(i)}" constant="true"/>')
            return "".join(out)
        factors = _ci(net.parameter_names[rx.rate]) + "".join(_ci(net.species[i]) for i in rx.reactants)
        lists = ""
        if rx.reactants:
            lists += f'<listOfReactants>{refs(list(rx.reactants))}</listOfReactants>'
        if rx.products:
            lists += f'<listOfProducts>{refs(list(rx.products))}</listOfProducts>'
        L.append(f'      <reaction id="r{r}" reversible="false">{lists}'
                 f'<kineticLaw><math xmlns="{MATHML_NS}"><apply><times/>{factors}</apply></math></kineticLaw>'
                 '</reaction>')
    L.append('    </listOfReactions>')
    L.append('  </model>')
    L.append('</sbml>')
    return "\n".join(L) + "\n"
