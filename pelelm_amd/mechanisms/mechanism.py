"""Mechanism description: species, NASA-7 thermo and elementary reactions.

A mechanism file is a small YAML document (see ``drm19.yaml``) holding what a
CHEMKIN input + thermo file hold for the mechanisms PeleLM compiles in
(``Tools/CMake/BuildPeleLMExe.cmake:70-73``): elements, species with element
composition and two-range 7-coefficient NASA polynomials, and reactions with
modified-Arrhenius parameters (A [mol-cm-s], beta, Ea [cal/mol]), third-body
efficiencies and Lindemann / Troe fall-off.  The kinetics these tables feed are the
ones specified in ``Docs/source/manual/Model.rst:227-296``.

This module only parses and validates; ``codegen.py`` turns a parsed mechanism into
the flat numeric table (``<name>.tbl``) and the generated HIP rate/Jacobian code.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field

import yaml

# CHEMKIN-II atomic weights [g/mol] (the values CHEMKIN-derived mechanism codes use).
ATOMIC_WEIGHTS = {"H": 1.00797, "O": 15.9994, "C": 12.01115, "N": 14.0067, "AR": 39.948, "HE": 4.0026}

# CHEMKIN-II gas constant.  The PMF profiles PeleLM ships satisfy the ideal-gas law
# with this value to 3e-8 relative (Exec/RegTests/FlameSheet/drm19_pmf.dat rows), but
# only to 5.7e-6 with the CODATA-2018 value, so the mechanism tables use it.
R_CGS = 8.314510e7              # erg/(mol K)
R_CAL = R_CGS / 4.184e7         # cal/(mol K)
P_ATM = 1.01325e6               # dyn/cm^2

HERE = os.path.dirname(os.path.abspath(__file__))


@dataclass
class Species:
    name: str
    composition: dict
    T: tuple          # (Tlo, Tmid, Thi)
    low: list         # 7 coefficients, T < Tmid
    high: list        # 7 coefficients, T >= Tmid
    W: float = 0.0


@dataclass
class Reaction:
    equation: str
    reactants: dict   # species index -> stoichiometric coefficient (int)
    products: dict
    reversible: bool
    kind: str         # "elementary" | "three-body" | "falloff"
    A: float
    beta: float
    Ea: float         # cal/mol
    low: tuple | None = None          # (A0, beta0, Ea0) for falloff
    troe: tuple | None = None         # (a, T3, T1[, T2])
    efficiencies: dict = field(default_factory=dict)   # species index -> efficiency (default 1)
    duplicate: bool = False

    @property
    def has_third_body(self) -> bool:
        return self.kind in ("three-body", "falloff")


@dataclass
class Mechanism:
    name: str
    elements: list
    species: list
    reactions: list

    @property
    def nspec(self) -> int:
        return len(self.species)

    @property
    def nreac(self) -> int:
        return len(self.reactions)

    def index(self, name: str) -> int:
        for i, s in enumerate(self.species):
            if s.name == name:
                return i
        raise KeyError(name)

    @property
    def names(self) -> list:
        return [s.name for s in self.species]

    @property
    def W(self) -> list:
        return [s.W for s in self.species]


_TOKEN = re.compile(r"^(\d*)\s*(.+)$")


def _parse_side(text: str, names: list) -> tuple[dict, str | None]:
    """Parse one side of a reaction equation. Returns ({idx: coef}, third-body marker)."""
    marker = None
    t = text.strip()
    m = re.search(r"\(\+\s*M\s*\)", t)
    if m:
        marker = "(+M)"
        t = (t[: m.start()] + t[m.end():]).strip()
    out: dict = {}
    for tok in [x.strip() for x in re.split(r"\s+\+\s+", " " + t + " ") if x.strip()]:
        if tok == "M":
            marker = "M"
            continue
        mm = _TOKEN.match(tok)
        coef = int(mm.group(1)) if mm.group(1) else 1
        name = mm.group(2).strip()
        if name not in names:
            raise ValueError(f"unknown species {name!r} in {text!r}")
        idx = names.index(name)
        out[idx] = out.get(idx, 0) + coef
    return out, marker


def load(path_or_name: str) -> Mechanism:
    path = path_or_name
    if not os.path.exists(path):
        path = os.path.join(HERE, f"{path_or_name}.yaml")
    with open(path) as f:
        doc = yaml.safe_load(f)
    elements = list(doc["elements"])
    species = []
    for s in doc["species"]:
        th = s["thermo"]
        sp = Species(s["name"], dict(s["composition"]), tuple(th["T"]),
                     [float(x) for x in th["low"]], [float(x) for x in th["high"]])
        for el in sp.composition:
            if el not in elements:
                raise ValueError(f"{sp.name}: element {el} not declared")
        sp.W = sum(ATOMIC_WEIGHTS[el] * n for el, n in sp.composition.items())
        species.append(sp)
    names = [s.name for s in species]
    reactions = []
    for r in doc["reactions"]:
        eq = r["eq"]
        if "<=>" in eq:
            lhs, rhs = eq.split("<=>")
            rev = True
        elif "=>" in eq:
            lhs, rhs = eq.split("=>")
            rev = False
        else:
            raise ValueError(eq)
        reac, m1 = _parse_side(lhs, names)
        prod, m2 = _parse_side(rhs, names)
        kind = r.get("type", "elementary")
        if kind == "three-body" and not (m1 == "M" and m2 == "M"):
            raise ValueError(f"three-body reaction without M: {eq}")
        if kind == "falloff" and not (m1 == "(+M)" and m2 == "(+M)"):
            raise ValueError(f"falloff reaction without (+M): {eq}")
        if kind == "elementary" and (m1 or m2):
            raise ValueError(f"M in elementary reaction: {eq}")
        A, b, Ea = (float(x) for x in r["rate"])
        eff = {names.index(k): float(v) for k, v in (r.get("efficiencies") or {}).items()}
        rx = Reaction(eq, reac, prod, rev, kind, A, b, Ea,
                      low=tuple(float(x) for x in r["low"]) if "low" in r else None,
                      troe=tuple(float(x) for x in r["troe"]) if "troe" in r else None,
                      efficiencies=eff, duplicate=bool(r.get("duplicate", False)))
        if kind == "falloff" and rx.low is None:
            raise ValueError(f"falloff without low-pressure limit: {eq}")
        # element balance
        for el in elements:
            lhs_n = sum(species[i].composition.get(el, 0) * c for i, c in reac.items())
            rhs_n = sum(species[i].composition.get(el, 0) * c for i, c in prod.items())
            if lhs_n != rhs_n:
                raise ValueError(f"element {el} not balanced in {eq}")
        reactions.append(rx)
    return Mechanism(doc["name"], elements, species, reactions)


# ---------------------------------------------------------------------------
# Plain-numpy thermo helpers used by the generator and by host-side checks.
# ---------------------------------------------------------------------------
def nasa_coeffs(sp: Species, T: float) -> list:
    return sp.low if T < sp.T[1] else sp.high


def cp_R(sp: Species, T: float) -> float:
    a = nasa_coeffs(sp, T)
    return a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])))


def h_RT(sp: Species, T: float) -> float:
    a = nasa_coeffs(sp, T)
    return a[0] + T * (a[1] / 2 + T * (a[2] / 3 + T * (a[3] / 4 + T * a[4] / 5))) + a[5] / T


def s_R(sp: Species, T: float) -> float:
    import math
    a = nasa_coeffs(sp, T)
    return a[0] * math.log(T) + T * (a[1] + T * (a[2] / 2 + T * (a[3] / 3 + T * a[4] / 4))) + a[6]
