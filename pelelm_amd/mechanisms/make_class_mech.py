"""Write heptane88.yaml and dodecane56.yaml: mechanisms with the species sets and orders of
the reference's heptane_lu and dodecane_wang flame profiles.

PeleLM's FlameSheet case runs Lu's 88-species skeletal n-heptane mechanism and Wang's
56-species n-dodecane mechanism (species lists and state order:
Exec/RegTests/FlameSheet/heptaneLu_pmf.dat:1 and dodecaneWang_pmf.dat:1).  Their rate and
thermo data are not in the reference tree and cannot be restated from memory, so each
mechanism here keeps the profile's species list and order and builds its chemistry from
what this repository already holds plus class rules over compositions:

* thermo: the NASA-7 entries of drm19, LiDryer and dodecane53 for the species they share
  (under the profile's names), standard enthalpies and entropies of formation at 298 K
  for the other C0-C4 species (ATcT/NIST-JANAF-level values) and, for the C5+ and
  oxygenated heptane-class species, group additivity around the n-alkane series; the
  heat capacity of every new species from the rigid-rotor/harmonic-oscillator estimate of
  make_dodecane53.py, fitted in two ranges;
* reactions: every reaction of drm19 (argon dropped), LiDryer's H2O2 reactions and the
  class-generic chemistry of dodecane53 whose species all exist under the target's names;
  then, for each species, class rules over compositions: H-abstraction by H, O, OH, HO2,
  CH3 and O2 from each stable species to each radical one H lighter, C-C / O-O fission of
  stable species into two radicals, beta-scission of radicals into a stable species and a
  radical, radical + H recombination, radical + O2 addition to the peroxy species of that
  composition, and isomerisation between species of one composition (RO2 <=> QOOH);
  species still unconnected decompose into a listed species plus small fragments.

Every reaction is element-balanced (checked) and reversible through equilibrium constants
from this thermo, with activation energies raised where a reverse rate would pass the
collision limit, so the invariants the mechanism tests check hold; parity with
PelePhysics' heptane_lu / dodecane_wang is unpinned.  Run
`python -m pelelm_amd.mechanisms.make_class_mech` to regenerate both (deterministic).
"""
from __future__ import annotations

import itertools
import os
import re

import yaml

from .make_dodecane53 import R_KCAL, cap_reverse, cp_rrho, load_yaml, nasa_fit

HERE = os.path.dirname(os.path.abspath(__file__))

HEPTANE88 = ["H", "H2", "O", "O2", "OH", "H2O", "CO", "HCO", "CO2", "CH3", "CH4", "HO2", "H2O2", "CH2O", "CH3O",
             "C2H6", "C2H4", "C2H5", "C2H2", "C2H3", "CH2CO", "HCCO", "CH3CO", "CH2CHO", "CH3CHO", "C3H4-A", "C3H6",
             "C4H6", "NC3H7", "C4H7", "C4H8-1", "PC4H9", "CH3COCH2", "C2H5CHO", "C2H5CO", "C5H9", "C5H10-1",
             "C2H5O", "CH3O2", "CH3O2H", "C2H3CO", "C2H3CHO", "C3H5O", "C4H7O", "C4H8OOH1-3O2", "C4H8OOH1-3",
             "PC4H9O2", "C3H5-A", "C3H3", "C3H2", "CH2(S)", "NC4KET13", "NC3H7CHO", "NC3H7CO", "C2H5COCH2",
             "NC3H7COCH2", "NC4H9CHO", "NC4H9CO", "NC7H16", "C7H15-1", "C7H15-2", "C7H15-3", "C7H15-4", "C7H14-2",
             "C7H14-3", "C7H15O2-1", "C7H15O2-2", "C7H15O2-3", "C7H15O2-4", "C7H14OOH1-3", "C7H14OOH2-3",
             "C7H14OOH2-4", "C7H14OOH3-2", "C7H14OOH3-4", "C7H14OOH3-5", "C7H14OOH4-2", "C7H14OOH4-3",
             "C7H14OOH1-3O2", "C7H14OOH2-4O2", "C7H14OOH3-5O2", "C7H14OOH4-2O2", "C7H14O1-3", "C7H14O2-4",
             "NC7KET13", "NC7KET24", "NC7KET35", "NC7KET42", "N2"]

DODECANE56 = ["H", "O", "OH", "HO2", "H2", "H2O", "H2O2", "O2", "CH", "CH2", "CH2*", "CH3", "CH4", "HCO", "CH2O",
              "CH3O", "CH2OH", "CH3OH", "CO", "CO2", "C2H", "C2H2", "C2H3", "C2H4", "C2H5", "C2H6", "HCCO", "CH2CO",
              "CH3CO", "CH2CHO", "CH3CHO", "C3H3", "pC3H4", "aC3H4", "aC3H5", "CH3CCH2", "C3H6", "nC3H7", "iC3H7",
              "C2H3CHO", "C4H2", "iC4H3", "C4H4", "iC4H5", "C4H5-2", "C4H6", "C4H612", "C4H6-2", "C4H7", "C4H81",
              "pC4H9", "NC12H26", "C6H12", "C6H11", "C5H10", "N2"]

# the names the source mechanisms use for the same species
SOURCE_NAME = {"CH2(S)": ["CH2(S)", "CH2D"], "CH2*": ["CH2(S)", "CH2D"], "C3H5-A": ["aC3H5"], "aC3H5": ["aC3H5"],
               "C4H8-1": ["C4H81"], "C4H81": ["C4H81"], "NC3H7": ["nC3H7"], "nC3H7": ["nC3H7"],
               "PC4H9": ["pC4H9"], "pC4H9": ["pC4H9"], "C5H10-1": ["C5H10"], "C5H10": ["C5H10"],
               "C6H12": ["C6H12"]}

# compositions the formula parser cannot read from the name
FIXED_COMP = {"NC7KET13": {"C": 7, "H": 14, "O": 3}, "NC7KET24": {"C": 7, "H": 14, "O": 3},
              "NC7KET35": {"C": 7, "H": 14, "O": 3}, "NC7KET42": {"C": 7, "H": 14, "O": 3},
              "NC4KET13": {"C": 4, "H": 8, "O": 3}, "C4H612": {"C": 4, "H": 6}, "CH2(S)": {"C": 1, "H": 2},
              "CH2*": {"C": 1, "H": 2}, "C4H81": {"C": 4, "H": 8}, "C7H14O1-3": {"C": 7, "H": 14, "O": 1},
              "C7H14O2-4": {"C": 7, "H": 14, "O": 1}}

# standard enthalpy (kcal/mol) and entropy (J/mol/K) of formation at 298 K, C0-C4 species
# the source mechanisms lack (ATcT / NIST-JANAF-level values)
H298_S298 = {
    "CH": (142.5, 183.0), "CH2OH": (-4.1, 244.2), "CH3OH": (-48.0, 239.9), "C2H": (135.0, 207.4),
    "HCCO": (42.4, 254.9), "CH2CO": (-11.4, 241.8), "CH3CO": (-2.4, 267.5), "CH3CHO": (-39.7, 263.8),
    "C3H3": (83.0, 257.0), "C3H2": (130.0, 240.0), "pC3H4": (44.3, 248.2), "aC3H4": (45.6, 243.9),
    "C3H4-A": (45.6, 243.9), "CH3CCH2": (60.0, 270.0), "iC3H7": (21.5, 289.0), "C4H2": (111.0, 250.0),
    "iC4H3": (119.0, 270.0), "C4H4": (70.0, 279.0), "iC4H5": (75.0, 293.0), "C4H5-2": (74.0, 293.0),
    "C4H6": (26.3, 278.7), "C4H612": (38.8, 293.0), "C4H6-2": (34.7, 283.3), "CH3COCH2": (-8.0, 290.0),
    "C2H5CHO": (-45.1, 304.5), "C2H5CO": (-7.6, 320.0), "C2H5O": (-3.3, 274.0), "CH3O2": (2.9, 270.0),
    "CH3O2H": (-31.3, 280.0), "C2H3CO": (19.5, 290.0), "C3H5O": (18.0, 300.0), "C4H7O": (8.0, 340.0),
    "C6H11": (30.0, 380.0)}


def composition(name: str) -> dict:
    if name in FIXED_COMP:
        return dict(FIXED_COMP[name])
    body = re.sub(r"(?<=OOH)\d-\d", "", name)   # hydroperoxide site tags: C7H14OOH2-4O2
    body = re.sub(r"-\w+$", "", body)            # isomer suffixes: C7H15-2, C3H5-A, C7H15O2-1
    for pre in ("PX", "SX", "S3X"):
        if body.startswith(pre):
            body = body[len(pre):]
    while body and body[0].islower():            # iC4H5, aC3H4, pC4H9, nC3H7
        body = body[1:]
    if body[:2] in ("NC", "PC"):                 # normal / primary
        body = body[1:]
    comp, i = {}, 0
    while i < len(body):
        el = body[i]
        i += 1
        assert el in "CHON", (name, body)
        num = ""
        while i < len(body) and body[i].isdigit():
            num += body[i]
            i += 1
        comp[el] = comp.get(el, 0) + (int(num) if num else 1)
    return comp


def is_radical(comp: dict) -> bool:
    return comp.get("H", 0) % 2 == 1 or comp in ({"C": 1, "H": 2}, {"O": 1}, {"O": 2}, {"C": 1}, {"C": 3, "H": 2})


def alkane_h(n):
    return -10.0 - 4.99 * n            # C2H6 -20.0, n-C7H16 -44.9


def alkane_s(n):
    return 229.5 + 40.0 * (n - 2)      # J/mol/K


def group_additivity(name: str, comp: dict):
    """(h298 kcal/mol, s298 J/mol/K) of the heptane-class and C5+ species: offsets from the
    n-alkane of the same carbon count (C-H bond 98 kcal/mol for alkyl radicals, 30 kcal/mol
    per C=C, R-OO 35 kcal/mol, 1,5 H shift +12 kcal/mol, carbonyl from CH2 -21 kcal/mol,
    hydroperoxide -OOH for -H -13 kcal/mol, cyclic ether for two H -19 kcal/mol)."""
    n = comp.get("C", 0)
    H, O = comp.get("H", 0), comp.get("O", 0)
    h0, s0 = alkane_h(n), alkane_s(n)
    if name in H298_S298:
        return H298_S298[name]
    if name.startswith("NC") and H == 2 * n + 2 and O == 0:                 # n-alkane
        return h0, s0
    if O == 0 and H == 2 * n + 1:                                            # alkyl
        return h0 + 46.0 + (1.5 if name.endswith("-1") or name.startswith("P") else 0.0), s0 + 10.0
    if O == 0 and H == 2 * n:                                                # alkene
        return h0 + 30.5 - (1.5 if not name.endswith("-1") else 0.0), s0 - 5.0
    if O == 0 and H == 2 * n - 1:                                            # alkenyl / allylic
        return h0 + 30.5 + 36.0, s0
    if "KET" in name:                                                        # ketohydroperoxide
        return h0 - 21.0 - 13.0 - 13.0, s0 + 70.0
    if "OOH" in name and name.endswith("O2"):                               # O2QOOH
        return h0 + 46.0 - 35.0 - 13.0, s0 + 110.0
    if "OOH" in name:                                                        # QOOH
        return h0 + 46.0 - 13.0, s0 + 70.0
    if name.endswith("O2") or "O2-" in name:                                 # RO2
        return h0 + 46.0 - 35.0, s0 + 50.0
    if O == 1 and H == 2 * n and "CHO" not in name and "CO" not in name:     # cyclic ether
        return h0 - 19.0, s0 - 20.0
    if name.endswith("CHO") or (O == 1 and H == 2 * n and name.endswith("CH2") is False):   # aldehyde
        return h0 - 21.0 + 1.0, s0 + 5.0
    if name.endswith("CO") and H == 2 * n - 1:                                # acyl
        return h0 - 21.0 + 1.0 + 37.0, s0 + 10.0
    if name.endswith("COCH2"):                                               # ketonyl
        return h0 - 21.0 + 46.0 - 6.0, s0 + 10.0
    raise KeyError(name)


def side_species(side: str) -> list:
    out = []
    for tok in side.replace("(+M)", "").split("+"):
        tok = tok.strip()
        if tok in ("M", ""):
            continue
        parts = tok.split()
        out += [parts[1]] * int(parts[0]) if len(parts) == 2 else [tok]
    return out


def r(eq, A, b, Ea, **kw):
    d = {"eq": eq, "rate": [float(A), float(b), float(Ea)]}
    d.update(kw)
    return d


def ckey(comp: dict) -> tuple:
    return tuple(sorted((k, v) for k, v in comp.items() if v))


def add_comp(*comps) -> dict:
    out = {}
    for c in comps:
        for k, v in c.items():
            out[k] = out.get(k, 0) + v
    return {k: v for k, v in out.items() if v}


def heavy(comp: dict) -> int:
    return comp.get("C", 0) + comp.get("O", 0)


# lumped fuel chemistry where the profile's species set holds no fuel radical (Wang's
# n-dodecane set): abstraction straight to an alkene-alkyl split of the dodecyl radical,
# and C-C fission into three fragments
LUMPED = {"dodecane56": [
    ("NC12H26 + H <=> H2 + C6H12 + C2H4 + pC4H9", 2.5e7, 2.4, 4471.0),
    ("NC12H26 + O <=> OH + C6H12 + C2H4 + pC4H9", 1.0e7, 2.5, 2830.0),
    ("NC12H26 + OH <=> H2O + C6H12 + C2H4 + pC4H9", 4.0e7, 2.0, -400.0),
    ("NC12H26 + HO2 <=> H2O2 + C6H12 + C2H4 + pC4H9", 3.0e5, 2.6, 13910.0),
    ("NC12H26 + CH3 <=> CH4 + C6H12 + C2H4 + pC4H9", 30.0, 3.5, 5480.0),
    ("NC12H26 + O2 <=> HO2 + C6H12 + C2H4 + pC4H9", 4.0e14, 0.0, 47590.0),
    ("NC12H26 <=> pC4H9 + nC3H7 + C5H10", 5.0e16, 0.0, 82000.0)]}


def class_reactions(names: list, comps: dict, existing: list) -> list:
    """Reactions from the class rules for the species of `names` (see the module docstring)."""
    by_comp = {}
    for s in names:
        by_comp.setdefault(ckey(comps[s]), []).append(s)
    have = set(names)
    covered = set()
    for rx in existing:
        lhs, rhs = rx["eq"].split("<=>")
        covered.update(side_species(lhs) + side_species(rhs))
    small = {"H", "O", "OH", "HO2", "CH3", "O2", "H2", "H2O", "H2O2", "CH4", "N2", "CO", "CO2"}
    R = []
    seen = set()

    def push(eq, A, b, Ea):
        lhs, rhs = eq.split("<=>")
        key = (tuple(sorted(side_species(lhs))), tuple(sorted(side_species(rhs))))
        if key in seen or key[::-1] in seen or key[0] == key[1]:
            return
        seen.add(key)
        R.append(r(eq, A, b, Ea))

    for rx in existing:
        lhs, rhs = rx["eq"].split("<=>")
        seen.add((tuple(sorted(side_species(lhs))), tuple(sorted(side_species(rhs)))))
    abstract = (("H", "H2", 1.3e6, 2.4, 4471.0), ("O", "OH", 5.5e5, 2.5, 2830.0), ("OH", "H2O", 2.4e6, 2.0, -400.0),
                ("HO2", "H2O2", 1.6e4, 2.6, 13910.0), ("CH3", "CH4", 1.5, 3.5, 5480.0),
                ("O2", "HO2", 2.0e13, 0.0, 47590.0))
    ordered = sorted(names, key=lambda s: (heavy(comps[s]), names.index(s)))
    for s in ordered:
        if s in small:
            continue
        c = comps[s]
        rad = is_radical(c)
        nC = c.get("C", 0)
        if not rad:
            # H-abstraction to every radical one H lighter
            minus_h = add_comp(c, {"H": -1})
            rads = by_comp.get(ckey(minus_h), [])
            for rd in rads:
                share = 1.0 / len(rads)
                for x, xh, A, b, Ea in abstract:
                    push(f"{s} + {x} <=> {rd} + {xh}", A * max(c.get("H", 1), 1) * share, b, Ea)
            # fission into two radicals (C-C, or O-O for hydroperoxides)
            pairs = []
            for a, b_ in itertools.combinations_with_replacement(names, 2):
                if is_radical(comps[a]) and is_radical(comps[b_]) and \
                        ckey(add_comp(comps[a], comps[b_])) == ckey(c) and a not in ("H",):
                    pairs.append((abs(heavy(comps[a]) - heavy(comps[b_])), a, b_))
            pairs.sort()
            for _, a, b_ in pairs[:2]:
                Ea = 43000.0 if ("OOH" in s or "KET" in s or "O2H" in s) and "OH" in (a, b_) else 82000.0
                push(f"{s} <=> {a} + {b_}", 5.0e16, 0.0, Ea)
        else:
            # beta-scission into a stable species and a radical
            pairs = []
            for a in names:
                for b_ in names:
                    ca, cb = comps[a], comps[b_]
                    if is_radical(ca) or not is_radical(cb) or b_ == s or a == s:
                        continue
                    if ckey(add_comp(ca, cb)) == ckey(c):
                        pairs.append((abs(heavy(ca) - heavy(cb)), a, b_))
            pairs.sort()
            for _, a, b_ in pairs[:2]:
                Ea = 24000.0 if b_ == "OH" else 29000.0
                push(f"{s} <=> {a} + {b_}", 2.5e13, 0.0, Ea)
            # recombination with H
            for st in by_comp.get(ckey(add_comp(c, {"H": 1})), [])[:1]:
                if st != s:
                    push(f"{s} + H <=> {st}", 1.0e14, 0.0, 0.0)
            # O2 addition to the peroxy species of that composition, and HO2 elimination
            if nC >= 1 and c.get("O", 0) <= 2:
                for ro2 in by_comp.get(ckey(add_comp(c, {"O": 2})), [])[:1]:
                    push(f"{s} + O2 <=> {ro2}", 4.5e12, 0.0, 0.0)
            for st in by_comp.get(ckey(add_comp(c, {"H": -1})), [])[:1]:
                if not is_radical(comps[st]):
                    push(f"{s} + O2 <=> {st} + HO2", 1.0e12, 0.0, 5000.0)
        # isomerisation within one composition (RO2 <=> QOOH, alkyl isomers): the first
        # species of the composition with each later one
        same = by_comp[ckey(c)]
        if len(same) > 1 and s != same[0]:
            push(f"{same[0]} <=> {s}", 2.5e10, 0.0, 21000.0)
    # species still unconnected: split into a listed species plus small fragments
    frags = [("OH",), ("H",), ("CH3",), ("HCO",), ("CH2O", "OH"), ("CO", "OH"), ("OH", "OH"), ("H2O",), ("HO2",),
             ("CO",), ("CH2O",), ("C2H4",), ("CH2O", "H"), ("CH2O", "HCO"), ("CO", "H")]
    involved = set(covered)
    for rx in R:
        lhs, rhs = rx["eq"].split("<=>")
        involved.update(side_species(lhs) + side_species(rhs))
    for s in ordered:
        if s in involved:
            continue
        c = comps[s]
        done = False
        for fr in frags:
            if not all(f in have for f in fr):
                continue
            rest = add_comp(c, *[{k: -v for k, v in comps[f].items()} for f in fr])
            if any(v < 0 for v in rest.values()) or not rest:
                continue
            for big in by_comp.get(ckey(rest), []):
                if big == s:
                    continue
                push(f"{s} <=> {big} + {' + '.join(fr)}", 1.0e14, 0.0, 40000.0)
                involved.add(s)
                done = True
                break
            if done:
                break
        assert done, f"no reaction found for {s}"
    return R


def balanced(eq: str, comps: dict) -> bool:
    lhs, rhs = eq.split("<=>")
    sides = []
    for side in (lhs, rhs):
        tot = {}
        for sp in side_species(side):
            for k, v in comps[sp].items():
                tot[k] = tot.get(k, 0) + v
        sides.append(tot)
    return sides[0] == sides[1]


def build(name: str, species: list) -> dict:
    srcs = [load_yaml("drm19"), load_yaml("LiDryer"), load_yaml("dodecane53")]
    known = {}
    for m in srcs:
        for s in m["species"]:
            known.setdefault(s["name"], s)
    comps = {s: composition(s) for s in species}
    # the target's name of each source species
    to_target = {}
    for s in species:
        for src_name in SOURCE_NAME.get(s, [s]):
            to_target[src_name] = s
    out_species = []
    for s in species:
        src = next((known[n] for n in SOURCE_NAME.get(s, [s]) if n in known), None)
        if src is not None:
            assert src["composition"] == comps[s], (s, src["composition"], comps[s])
            out_species.append({"name": s, "composition": comps[s], "thermo": src["thermo"]})
            continue
        h298, s298 = group_additivity(s, comps[s])
        lo, hi = nasa_fit(cp_rrho(comps[s], linear=s in ("C2H", "C4H2", "HCCO", "C3H2")), h298, s298 / 8.31446261815324)
        out_species.append({"name": s, "composition": comps[s],
                            "thermo": {"T": [300.0, 1000.0, 3500.0], "low": lo, "high": hi}})
    have = set(species)

    def rename(eq: str):
        toks = eq.replace("(+M)", " (+M) ").split()
        res = []
        for t in toks:
            if t in ("+", "<=>", "=>", "M", "(+M)") or t.replace(".", "").isdigit():
                res.append(t)
                continue
            tn = to_target.get(t, t)
            if tn not in have:
                return None
            res.append(tn)
        return " ".join(res).replace(" (+M)", " (+M)")

    reactions, seen = [], set()
    for m in srcs:
        for rx in m["reactions"]:
            if m["name"] == "LiDryer" and "H2O2" not in rx["eq"]:
                continue
            eq = rename(rx["eq"])
            if eq is None:
                continue
            key = eq.replace(" ", "")
            if key in seen:
                continue
            seen.add(key)
            rx = dict(rx)
            rx["eq"] = eq
            if "efficiencies" in rx:
                rx["efficiencies"] = {to_target.get(k, k): v for k, v in rx["efficiencies"].items()
                                      if to_target.get(k, k) in have}
            reactions.append(rx)
    thermo = {s["name"]: s["thermo"] for s in out_species}
    reactions += [cap_reverse(r(*rx), thermo) for rx in LUMPED.get(name, [])]
    extra = class_reactions(species, comps, reactions)
    reactions += [cap_reverse(rx, thermo) for rx in extra]
    for rx in reactions:
        assert balanced(rx["eq"], comps), rx["eq"]
    return {"name": name, "elements": ["O", "H", "C", "N"], "species": out_species, "reactions": reactions}


HEADER = """# {name}: a {ns}-species mechanism with the species and state order of
# Exec/RegTests/FlameSheet/{pmf}:1 ({desc}, whose data the reference tree does not hold).
#
# Provenance (generated by make_class_mech.py, deterministic): the NASA-7 thermo and
# reactions of drm19, LiDryer (H2O2) and dodecane53 for the species they share, 298 K
# formation enthalpies/entropies and group additivity with a rigid-rotor/harmonic-
# oscillator heat capacity for the others, and class rules over compositions (H-abstraction,
# fission, beta-scission, recombination, O2 addition, isomerisation).  Element-balanced
# and reversible; parity with PelePhysics' mechanism is unpinned.
"""


def write(mech: dict, pmf: str, desc: str) -> str:
    path = os.path.join(HERE, f"{mech['name']}.yaml")
    with open(path, "w") as f:
        f.write(HEADER.format(name=mech["name"], ns=len(mech["species"]), pmf=pmf, desc=desc))
        f.write(f"name: {mech['name']}\n")
        f.write(f"elements: [{', '.join(mech['elements'])}]\n")
        f.write("species:\n")
        for s in mech["species"]:
            comp = ", ".join(f"{k}: {v}" for k, v in s["composition"].items())
            th = s["thermo"]
            fmt = lambda v: ", ".join(f"{x:.8e}" for x in v)  # noqa: E731
            f.write(f"  - name: {s['name']}\n    composition: {{{comp}}}\n")
            f.write(f"    thermo: {{T: [{', '.join(f'{t:.1f}' for t in th['T'])}],\n")
            f.write(f"      low:  [{fmt(th['low'])}],\n")
            f.write(f"      high: [{fmt(th['high'])}]}}\n")
        f.write("reactions:\n")
        for rx in mech["reactions"]:
            f.write("  - " + yaml.safe_dump(rx, default_flow_style=True, sort_keys=False, width=10000).strip() + "\n")
    return path


def main():
    for name, species, pmf, desc in (("heptane88", HEPTANE88, "heptaneLu_pmf.dat", "Lu's skeletal n-heptane mechanism"),
                                     ("dodecane56", DODECANE56, "dodecaneWang_pmf.dat",
                                      "Wang's n-dodecane mechanism")):
        mech = build(name, species)
        path = write(mech, pmf, desc)
        print(f"wrote {path}: {len(mech['species'])} species, {len(mech['reactions'])} reactions")


if __name__ == "__main__":
    main()
