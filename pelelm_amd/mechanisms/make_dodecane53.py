"""Write dodecane53.yaml: a 53-species n-dodecane/air mechanism of the dodecane_lu class.

PeleLM's FlameSheet case also runs Lu's 53-species skeletal n-dodecane mechanism
(species list and order: Exec/RegTests/FlameSheet/dodecaneLu_pmf.dat:1).  Its rate and
thermo data are not in the reference tree and cannot be restated from memory, so this
mechanism keeps that species set and order and builds the chemistry from:

* the H2/O2/CO/C1-C2 core of drm19 (drm19.yaml; reactions on argon dropped) and the
  H2O2 reactions of LiDryer (LiDryer.yaml), with their NASA-7 thermo;
* class-generic n-alkane reactions for everything above C2 -- fuel C-C fission and
  H-abstraction by H, O, OH, HO2, CH3 and O2 at primary and secondary sites, alkyl
  isomerisation, beta-scission to ethene and to alkene + H, alkene allylic C-C fission,
  allyl/vinyl/acrolein chemistry, and the lumped low-temperature branch
  R + O2 <=> RO2 <=> QOOH, QOOH + O2 <=> O2QOOH => ketohydroperoxide + OH;
* NASA-7 thermo for the species the core lacks, from a rigid-rotor/harmonic-oscillator
  heat capacity (mode frequencies by bond type) fitted in two ranges, and enthalpies
  and entropies of formation by group additivity.

Every reaction is element-balanced and reversible through equilibrium constants from
this thermo, so the invariants the mechanism tests check (mass and element
conservation, detailed balance, Jacobian) hold; its parity with PelePhysics'
dodecane_lu is unpinned.  Run `python -m pelelm_amd.mechanisms.make_dodecane53` to
regenerate the YAML (deterministic).
"""
from __future__ import annotations

import math
import os

import numpy as np
import yaml

HERE = os.path.dirname(os.path.abspath(__file__))
R_KCAL = 1.98720425864e-3   # kcal/(mol K)

SPECIES = ["NC12H26", "H", "O", "OH", "HO2", "H2", "H2O", "H2O2", "O2", "CH2", "CH2D", "CH3", "CH4", "HCO",
           "CH2O", "CH3O", "CO", "CO2", "C2H2", "C2H3", "C2H4", "C2H5", "C2H6", "CH2CHO", "aC3H5", "C3H6",
           "nC3H7", "C2H3CHO", "C4H7", "C4H81", "pC4H9", "C5H9", "C5H10", "PXC5H11", "C6H12", "PXC6H13",
           "C7H14", "PXC7H15", "C8H16", "PXC8H17", "C9H18", "PXC9H19", "C10H20", "PXC10H21", "PXC12H25",
           "SXC12H25", "S3XC12H25", "C12H24", "C12H25O2", "C12OOH", "O2C12H24OOH", "OC12H23OOH", "N2"]

# the core mechanisms' names for the same species
ALIAS = {"CH2D": "CH2(S)"}


def composition(name: str) -> dict:
    fixed = {"CH2D": {"C": 1, "H": 2}, "aC3H5": {"C": 3, "H": 5}, "nC3H7": {"C": 3, "H": 7},
             "pC4H9": {"C": 4, "H": 9}, "C4H81": {"C": 4, "H": 8}, "C2H3CHO": {"C": 3, "H": 4, "O": 1},
             "CH2CHO": {"C": 2, "H": 3, "O": 1}, "C12OOH": {"C": 12, "H": 25, "O": 2},
             "O2C12H24OOH": {"C": 12, "H": 25, "O": 4}, "OC12H23OOH": {"C": 12, "H": 24, "O": 3},
             "NC12H26": {"C": 12, "H": 26}}
    if name in fixed:
        return fixed[name]
    body = name
    for pre in ("PX", "SX", "S3X"):
        if body.startswith(pre):
            body = body[len(pre):]
    comp = {}
    i = 0
    while i < len(body):
        el = body[i]
        i += 1
        num = ""
        while i < len(body) and body[i].isdigit():
            num += body[i]
            i += 1
        comp[el] = comp.get(el, 0) + (int(num) if num else 1)
    return comp


# ---------------------------------------------------------------------------
# thermo for species the core mechanisms lack
# ---------------------------------------------------------------------------
def cp_rrho(comp: dict, linear: bool = False) -> callable:
    """cp/R(T) of a rigid rotor with harmonic modes assigned by bond type: per H one
    stretch (4300 K) and two bends (2000 K), per O one mode at 1500 K; the remaining
    3 n - 6 modes split between skeletal stretches (1400 K) and bends/torsions (500 K)."""
    n = sum(comp.values())
    nH, nO = comp.get("H", 0), comp.get("O", 0)
    modes = 3 * n - (5 if linear else 6)
    theta = [4300.0] * nH + [2000.0] * (2 * nH) + [1500.0] * nO
    theta = theta[:modes]
    rest = modes - len(theta)
    theta += [1400.0] * (rest // 2) + [500.0] * (rest - rest // 2)
    th = np.array(theta)
    base = 3.5 if linear else 4.0

    def cp(T):
        x = th[None, :] / np.asarray(T, dtype=float)[:, None]
        return base + np.sum(x * x * np.exp(x) / np.expm1(x) ** 2, axis=1)
    return cp


def nasa_fit(cp, h298_kcal: float, s298_R: float, Tlo=300.0, Tmid=1000.0, Thi=3500.0):
    """Two-range NASA-7 coefficients: cp/R polynomial fits (the high range shifted to meet
    the low one at Tmid), a6/a7 from h(298) and s(298), continued at Tmid."""
    def fit(a, b):
        T = np.linspace(a, b, 200)
        return np.polynomial.polynomial.polyfit(T, cp(T), 4)
    lo = fit(Tlo, Tmid)
    hi = fit(Tmid, Thi)
    pv = lambda c, T: sum(c[k] * T ** k for k in range(5))  # noqa: E731
    hi[0] += pv(lo, Tmid) - pv(hi, Tmid)

    def h_int(c, T):   # integral of cp/R dT
        return sum(c[k] * T ** (k + 1) / (k + 1) for k in range(5))

    def s_int(c, T):   # integral of cp/(R T) dT
        return c[0] * math.log(T) + sum(c[k] * T ** k / k for k in range(1, 5))
    T0 = 298.15
    a6_lo = h298_kcal / R_KCAL - h_int(lo, T0)
    a7_lo = s298_R - s_int(lo, T0)
    a6_hi = a6_lo + h_int(lo, Tmid) - h_int(hi, Tmid)
    a7_hi = a7_lo + s_int(lo, Tmid) - s_int(hi, Tmid)
    return [float(v) for v in lo] + [a6_lo, a7_lo], [float(v) for v in hi] + [a6_hi, a7_hi]


def group_additivity(name: str, comp: dict):
    """(h298 kcal/mol, s298 J/mol/K) by group additivity around the n-alkane series."""
    nC = comp.get("C", 0)
    alkane_h = lambda n: -10.0 - 4.99 * n        # noqa: E731  (C2H6 -20.0, n-C12H26 -69.9)
    alkane_s = lambda n: 229.5 + 40.0 * (n - 2)  # noqa: E731  (C2H6 229.5, +40 per CH2)
    table = {"C2H2": (54.2, 200.9), "C2H3": (71.6, 233.6), "CH2CHO": (3.0, 260.0), "aC3H5": (40.0, 257.0),
             "C2H3CHO": (-18.0, 280.0), "C4H7": (30.0, 300.0), "C5H9": (25.0, 340.0),
             "NC12H26": (alkane_h(12), alkane_s(12))}
    if name in table:
        return table[name]
    if name.startswith("PX") or name in ("nC3H7", "pC4H9"):        # primary n-alkyl
        return alkane_h(nC) + 48.9, alkane_s(nC) + 10.0
    if name.startswith("SX") or name.startswith("S3X"):              # secondary n-alkyl
        return alkane_h(nC) + 46.4, alkane_s(nC) + 10.0
    if name in ("C3H6", "C4H81") or (name.startswith("C") and comp.get("H") == 2 * nC and "O" not in comp):
        return 20.0 - 5.0 * nC, 267.0 + 40.0 * (nC - 3)             # 1-alkenes
    if name == "C12H25O2":
        return alkane_h(12) + 48.9 - 34.0, alkane_s(12) + 60.0
    if name == "C12OOH":
        return alkane_h(12) + 48.9 - 24.0, alkane_s(12) + 70.0
    if name == "O2C12H24OOH":
        return alkane_h(12) + 48.9 - 58.0, alkane_s(12) + 120.0
    if name == "OC12H23OOH":
        return alkane_h(12) + 48.9 - 58.0 - 9.3 - 10.0, alkane_s(12) + 90.0
    raise KeyError(name)


# ---------------------------------------------------------------------------
# reactions
# ---------------------------------------------------------------------------
def r(eq, A, b, Ea, **kw):
    d = {"eq": eq, "rate": [float(A), float(b), float(Ea)]}
    d.update(kw)
    return d


def alkane_class_reactions():
    R = []
    fuel = "NC12H26"
    # fuel C-C fission
    for a, b in (("PXC5H11", "PXC7H15"), ("PXC6H13", "PXC6H13"), ("PXC8H17", "pC4H9"), ("PXC9H19", "nC3H7"),
                 ("PXC10H21", "C2H5")):
        R.append(r(f"{fuel} <=> {a} + {b}", 5.0e16, 0.0, 82000.0))
    # H-abstraction: sites primary (6 H -> PXC12H25), secondary next to the end (4 H ->
    # SXC12H25) and interior secondary (16 H -> S3XC12H25); per-H rate parameters
    per_h = {  # abstractor: (prim A n Ea, sec A n Ea, product)
        "H": ((2.2e5, 2.54, 6756.0), (1.1e5, 2.40, 4471.0), "H2"),
        "O": ((1.7e5, 2.55, 5000.0), (1.9e5, 2.40, 2750.0), "OH"),
        "OH": ((1.7e6, 1.80, 1430.0), (1.2e6, 1.80, 240.0), "H2O"),
        "HO2": ((6.8e3, 2.55, 16490.0), (1.6e4, 2.60, 13910.0), "H2O2"),
        "CH3": ((2.5e-1, 3.65, 7154.0), (2.5e-1, 3.46, 5480.0), "CH4"),
        "O2": ((3.5e12, 0.0, 50930.0), (2.0e12, 0.0, 47590.0), "HO2"),
    }
    for x, (prim, sec, xh) in per_h.items():
        for rad, nh, p in (("PXC12H25", 6, prim), ("SXC12H25", 4, sec), ("S3XC12H25", 16, sec)):
            R.append(r(f"{fuel} + {x} <=> {rad} + {xh}", p[0] * nh, p[1], p[2]))
    # dodecyl isomerisation and decomposition
    R.append(r("PXC12H25 <=> S3XC12H25", 3.7e11, 0.0, 20800.0))
    R.append(r("SXC12H25 <=> S3XC12H25", 1.0e11, 0.0, 23000.0))
    R.append(r("PXC12H25 <=> C2H4 + PXC10H21", 2.5e13, 0.0, 28800.0))
    R.append(r("SXC12H25 <=> C3H6 + PXC9H19", 2.0e13, 0.0, 28400.0))
    R.append(r("S3XC12H25 <=> C4H81 + PXC8H17", 1.0e13, 0.0, 28400.0))
    R.append(r("S3XC12H25 <=> C5H10 + PXC7H15", 1.0e13, 0.0, 28400.0))
    R.append(r("PXC12H25 <=> C12H24 + H", 2.0e13, 0.0, 35500.0))
    R.append(r("SXC12H25 <=> C12H24 + H", 2.0e13, 0.0, 35000.0))
    # n-alkyl beta-scission to ethene and to 1-alkene + H
    chain = {3: "nC3H7", 4: "pC4H9", 5: "PXC5H11", 6: "PXC6H13", 7: "PXC7H15", 8: "PXC8H17", 9: "PXC9H19",
             10: "PXC10H21"}
    alkene = {3: "C3H6", 4: "C4H81", 5: "C5H10", 6: "C6H12", 7: "C7H14", 8: "C8H16", 9: "C9H18", 10: "C10H20"}
    for n in range(3, 11):
        lower = "CH3" if n == 3 else ("C2H5" if n == 4 else chain[n - 2])
        R.append(r(f"{chain[n]} <=> C2H4 + {lower}", 2.5e13, 0.0, 28800.0))
        R.append(r(f"{chain[n]} <=> {alkene[n]} + H", 2.0e13, 0.0, 35500.0))
    # 1-alkene allylic C-C fission
    alkyl = {1: "CH3", 2: "C2H5", 3: "nC3H7", 4: "pC4H9", 5: "PXC5H11", 6: "PXC6H13", 7: "PXC7H15",
             9: "PXC9H19"}
    for n, ene in ((4, "C4H81"), (5, "C5H10"), (6, "C6H12"), (7, "C7H14"), (8, "C8H16"), (9, "C9H18"),
                   (10, "C10H20"), (12, "C12H24")):
        R.append(r(f"{ene} <=> aC3H5 + {alkyl[n - 3]}", 2.5e16, 0.0, 71000.0))
    # C3-C5 unsaturated chemistry
    R += [
        r("C3H6 <=> aC3H5 + H", 2.0e61, -13.26, 118500.0),
        r("C3H6 + OH <=> aC3H5 + H2O", 3.1e6, 2.0, -298.0),
        r("C3H6 + H <=> aC3H5 + H2", 1.7e5, 2.5, 2490.0),
        r("C3H6 + O <=> CH2CHO + CH3", 1.58e7, 1.76, -1216.0),
        r("aC3H5 + O2 <=> C2H3CHO + OH", 1.82e13, -0.41, 22859.0),
        r("aC3H5 + HO2 <=> C2H3CHO + H2O", 1.0e12, 0.0, 0.0),
        r("C2H3CHO + OH <=> C2H3 + CO + H2O", 9.24e6, 1.5, -962.0),
        r("C2H3CHO + H <=> C2H3 + CH2O", 1.0e13, 0.0, 3500.0),
        r("C4H81 + H <=> C4H7 + H2", 6.5e5, 2.54, 6756.0),
        r("C4H81 + OH <=> C4H7 + H2O", 7.0e2, 2.66, -527.0),
        r("C4H7 <=> C2H4 + C2H3", 1.0e11, 0.0, 37000.0),
        r("C4H7 + H <=> C4H81", 3.6e13, 0.0, 0.0),
        r("C5H10 + H <=> C5H9 + H2", 2.8e7, 2.0, 1500.0),
        r("C5H9 <=> C2H4 + aC3H5", 2.5e13, 0.0, 45000.0),
        r("C5H9 + H <=> C5H10", 3.6e13, 0.0, 0.0),
        # C2H2 / C2H3 / CH2CHO
        r("C2H3 + O2 <=> CH2O + HCO", 4.58e16, -1.39, 1015.0),
        r("C2H3 + O2 <=> CH2CHO + O", 3.03e11, 0.29, 11.0),
        r("C2H3 + H <=> C2H2 + H2", 3.0e13, 0.0, 0.0),
        r("C2H3 <=> C2H2 + H", 3.86e8, 1.62, 37048.0),
        r("C2H4 + OH <=> C2H3 + H2O", 3.6e6, 2.0, 2500.0),
        r("C2H4 + H <=> C2H3 + H2", 1.33e6, 2.53, 12240.0),
        r("C2H2 + O <=> CH2 + CO", 6.1e6, 2.0, 1900.0),
        r("C2H2 + OH <=> CH3 + CO", 4.83e-4, 4.0, -2000.0),
        r("CH2CHO <=> CH3 + CO", 7.8e41, -9.147, 46900.0),
        r("CH2CHO + O2 <=> CH2O + CO + OH", 1.81e10, 0.0, 0.0),
        r("CH2CHO + H <=> CH3 + HCO", 9.0e13, 0.0, 0.0),
    ]
    # low-temperature branch (lumped over the dodecyl isomers)
    for rad in ("PXC12H25", "SXC12H25", "S3XC12H25"):
        R.append(r(f"{rad} + O2 <=> C12H25O2", 4.5e12, 0.0, 0.0))
    R += [
        r("C12H25O2 <=> C12OOH", 2.5e10, 0.0, 20450.0),
        r("C12H25O2 <=> C12H24 + HO2", 4.0e12, 0.0, 30000.0),
        r("C12OOH <=> C12H24 + HO2", 1.0e12, 0.0, 20000.0),
        r("C12OOH + O2 <=> O2C12H24OOH", 4.6e11, 0.0, 0.0),
        r("O2C12H24OOH <=> OC12H23OOH + OH", 2.5e10, 0.0, 21400.0),
        r("OC12H23OOH <=> OH + CH2CHO + CH2O + C9H18", 1.0e16, 0.0, 39000.0),
    ]
    return R


def nasa_g_RT(th: dict, T: float) -> tuple:
    """(h/RT, s/R) of a NASA-7 entry at T."""
    a = th["low"] if T < th["T"][1] else th["high"]
    h = a[0] + a[1] * T / 2 + a[2] * T ** 2 / 3 + a[3] * T ** 3 / 4 + a[4] * T ** 4 / 5 + a[5] / T
    s = a[0] * math.log(T) + a[1] * T + a[2] * T ** 2 / 2 + a[3] * T ** 3 / 3 + a[4] * T ** 4 / 4 + a[6]
    return h, s


def side_species(side: str) -> list:
    out = []
    for tok in side.replace("(+M)", "").split("+"):
        tok = tok.strip()
        if tok in ("M", ""):
            continue
        parts = tok.split()
        out += [parts[1]] * int(parts[0]) if len(parts) == 2 else [tok]
    return out


# collision-limit ceilings for the reverse rate constants of the class-generic reactions
# (CGS: 1/s, cm^3/mol/s, cm^6/mol^2/s, cm^9/mol^3/s by reaction order)
K_CEILING = {1: 1.0e15, 2: 2.0e14, 3: 1.0e17, 4: 1.0e20}
T_GRID = np.linspace(250.0, 3500.0, 131)


def cap_reverse(rx: dict, thermo: dict) -> dict:
    """Raise the activation energy of a reaction whose reverse rate constant, from the
    equilibrium constant of this thermo, would exceed the collision limit anywhere in
    250-3500 K (an endothermic direction must carry at least its reaction enthalpy)."""
    lhs, rhs = rx["eq"].split("<=>")
    sl, sr = side_species(lhs), side_species(rhs)
    dn = len(sr) - len(sl)
    A, b, Ea = rx["rate"]
    worst = 0.0
    for T in T_GRID:
        g = lambda sp: nasa_g_RT(thermo[sp], T)  # noqa: E731
        dh = sum(g(s)[0] for s in sr) - sum(g(s)[0] for s in sl)
        ds = sum(g(s)[1] for s in sr) - sum(g(s)[1] for s in sl)
        lnKc = ds - dh + dn * math.log(1.01325e6 / (8.31446261815324e7 * T))
        lnkr = math.log(A) + b * math.log(T) - Ea / (R_KCAL * 1e3 * T) - lnKc
        worst = max(worst, R_KCAL * 1e3 * T * (lnkr - math.log(K_CEILING[len(sr)])))
    if worst > 0.0:
        rx = dict(rx)
        rx["rate"] = [A, b, float(math.ceil(Ea + worst))]
    return rx


def load_yaml(name):
    with open(os.path.join(HERE, f"{name}.yaml")) as f:
        return yaml.safe_load(f)


def build() -> dict:
    drm = load_yaml("drm19")
    li = load_yaml("LiDryer")
    known = {s["name"]: s for s in drm["species"]}
    for s in li["species"]:
        known.setdefault(s["name"], s)
    species = []
    for name in SPECIES:
        src = known.get(ALIAS.get(name, name))
        comp = composition(name)
        if src is not None:
            assert src["composition"] == comp, name
            species.append({"name": name, "composition": comp, "thermo": src["thermo"]})
            continue
        h298, s298 = group_additivity(name, comp)
        lo, hi = nasa_fit(cp_rrho(comp, linear=(name == "C2H2")), h298, s298 / 8.31446261815324)
        species.append({"name": name, "composition": comp,
                        "thermo": {"T": [300.0, 1000.0, 3500.0], "low": lo, "high": hi}})
    names = set(SPECIES)

    def rename(eq):
        return eq.replace("CH2(S)", "CH2D")

    reactions = []
    for rx in drm["reactions"]:
        if "AR" in rx["eq"].replace("(", " ").split():
            continue
        rx = dict(rx)
        rx["eq"] = rename(rx["eq"])
        if "efficiencies" in rx:
            rx["efficiencies"] = {k.replace("CH2(S)", "CH2D"): v
                                  for k, v in rx["efficiencies"].items() if k != "AR"}
        reactions.append(rx)
    for rx in li["reactions"]:
        if "H2O2" in rx["eq"]:
            reactions.append(dict(rx))
    thermo = {s["name"]: s["thermo"] for s in species}
    reactions += [cap_reverse(rx, thermo) for rx in alkane_class_reactions()]
    # every species named in a reaction must exist, every reaction must balance
    for rx in reactions:
        lhs, rhs = rx["eq"].split("<=>")
        counts = []
        for side in (lhs, rhs):
            tot = {}
            for tok in side.replace("(+M)", "").split("+"):
                tok = tok.strip()
                if tok in ("M", ""):
                    continue
                coef = 1
                parts = tok.split()
                if len(parts) == 2:
                    coef, tok = int(parts[0]), parts[1]
                assert tok in names, (tok, rx["eq"])
                for el, c in composition(tok).items():
                    tot[el] = tot.get(el, 0) + coef * c
            counts.append(tot)
        assert counts[0] == counts[1], rx["eq"]
    return {"name": "dodecane53", "elements": ["O", "H", "C", "N"], "species": species, "reactions": reactions}


HEADER = """# dodecane53: a 53-species n-dodecane/air mechanism of the dodecane_lu class, with the
# species and order of Exec/RegTests/FlameSheet/dodecaneLu_pmf.dat:1 (Lu's skeletal
# mechanism, whose data the reference tree does not hold).
#
# Provenance (generated by make_dodecane53.py, deterministic): drm19's H2/O2/CO/C1-C2 core
# (argon reactions dropped) and LiDryer's H2O2 reactions with their NASA-7 thermo;
# class-generic n-alkane chemistry above C2; NASA-7 thermo for the other species from a
# rigid-rotor/harmonic-oscillator heat capacity and group-additivity enthalpies and
# entropies.  Element-balanced and reversible; parity with PelePhysics' dodecane_lu is
# unpinned.
"""


def main():
    mech = build()
    path = os.path.join(HERE, "dodecane53.yaml")
    with open(path, "w") as f:
        f.write(HEADER)
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
    print(f"wrote {path}: {len(mech['species'])} species, {len(mech['reactions'])} reactions")


if __name__ == "__main__":
    main()
