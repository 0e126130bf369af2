"""Emit straight-line HIP device code for one mechanism (one cell per lane).

The generated header defines ``struct mech_<name>`` with static members:

* ``NS``, ``NR``, ``N1`` and the constant tables host code needs (molecular weights,
  NASA-7 coefficients, species names);
* ``thermo_cph(T, cpR, hRT)`` -- cp/R and h/RT of every species (HY2T, energy equation);
* ``gibbs(T, logT, gRT)`` -- g/RT;
* ``wdot(C, T, logT, invT, gRT, wd)`` -- net molar production rates [mol/cm^3/s];
* ``jacobian(C, T, logT, invT, gRT, hRT, wd, dwdT, acc)`` -- the analytic derivatives
  d(wdot_i W_i)/d(rhoY_k) accumulated through ``acc(i, k, value)`` plus dwdot/dT.

NASA-7 range selection is branch-free (coefficients chosen with a select on
T < Tmid) so lanes on both sides of the break temperature stay converged.  Constant
folding: k_f with beta = Ea = 0 is a literal; Troe terms with T*** = 1e-30 / T* = 1e30
vanish / become constants; third-body sums start from the total concentration and
add (eff - 1) C_k only for species with non-unit efficiency.
"""
from __future__ import annotations

import json
import math
import os
import re

from . import mechanism as M


def _d(x: float) -> str:
    r = repr(float(x))
    if "e" not in r and "." not in r and "inf" not in r and "nan" not in r:
        r += ".0"
    return r


def _cname(name: str) -> str:
    return name.replace("(", "_").replace(")", "_").replace("-", "m").replace("*", "s")


def _prod(terms: list) -> str:
    return " * ".join(terms) if terms else "1.0"


def _power_terms(side: dict) -> list:
    t = []
    for i, c in sorted(side.items()):
        t += [f"C[{i}]"] * c
    return t


def _dpower_terms(side: dict, k: int) -> tuple[float, list]:
    """d/dC_k of prod C^nu: coefficient and remaining factors."""
    c = side[k]
    t = []
    for i, ci in sorted(side.items()):
        n = ci - 1 if i == k else ci
        t += [f"C[{i}]"] * n
    return float(c), t


def _thermo_groups(mech: M.Mechanism) -> dict:
    groups: dict = {}
    for k, s in enumerate(mech.species):
        groups.setdefault(s.T[1], []).append(k)
    return groups


def gen_thermo(mech: M.Mechanism) -> str:
    out = []
    out.append("static __device__ __forceinline__ void thermo_cph(const double T, double* __restrict__ cpR, "
               "double* __restrict__ hRT) {")
    out.append("  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T, iT = 1.0 / T;")
    for tmid, ks in _thermo_groups(mech).items():
        out.append(f"  {{ const bool lo = T < {_d(tmid)};")
        for k in ks:
            s = mech.species[k]
            a = [f"(lo ? {_d(s.low[i])} : {_d(s.high[i])})" for i in range(7)]
            out.append(f"    {{ const double a0 = {a[0]}, a1 = {a[1]}, a2 = {a[2]}, a3 = {a[3]}, a4 = {a[4]}, a5 = {a[5]};")
            out.append(f"      cpR[{k}] = a0 + a1 * T + a2 * T2 + a3 * T3 + a4 * T4;")
            out.append(f"      hRT[{k}] = a0 + 0.5 * a1 * T + (1.0 / 3.0) * a2 * T2 + 0.25 * a3 * T3 + 0.2 * a4 * T4 + a5 * iT; }}")
        out.append("  }")
    out.append("}")
    out.append("")
    out.append("static __device__ __forceinline__ void thermo_dcp(const double T, double* __restrict__ dcpR) {")
    out.append("  const double T2 = T * T, T3 = T2 * T;")
    for tmid, ks in _thermo_groups(mech).items():
        out.append(f"  {{ const bool lo = T < {_d(tmid)};")
        for k in ks:
            s_ = mech.species[k]
            a = [f"(lo ? {_d(s_.low[i])} : {_d(s_.high[i])})" for i in range(5)]
            out.append(f"    dcpR[{k}] = {a[1]} + 2.0 * {a[2]} * T + 3.0 * {a[3]} * T2 + 4.0 * {a[4]} * T3;")
        out.append("  }")
    out.append("}")
    out.append("")
    out.append("static __device__ __forceinline__ void gibbs(const double T, const double logT, double* __restrict__ gRT) {")
    out.append("  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T, iT = 1.0 / T;")
    for tmid, ks in _thermo_groups(mech).items():
        out.append(f"  {{ const bool lo = T < {_d(tmid)};")
        for k in ks:
            s = mech.species[k]
            a = [f"(lo ? {_d(s.low[i])} : {_d(s.high[i])})" for i in range(7)]
            out.append(f"    {{ const double a0 = {a[0]}, a1 = {a[1]}, a2 = {a[2]}, a3 = {a[3]}, a4 = {a[4]}, a5 = {a[5]}, a6 = {a[6]};")
            # g/RT = h/RT - s/R
            out.append(f"      gRT[{k}] = a0 * (1.0 - logT) - 0.5 * a1 * T - (1.0 / 6.0) * a2 * T2 - (1.0 / 12.0) * a3 * T3"
                       f" - 0.05 * a4 * T4 + a5 * iT - a6; }}")
        out.append("  }")
    out.append("}")
    return "\n".join(out)


def _kf_expr(A: float, b: float, Ta: float) -> str:
    if b == 0.0 and Ta == 0.0:
        return _d(A)
    parts = [_d(math.log(A))]
    if b != 0.0:
        parts.append(f"{_d(b)} * logT")
    if Ta != 0.0:
        parts.append(f"{_d(-Ta)} * invT")
    return f"exp({' + '.join(parts)})"


def _dlnk_expr(b: float, Ta: float) -> str:
    # d ln(A T^b exp(-Ta/T)) / dT = (b + Ta/T) / T
    if b == 0.0 and Ta == 0.0:
        return "0.0"
    if Ta == 0.0:
        return f"{_d(b)} * invT"
    if b == 0.0:
        return f"{_d(Ta)} * invT * invT"
    return f"({_d(b)} + {_d(Ta)} * invT) * invT"


def _mc_expr(r: M.Reaction, ns: int) -> str:
    terms = ["Ctot"]
    for k in range(ns):
        e = r.efficiencies.get(k, 1.0)
        if e != 1.0:
            terms.append(f"{_d(e - 1.0)} * C[{k}]")
    return " + ".join(terms)


def _troe_fc(r: M.Reaction) -> tuple[str, str]:
    """Fcent and dFcent/dT expressions with constant folding of the 1e-30 / 1e30 forms."""
    a = r.troe[0]
    T3, T1 = r.troe[1], r.troe[2]
    fc, dfc = [], []
    if (1.0 - a) != 0.0 and T3 > 1e-20:
        fc.append(f"{_d(1.0 - a)} * exp({_d(-1.0 / T3)} * T)")
        dfc.append(f"{_d(-(1.0 - a) / T3)} * exp({_d(-1.0 / T3)} * T)")
    if a != 0.0:
        if T1 >= 1e20:
            fc.append(_d(a))
        else:
            fc.append(f"{_d(a)} * exp({_d(-1.0 / T1)} * T)")
            dfc.append(f"{_d(-a / T1)} * exp({_d(-1.0 / T1)} * T)")
    if len(r.troe) == 4:
        T2 = r.troe[3]
        fc.append(f"exp({_d(-T2)} * invT)")
        dfc.append(f"{_d(T2)} * invT * invT * exp({_d(-T2)} * invT)")
    return (" + ".join(fc) or "0.0"), (" + ".join(dfc) or "0.0")


def _kc_exprs(r: M.Reaction) -> tuple[str, str, int]:
    """Sum of nu g/RT (products minus reactants), the same for h/RT, and sum nu."""
    def side(arr: str) -> str:
        out = ""
        for i, c in sorted(r.products.items()):
            t = f"{_d(c)} * {arr}[{i}]" if c != 1 else f"{arr}[{i}]"
            out += t if not out else f" + {t}"
        for i, c in sorted(r.reactants.items()):
            t = f"{_d(c)} * {arr}[{i}]" if c != 1 else f"{arr}[{i}]"
            out += f" - {t}"
        return out
    dnu = sum(r.products.values()) - sum(r.reactants.values())
    return side("gRT"), side("hRT"), dnu


def _refc_factor(dnu: int) -> str:
    # (P_atm / RT)^(-dnu) = (RT/P_atm)^dnu = refCinv^dnu
    if dnu == 0:
        return ""
    if dnu > 0:
        return " * " + " * ".join(["refCinv"] * dnu)
    return " * " + " * ".join(["refC"] * (-dnu))


def gen_wdot(mech: M.Mechanism) -> str:
    ns = mech.nspec
    out = []
    out.append("static __device__ __forceinline__ void wdot(const double* __restrict__ C, const double T, const double logT, "
               "const double invT, const double* __restrict__ gRT, double* __restrict__ wd) {")
    out.append(f"  for (int k = 0; k < {ns}; ++k) wd[k] = 0.0;")
    out.append(f"  const double refC = {_d(M.P_ATM / M.R_CGS)} * invT, refCinv = 1.0 / refC;")
    if any(r.has_third_body for r in mech.reactions):
        out.append("  double Ctot = 0.0;")
        out.append(f"  for (int k = 0; k < {ns}; ++k) Ctot += C[k];")
    for j, r in enumerate(mech.reactions):
        out.append(f"  {{ // {j}: {r.equation}")
        out.append(f"    const double kf = {_kf_expr(r.A, r.beta, r.Ea / M.R_CAL)};")
        if r.kind == "elementary":
            out.append("    const double keff = kf;")
        elif r.kind == "three-body":
            out.append(f"    const double keff = kf * ({_mc_expr(r, ns)});")
        else:
            A0, b0, E0 = r.low
            out.append(f"    const double k0 = {_kf_expr(A0, b0, E0 / M.R_CAL)};")
            out.append(f"    const double Pr = k0 * ({_mc_expr(r, ns)}) / kf;")
            if r.troe is not None:
                fc, _ = _troe_fc(r)
                out.append(f"    const double logFc = log10(fmax({fc}, 1e-300));")
                out.append("    const double cc = -0.4 - 0.67 * logFc, nn = 0.75 - 1.27 * logFc;")
                out.append("    const double u = log10(fmax(Pr, 1e-300)) + cc;")
                out.append("    const double f1 = u / (nn - 0.14 * u);")
                out.append("    const double F = exp(2.302585092994045684 * logFc / (1.0 + f1 * f1));")
                out.append("    const double keff = kf * Pr / (1.0 + Pr) * F;")
            else:
                out.append("    const double keff = kf * Pr / (1.0 + Pr);")
        pf = _prod(_power_terms(r.reactants))
        if r.reversible:
            gs, _, dnu = _kc_exprs(r)
            pr = _prod(_power_terms(r.products))
            out.append(f"    const double q = keff * ({pf} - {pr} * exp({gs}){_refc_factor(dnu)});")
        else:
            out.append(f"    const double q = keff * ({pf});")
        for i, c in sorted(r.reactants.items()):
            out.append(f"    wd[{i}] -= {_d(c) + ' * ' if c != 1 else ''}q;")
        for i, c in sorted(r.products.items()):
            out.append(f"    wd[{i}] += {_d(c) + ' * ' if c != 1 else ''}q;")
        out.append("  }")
    out.append("}")
    return "\n".join(out)


def gen_jacobian(mech: M.Mechanism) -> str:
    """Derivatives of f_i = W_i wdot_i with respect to rhoY_k (=W_k C_k) and T.

    Two phases, so that no Jacobian entry is accumulated in memory (a read-modify-
    write chain whose latency would dominate): per reaction j the rate of progress q_j,
    its T derivative and every distinct dq_j/dC_k go to write-once private arrays; then
    wd and dwdT are formed species by species, and the entries column by column (k
    ascending, rows i ascending, structural zeros included), each sum over reactions in
    reaction order.  acc(i, k, v) receives each entry v = sum_j W_i/W_k nu_ij dq_j/dC_k
    exactly once, and colend(k) follows the last row of column k.  dwdT is the molar
    dwdot/dT and wd the molar production rates at the same state.
    """
    ns = mech.nspec
    nr = mech.nreac
    W = mech.W
    out = []
    body = []
    nu_all = []          # per reaction: {species: net stoichiometric coefficient}
    entries: dict = {}   # (i, k) -> [(coefficient, D index)] in reaction order
    nd = 0
    if any(r.has_third_body for r in mech.reactions):
        body.append("  double Ctot = 0.0;")
        body.append(f"  for (int k = 0; k < {ns}; ++k) Ctot += C[k];")
    for j, r in enumerate(mech.reactions):
        body.append(f"  {{ // {j}: {r.equation}")
        body.append(f"    const double kf = {_kf_expr(r.A, r.beta, r.Ea / M.R_CAL)};")
        body.append(f"    const double dlnkf = {_dlnk_expr(r.beta, r.Ea / M.R_CAL)};")
        if r.kind == "elementary":
            body.append("    const double keff = kf, dkeffdT = kf * dlnkf;")
        elif r.kind == "three-body":
            body.append(f"    const double Mc = {_mc_expr(r, ns)};")
            body.append("    const double keff = kf * Mc, dkeffdT = keff * dlnkf, dkeffdM = kf;")
        else:
            A0, b0, E0 = r.low
            body.append(f"    const double Mc = {_mc_expr(r, ns)};")
            body.append(f"    const double k0 = {_kf_expr(A0, b0, E0 / M.R_CAL)};")
            body.append(f"    const double dlnk0 = {_dlnk_expr(b0, E0 / M.R_CAL)};")
            body.append("    const double Pr = k0 * Mc / kf, dPrdM = k0 / kf, dPrdT = Pr * (dlnk0 - dlnkf);")
            if r.troe is not None:
                fc, dfc = _troe_fc(r)
                body.append(f"    double Fc = {fc}, dFc = {dfc};")
                body.append("    if (Fc < 1e-300) { Fc = 1e-300; dFc = 0.0; }")
                body.append("    const double logFc = log10(Fc);")
                body.append("    const double cc = -0.4 - 0.67 * logFc, nn = 0.75 - 1.27 * logFc;")
                body.append("    const double PrS = fmax(Pr, 1e-300);")
                body.append("    const double u = log10(PrS) + cc, dd = nn - 0.14 * u, f1 = u / dd, den = 1.0 + f1 * f1;")
                body.append("    const double F = exp(2.302585092994045684 * logFc / den);")
                body.append("    const double df1du = nn / (dd * dd), df1dl = (-0.67 * dd + 1.1762 * u) / (dd * dd);")
                body.append("    const double dlFdlPr = -logFc * 2.0 * f1 / (den * den) * df1du;")
                body.append("    const double dlFdlFc = 1.0 / den - logFc * 2.0 * f1 / (den * den) * df1dl;")
                body.append("    const double dFdPr = F * dlFdlPr / PrS, dFdT = F * dlFdlFc * dFc / Fc;")
            else:
                body.append("    const double F = 1.0, dFdPr = 0.0, dFdT = 0.0;")
            body.append("    const double gg = Pr / (1.0 + Pr), dg = 1.0 / ((1.0 + Pr) * (1.0 + Pr));")
            body.append("    const double keff = kf * gg * F;")
            body.append("    const double dkeffdM = kf * (dg * F + gg * dFdPr) * dPrdM;")
            body.append("    const double dkeffdT = keff * dlnkf + kf * (dg * F + gg * dFdPr) * dPrdT + kf * gg * dFdT;")
        pf = _prod(_power_terms(r.reactants))
        body.append(f"    const double pf = {pf};")
        if r.reversible:
            gs, hs, dnu = _kc_exprs(r)
            pr = _prod(_power_terms(r.products))
            body.append(f"    const double pr = {pr};")
            body.append(f"    const double invKc = exp({gs}){_refc_factor(dnu)};")
            body.append(f"    const double dlnKc = ({hs} - {_d(dnu)}) * invT;")
            body.append("    const double net = pf - pr * invKc;")
            body.append("    const double dqdT = dkeffdT * net + keff * pr * invKc * dlnKc;")
        else:
            body.append("    const double net = pf;")
            body.append("    const double dqdT = dkeffdT * net;")
        body.append(f"    Q[{j}] = keff * net; QT[{j}] = dqdT;")
        nu: dict = {}
        for i, c in r.reactants.items():
            nu[i] = nu.get(i, 0) - c
        for i, c in r.products.items():
            nu[i] = nu.get(i, 0) + c
        nu = {i: v for i, v in nu.items() if v != 0}
        nu_all.append(nu)
        # dq/dC_k from the mass-action products
        dq: dict = {}
        for k in sorted(r.reactants):
            coef, terms = _dpower_terms(r.reactants, k)
            e = f"{_d(coef)} * {_prod(terms)}" if coef != 1.0 else _prod(terms)
            dq.setdefault(k, []).append(f"keff * {e}")
        if r.reversible:
            for k in sorted(r.products):
                coef, terms = _dpower_terms(r.products, k)
                e = f"{_d(coef)} * {_prod(terms)}" if coef != 1.0 else _prod(terms)
                dq.setdefault(k, []).append(f"- keff * invKc * {e}")
        if r.has_third_body:
            body.append("    const double tb = dkeffdM * net;")
            for k in range(ns):
                e = r.efficiencies.get(k, 1.0)
                if e != 0.0:
                    dq.setdefault(k, []).append(f"{_d(e)} * tb" if e != 1.0 else "tb")
        for k in sorted(dq):
            expr = " + ".join(dq[k]).replace("+ -", "-")
            body.append(f"    D[{nd}] = {expr};")
            for i, v in sorted(nu.items()):
                entries.setdefault((i, k), []).append((v * W[i] / W[k], nd))
            nd += 1
        body.append("  }")
    out.append("template <class Acc, class ColEnd>")
    out.append("static __device__ __forceinline__ void jacobian(const double* __restrict__ C, const double T, const double logT, "
               "const double invT, const double* __restrict__ gRT, const double* __restrict__ hRT, "
               "double* __restrict__ wd, double* __restrict__ dwdT, Acc&& acc, ColEnd&& colend) {")
    out.append(f"  const double refC = {_d(M.P_ATM / M.R_CGS)} * invT, refCinv = 1.0 / refC;")
    out.append(f"  double Q[{nr}], QT[{nr}], D[{max(nd, 1)}];")
    out += body
    # production rates and their T derivatives, species by species
    for i in range(ns):
        terms = [(nu_all[j][i], j) for j in range(nr) if i in nu_all[j]]
        if not terms:
            out.append(f"  wd[{i}] = 0.0; dwdT[{i}] = 0.0;")
            continue
        v, j = terms[0]
        out.append(f"  {{ double s = {_d(v)} * Q[{j}], t = {_d(v)} * QT[{j}];")
        for v, j in terms[1:]:
            out.append(f"    s += {_d(v)} * Q[{j}]; t += {_d(v)} * QT[{j}];")
        out.append(f"    wd[{i}] = s; dwdT[{i}] = t; }}")
    # Jacobian entries, column by column
    for k in range(ns):
        for i in range(ns):
            terms = entries.get((i, k))
            if not terms:
                out.append(f"  acc({i}, {k}, 0.0);")
                continue
            c, m = terms[0]
            out.append(f"  {{ double s = {_d(c)} * D[{m}];")
            for c, m in terms[1:]:
                out.append(f"    s += {_d(c)} * D[{m}];")
            out.append(f"    acc({i}, {k}, s); }}")
        out.append(f"  colend({k});")
    out.append("}")
    return "\n".join(out)


_TRANS = ("exp", "log10", "log", "pow", "sqrt")


def count_ops(code: str) -> dict:
    """Static operation count of generated straight-line code: every binary + - * /
    as written counts one flop (so a fused multiply-add counts two), fmax/fmin one, and
    exp / log / log10 / pow / sqrt are counted separately as transcendentals.  Loops in
    the generated code are the fixed per-species ones ("for (int k = 0; k < NS; ...)"),
    whose single-statement bodies are weighted by NS."""
    flops = 0
    trans = {t: 0 for t in _TRANS}
    ns = None
    for line in code.splitlines():
        line = line.split("//")[0]
        m = re.search(r"for \(int k = 0; k < (\d+); \+\+k\)(.*)", line)
        weight = 1
        if m:
            weight = int(m.group(1))
            line = m.group(2)
        # drop the ternary coefficient selects of the NASA polynomials
        line = re.sub(r"\(lo \? [^:]+ : [^)]+\)", "a", line)
        for t in _TRANS:
            n = len(re.findall(r"\b%s\(" % t, line))
            trans[t] += n * weight
        n = len(re.findall(r"\bfm(ax|in)\(", line))
        # binary operators: + * / anywhere; - when it follows an operand
        ops = len(re.findall(r"(?<=[\w\)\]\s])\s*[\+\*/](?!=)", line))
        ops += len(re.findall(r"(?<=[\w\)\]])\s+-\s+", line))
        ops += len(re.findall(r"[\+\-\*/]=", line))
        flops += (ops + n) * weight
    del ns
    return {"flops": flops, "transcendentals": sum(trans.values()), "by_function": trans}


def write_inc(mech: M.Mechanism, path: str) -> None:
    ns = mech.nspec
    name = mech.name
    cn = _cname(name)
    lines = [f"// GENERATED by pelelm_amd/mechanisms/hipgen.py from {name}.yaml -- do not edit.",
             f"// {ns} species, {mech.nreac} reactions.", "#pragma once", "",
             f"struct mech_{cn} {{", "",
             f"static constexpr int NS = {ns};", f"static constexpr int NR = {mech.nreac};",
             f"static constexpr int N1 = {ns + 1};",
             f"static constexpr const char* NAME = \"{name}\";",
             f"static constexpr double RU = {_d(M.R_CGS)};", f"static constexpr double PATM = {_d(M.P_ATM)};",
             "static __host__ __device__ __forceinline__ double mw(int k) {",
             "  constexpr double t[NS] = {" + ", ".join(_d(w) for w in mech.W) + "};",
             "  return t[k];", "}",
             "static __host__ __device__ __forceinline__ double imw(int k) {",
             "  constexpr double t[NS] = {" + ", ".join(_d(1.0 / w) for w in mech.W) + "};",
             "  return t[k];", "}",
             "static const char* species(int k) {",
             "  static const char* t[NS] = {" + ", ".join(f'"{s.name}"' for s in mech.species) + "};",
             "  return t[k];", "}",
             ""]
    th = gen_thermo(mech)
    wd = gen_wdot(mech)
    jac = gen_jacobian(mech)
    parts = th.split("static __device__ __forceinline__ void ")
    counts = {}
    for part in parts[1:]:
        fname = part.split("(")[0]
        counts[fname] = count_ops(part)
    counts["wdot"] = count_ops(wd)
    counts["jacobian"] = count_ops(jac)
    lines.append("// static operation counts (flops; transcendentals) of the generated functions")
    for k, v in counts.items():
        lines.append(f"static constexpr long OPS_{k.upper()} = {v['flops']};")
        lines.append(f"static constexpr long TRANS_{k.upper()} = {v['transcendentals']};")
    lines.append("")
    lines.append(th)
    lines.append("")
    lines.append(wd)
    lines.append("")
    lines.append(jac)
    lines.append("")
    lines.append(f"}};  // struct mech_{cn}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    ops_path = os.path.join(os.path.dirname(os.path.dirname(path)), "..", "mechanisms", f"{name}_ops.json")
    with open(os.path.normpath(ops_path), "w") as f:
        json.dump({"mechanism": name, "nspec": ns, "nreac": mech.nreac, "functions": counts}, f, indent=1)
