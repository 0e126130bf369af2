"""Golden vectors for the net production rates, computed independently of the oracle.

An mpmath (50-digit) evaluation of the rate law of Docs/source/manual/Model.rst:227-264
straight from the mechanism YAML (not from the flat table the oracle reads), at seeded
states spanning 300-2500 K and lean-to-rich compositions.  Writes wdot_<mech>.json.
Run: python tests/golden/make_wdot_golden.py [mechanism ...]
"""
import json
import os
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from pelelm_amd.mechanisms import mechanism as M  # noqa: E402

mp.mp.dps = 50


def nasa(sp, T):
    a = [mp.mpf(x) for x in (sp.low if T < sp.T[1] else sp.high)]
    cp = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])))
    h = a[0] + T * (a[1] / 2 + T * (a[2] / 3 + T * (a[3] / 4 + T * a[4] / 5))) + a[5] / T
    s = a[0] * mp.log(T) + T * (a[1] + T * (a[2] / 2 + T * (a[3] / 3 + T * a[4] / 4))) + a[6]
    return cp, h, s


def wdot(mech, C, T):
    T = mp.mpf(T)
    C = [mp.mpf(c) for c in C]
    g = []
    for sp in mech.species:
        _, h, s = nasa(sp, T)
        g.append(h - s)
    R = mp.mpf(M.R_CGS)
    Rc = R / mp.mpf(4.184e7)
    refC = mp.mpf(M.P_ATM) / (R * T)
    w = [mp.mpf(0)] * mech.nspec
    for r in mech.reactions:
        kf = mp.mpf(r.A) * T ** mp.mpf(r.beta) * mp.exp(-mp.mpf(r.Ea) / (Rc * T))
        if r.has_third_body:
            Mc = sum((mp.mpf(r.efficiencies.get(k, 1.0)) * C[k] for k in range(mech.nspec)), mp.mpf(0))
        if r.kind == "elementary":
            k = kf
        elif r.kind == "three-body":
            k = kf * Mc
        else:
            A0, b0, E0 = r.low
            k0 = mp.mpf(A0) * T ** mp.mpf(b0) * mp.exp(-mp.mpf(E0) / (Rc * T))
            Pr = k0 * Mc / kf
            F = mp.mpf(1)
            if r.troe is not None:
                a, T3, T1 = (mp.mpf(x) for x in r.troe[:3])
                Fc = (1 - a) * mp.exp(-T / T3) + a * mp.exp(-T / T1)
                if len(r.troe) == 4:
                    Fc += mp.exp(-mp.mpf(r.troe[3]) / T)
                lFc = mp.log10(Fc)
                c = -mp.mpf("0.4") - mp.mpf("0.67") * lFc
                n = mp.mpf("0.75") - mp.mpf("1.27") * lFc
                u = mp.log10(Pr) + c
                f1 = u / (n - mp.mpf("0.14") * u)
                F = mp.power(10, lFc / (1 + f1 * f1))
            k = kf * Pr / (1 + Pr) * F
        pf = mp.mpf(1)
        for i, nu in r.reactants.items():
            pf *= C[i] ** nu
        q = k * pf
        if r.reversible:
            pr = mp.mpf(1)
            for i, nu in r.products.items():
                pr *= C[i] ** nu
            dG = sum(nu * g[i] for i, nu in r.products.items()) - sum(nu * g[i] for i, nu in r.reactants.items())
            dnu = sum(r.products.values()) - sum(r.reactants.values())
            q -= k * pr * mp.exp(dG) * refC ** (-dnu)
        for i, nu in r.reactants.items():
            w[i] -= nu * q
        for i, nu in r.products.items():
            w[i] += nu * q
    return [float(x) for x in w]


def main():
    out = {}
    only = set(sys.argv[1:])
    for name, nstate in (("LiDryer", 12), ("drm19", 12), ("dodecane53", 12), ("dodecane56", 8), ("heptane88", 8)):
        if only and name not in only:
            continue
        mech = M.load(name)
        rng = np.random.default_rng(2024)
        W = np.array(mech.W)
        states = []
        for i in range(nstate):
            T = float(rng.uniform(300.0, 2500.0))
            Y = rng.random(mech.nspec) ** 4 + 1e-7
            Y /= Y.sum()
            rho = M.P_ATM / (M.R_CGS * T * float(np.sum(Y / W)))
            C = (rho * Y / W).tolist()
            states.append({"T": T, "C": C, "wdot": wdot(mech, C, T)})
        out[name] = states
        with open(os.path.join(HERE, f"wdot_{name}.json"), "w") as f:
            json.dump({"mechanism": name, "states": states}, f)
        print(name, "done")


if __name__ == "__main__":
    main()
