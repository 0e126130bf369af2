"""Mechanism code generator.

Turns a parsed mechanism (``mechanism.py``) into

* ``<name>.tbl`` -- a flat whitespace-separated numeric table (species, NASA-7
  thermo, reactions with integer stoichiometry, third-body efficiencies, fall-off
  parameters, activation temperatures in K).  Host code reads it at run time.
* ``../csrc/mech/<name>.inc`` -- straight-line HIP device code for one cell per
  lane: species thermo, net production rates and the analytic Jacobian of the
  constant-pressure reactor right-hand side, plus the constant tables the host
  side needs.  This is the role FUEGO's generated ``mechanism.cpp`` plays for
  PelePhysics (``Docs/source/manual/Model.rst:299-306``).

Run ``python -m pelelm_amd.mechanisms.codegen`` to regenerate every mechanism.
"""
from __future__ import annotations

import os
import sys

from . import mechanism as M

KIND_CODE = {"elementary": 0, "three-body": 1, "falloff": 2}


def _fmt(x: float) -> str:
    return repr(float(x))


def write_table(mech: M.Mechanism, path: str) -> None:
    """Flat table.  Line formats (one record per line):

    ``mech <name> <nelem> <nspec> <nreac> <R_cgs> <P_atm>``
    ``elem <name>`` x nelem
    ``spec <name> <W> <Tlo> <Tmid> <Thi> <low x7> <high x7> <elem counts x nelem>``
    ``reac <kind 0/1/2> <rev> <A> <beta> <Ta[K]> <nr> (idx coef)*nr <np> (idx coef)*np
           <A0> <beta0> <Ta0> <ntroe> <troe x ntroe> <neff> (idx eff)*neff``

    kind: 0 elementary, 1 three-body, 2 fall-off (Lindemann when ntroe == 0, Troe otherwise).
    """
    lines = [f"mech {mech.name} {len(mech.elements)} {mech.nspec} {mech.nreac} "
             f"{_fmt(M.R_CGS)} {_fmt(M.P_ATM)}"]
    for e in mech.elements:
        lines.append(f"elem {e}")
    for s in mech.species:
        ec = " ".join(str(s.composition.get(e, 0)) for e in mech.elements)
        lines.append("spec " + " ".join([s.name, _fmt(s.W), _fmt(s.T[0]), _fmt(s.T[1]), _fmt(s.T[2])]
                                          + [_fmt(a) for a in s.low] + [_fmt(a) for a in s.high]) + " " + ec)
    for r in mech.reactions:
        parts = [str(KIND_CODE[r.kind]), str(int(r.reversible)), _fmt(r.A), _fmt(r.beta), _fmt(r.Ea / M.R_CAL)]
        parts.append(str(len(r.reactants)))
        for i, c in sorted(r.reactants.items()):
            parts += [str(i), str(c)]
        parts.append(str(len(r.products)))
        for i, c in sorted(r.products.items()):
            parts += [str(i), str(c)]
        if r.low is not None:
            parts += [_fmt(r.low[0]), _fmt(r.low[1]), _fmt(r.low[2] / M.R_CAL)]
        else:
            parts += ["0.0", "0.0", "0.0"]
        troe = list(r.troe) if r.troe is not None else []
        parts.append(str(len(troe)))
        parts += [_fmt(t) for t in troe]
        eff = sorted(r.efficiencies.items()) if r.has_third_body else []
        parts.append(str(len(eff)))
        for i, e in eff:
            parts += [str(i), _fmt(e)]
        lines.append("reac " + " ".join(parts))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main(argv=None) -> int:
    from . import hipgen
    names = argv or ["LiDryer", "drm19", "dodecane53", "dodecane56", "heptane88"]
    here = M.HERE
    incdir = os.path.join(os.path.dirname(here), "csrc", "mech")
    os.makedirs(incdir, exist_ok=True)
    for n in names:
        mech = M.load(n)
        write_table(mech, os.path.join(here, f"{n}.tbl"))
        hipgen.write_inc(mech, os.path.join(incdir, f"{n}.inc"))
        print(f"generated {n}: {mech.nspec} species, {mech.nreac} reactions")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
