"""Chem-advance throughput of the MI355X reactor (one JSON line on stdout).

Workload: drm19 (21 species) on a 128^3 box -- 2,097,152 independent cells at 1 atm
with the 60 % cold / 30 % near-equilibrium / 10 % flame-front stiffness mix, spatially
shuffled, advanced over dt = 1e-5 s at rtol 1e-6, atol 1e-10 (seeded Philox inputs,
seed 104).  A step is one ReactorBase::react call over the whole box with its inputs
already resident in HBM (a device-to-device restore of the inputs precedes every call
and is inside the timed region; it moves 0.4 GB, < 1 % of a step).

Forms measured (each is a complete, independently parity-tested reactor):
  plain  -- FP64 BDF, fixed cell-to-lane mapping: finished cells idle until the box is done
  queue  -- FP64 BDF, compacted work lists: finished cells leave the pipeline
`value` is the fastest form's cells/s.  The CPU baseline is this project's CPU oracle
(CVODE-algorithm restatement) on a bounded sample of the same cells, on the host
threads OMP_NUM_THREADS allows.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "chem-advance cells/s, drm19 128³, rtol 1e-6: plain FP64 vs MLP+FP32-LU"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def rhs_ops(mech: str):
    """Algorithmic flops of one right-hand-side evaluation (fixed part, per HY2T
    iteration) and its transcendentals, from the generator's static counts."""
    with open(os.path.join(ROOT, "pelelm_amd", "mechanisms", f"{mech}_ops.json")) as f:
        d = json.load(f)["functions"]
    N = json.load(open(os.path.join(ROOT, "pelelm_amd", "mechanisms", f"{mech}_ops.json")))["nspec"]
    cph = d["thermo_cph"]["flops"]
    # rho, 1/rho, Y; after HY2T: log T, 1/T, C, gibbs, wdot, cp/h, f, rho cp, h.wdot, fT
    fixed = (2 * N + 1) + 2 + N + d["gibbs"]["flops"] + d["wdot"]["flops"] + cph + 7 * N + 5
    per_iter = cph + 6 * N + 5
    trans = 1 + d["wdot"]["transcendentals"]
    return fixed, per_iter, trans


def rhs_bytes(N: int) -> int:
    """HBM bytes one k_eval cell-evaluation moves: zn0, zn1, acor, forcing, the lane
    state words read (96 max, 66 used) and y, f, T written."""
    N1 = N + 1
    return 8 * (3 * N1 + N + 66 + 2 * N1 + 1)


def run_form(form, dev_in, ncell, nspec, dt, steps, warmup, profile):
    import torch
    from pelelm_amd import reactor
    r = reactor.ReactorBase.create("ReactorCvode")
    r.set("mi.mechanism", "drm19")
    r.set("ode.rtol", 1e-6)
    r.set("ode.atol", 1e-10)
    r.set("cvode.max_order", 4)
    r.set("mi.schedule", form)
    r.set("mi.profile", 1 if profile else 0)
    r.init(2, ncell)
    rY = torch.empty_like(dev_in["rY"])
    T = torch.empty_like(dev_in["T"])
    rE = torch.empty_like(dev_in["rE"])
    FC = torch.zeros(ncell, dtype=torch.float64, device="cuda")

    def step():
        rY.copy_(dev_in["rY"])
        T.copy_(dev_in["T"])
        rE.copy_(dev_in["rE"])
        r.react(rY, None, T, rE, None, FC, None, dt, 0.0)

    for w in range(warmup):
        step()
        torch.cuda.synchronize()
        log(f"{form}: warmup {w + 1}/{warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for k in range(steps):
        step()
        stats.append(r.stats())   # synchronises the call
        log(f"{form}: step {k + 1}/{steps}  {stats[-1]['ms_kernel']:.1f} ms device, {stats[-1]['ntrips']} trips")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    r.close()
    ms = (t1 - t0) * 1e3 / steps
    out = {"cells_per_s": ncell / (ms * 1e-3), "ms_per_step": ms, "stats": stats[-1],
           "nfail": max(s["nfail"] for s in stats),
           "result": {"rY": rY, "T": T, "rE": rE, "FC": FC}}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="drm19_128_mix")
    ap.add_argument("--ncell", type=int, default=None, help="override the cell count (default: the workload's)")
    ap.add_argument("--forms", default="plain,queue")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline sample duration")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="skip per-kernel event timing")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the reactor's cells are independent, so a multi-GPU run is "
              f"N copies of this one on disjoint boxes with no communication; --gpus {args.gpus} is not run.",
              file=sys.stderr)
        return 2

    import torch
    from pelelm_amd import cells
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    t0 = time.time()
    d = cells.generate_workload(args.workload, ncell=args.ncell)
    ncell = d["T"].shape[0]
    nspec = d["rY"].shape[0]
    log(f"generated {ncell} cells of {args.workload} in {time.time() - t0:.1f}s")
    dev_in = {k: torch.tensor(d[k], device="cuda") for k in ("rY", "T", "rE")}

    forms = {}
    for form in args.forms.split(","):
        forms[form] = run_form(form, dev_in, ncell, nspec, d["dt"], args.steps, args.warmup,
                               profile=not args.no_profile)
        log(f"{form}: {forms[form]['cells_per_s']:.4g} cells/s, {forms[form]['ms_per_step']:.1f} ms/step")
    best = max(forms, key=lambda f: forms[f]["cells_per_s"])
    bf = forms[best]
    st = bf["stats"]

    # roofline of the dominant kernel (the right-hand-side kernel when profiled)
    roofline = None
    if not args.no_profile:
        kinds = ["rhs (k_eval)", "jacobian (k_setup)", "newton/step (k_consume)", "LU (k_lu_coop)"]
        kms = st["kernel_ms"]
        dom = int(np.argmax(kms))
        fixed, per_iter, trans = rhs_ops("drm19")
        evals = st["kernel_cells"][0]
        flops = evals * fixed + st["nhy2t_iters"] * per_iter
        per_launch_flops = flops / max(1, st["kernel_launches"][0])
        avg_ms = kms[0] / max(1, st["kernel_launches"][0])
        achieved = per_launch_flops / (avg_ms * 1e-3) / 1e12
        roofline = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                    "kernel": "k_eval<mech_drm19> (right-hand side)",
                    "dominant_kernel": kinds[dom],
                    "kernel_ms": {kinds[i]: kms[i] for i in range(4)},
                    "kernel_launches": {kinds[i]: st["kernel_launches"][i] for i in range(4)},
                    "avg_launch_ms_rhs": avg_ms,
                    "rhs_flops_per_eval_fixed": fixed, "rhs_flops_per_hy2t_iter": per_iter,
                    "rhs_transcendentals_per_eval": trans,
                    "rhs_evals": evals, "hy2t_iters": st["nhy2t_iters"],
                    "rhs_transcendentals_per_s": evals * trans / (kms[0] * 1e-3),
                    "rhs_hbm_bytes_per_eval": rhs_bytes(nspec),
                    "rhs_hbm_GBs": evals * rhs_bytes(nspec) / (kms[0] * 1e-3) / 1e9}

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        m = O.Mech(d["mech"])
        nthr = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        n0 = min(ncell, 16384)
        t0 = time.perf_counter()
        ref = m.react(d["rY"][:, :n0], d["T"][:n0], d["rE"][:n0], d["dt"], rtol=1e-6, atol=1e-10,
                      max_steps=20000, nthreads=nthr)
        rate = n0 / (time.perf_counter() - t0)
        n1 = int(min(ncell, max(n0, rate * args.cpu_seconds)))
        if n1 > n0:
            t0 = time.perf_counter()
            ref = m.react(d["rY"][:, :n1], d["T"][:n1], d["rE"][:n1], d["dt"], rtol=1e-6, atol=1e-10,
                          max_steps=20000, nthreads=nthr)
            rate = n1 / (time.perf_counter() - t0)
        else:
            n1 = n0
        log(f"cpu oracle: {n1} cells at {rate:.4g} cells/s on {nthr} threads")
        gT = bf["result"]["T"][:n1].cpu().numpy()
        gY = bf["result"]["rY"][:, :n1].cpu().numpy()
        tol = 1e-6 * np.abs(ref["rY"]) + 1e-10
        cpu = {"value": rate, "unit": "cells/s", "cores": nthr, "kind": "port",
               "sample": f"first {n1} cells of the same {args.workload} box (same seed, dt, tolerances); "
                         f"CVODE-algorithm FP64 BDF restatement, OpenMP dynamic schedule",
               "gpu_vs_cpu_max_abs_dT_K": float(np.abs(gT - ref["T"]).max()),
               "gpu_vs_cpu_max_rel_err_over_tol": float((np.abs(gY - ref["rY"]) / tol).max())}

    line = {
        "metric": METRIC, "value": bf["cells_per_s"], "unit": "cells/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": bf["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded Philox cells, fresh CH4/air (phi U[0.6,1.2]) blended with HP-equilibrium products",
        "config": {"workload": args.workload, "mechanism": "drm19", "ncell": ncell, "box": "128^3" if ncell == 128 ** 3
                   else f"{ncell} cells", "dt": d["dt"], "rtol": 1e-6, "atol": 1e-10, "max_order": 4,
                   "form": best, "integrator": "BDF (CVODE algorithm), dense analytic-Jacobian LU, FP64"},
        "forms": {f: {"cells_per_s": v["cells_per_s"], "ms_per_step": v["ms_per_step"], "nfail": v["nfail"],
                      "steps_per_cell": v["stats"]["nsteps"] / max(1, v["stats"]["ncells"]),
                      "rhs_per_cell": v["stats"]["nfe"] / max(1, v["stats"]["ncells"]),
                      "lu_per_cell": v["stats"]["nsetups"] / max(1, v["stats"]["ncells"]),
                      "trips": v["stats"]["ntrips"], "device_ms": v["stats"]["ms_kernel"]}
                  for f, v in forms.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
