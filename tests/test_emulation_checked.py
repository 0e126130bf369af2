"""The reactor's device code, run serially on the CPU under AddressSanitizer and
UndefinedBehaviorSanitizer, with every workspace slot, matrix-region word, packed PERM
int and FP32 factor index checked against its cell's extent (PELE_CHECK_BOUNDS in
reactor_bdf.hip) and the workspace filled with NaNs before the run.

A sanitised runtime must come first in its process, so the emulation is built as a
stand-alone executable (cpu_emulation/emu_check.cpp) and driven through files.  The
one-lane LU and solves run in FP64 and in FP32 with FP64 refinement, and GMRES, on
every mechanism, including the two whose ODE size is odd (57 and 89 unknowns: the PERM
ints then fill half of their last word).  Any out-of-extent index, heap overflow,
use-after-free or undefined operation aborts the executable; the results must still
match the oracle as in test_kernel_emulation.py.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from pelelm_amd import cells

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "cpu_emulation")
EXE = os.path.join(EMU, "_build", "emu_check")
SRC = [os.path.join(EMU, "emu_check.cpp"), os.path.join(EMU, "emu.cpp"), os.path.join(EMU, "hip", "hip_runtime.h"),
       os.path.join(os.path.dirname(HERE), "pelelm_amd", "csrc", "reactor_bdf.hip"),
       os.path.join(os.path.dirname(HERE), "pelelm_amd", "csrc", "reactor_internal.h"),
       os.path.join(os.path.dirname(HERE), "pelelm_amd", "csrc", "mlp_internal.h")]
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.fixture(scope="module")
def exe():
    stale = not os.path.exists(EXE) or any(os.path.getmtime(s) > os.path.getmtime(EXE) for s in SRC)
    if stale:
        os.makedirs(os.path.dirname(EXE), exist_ok=True)
        # -O0 keeps the build of every mechanism's generated code under a minute
        subprocess.run([CLANG, "-O0", "-g", "-std=c++17", "-fsanitize=address,undefined",
                        "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-DPELE_CHECK_BOUNDS",
                        "-I" + EMU, "-o", EXE, SRC[0]], check=True)
    return EXE


def _run(exe, tmp_path, wl, mech, n, lu, gmres, maxl=5, max_order=5):
    d = cells.generate_workload(wl, ncell=n)
    ns = d["rY"].shape[0]
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(inp, "wb") as f:
        for a in (d["rY"], d["T"], d["rE"]):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    r = subprocess.run([exe, mech, str(n), repr(float(d["dt"])), "1e-6", "1e-10", str(lu), "3", str(gmres),
                        str(maxl), inp, out], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    b = np.fromfile(out, dtype=np.float64, count=ns * n + 3 * n)
    cnt = np.fromfile(out, dtype=np.uint64, count=16, offset=8 * (ns * n + 3 * n))
    res = {"rY": b[:ns * n].reshape(ns, n), "T": b[ns * n:ns * n + n], "FC": b[ns * n + 2 * n:], "cnt": cnt}
    ref = O.Mech(mech).react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10, max_steps=20000,
                             max_order=max_order, **({"solve": "gmres", "maxl": maxl} if gmres else {}))
    return res, ref


@pytest.mark.parametrize("wl,mech,n,lu,gmres", [
    ("lidryer_16", "LiDryer", 64, 1, 0),
    ("drm19_128_mix", "drm19", 128, 0, 0),
    ("drm19_128_mix", "drm19", 128, 1, 0),
    ("drm19_128_mix", "drm19", 64, 0, 1),
    ("dodecane53_64_mix", "dodecane53", 48, 0, 0),
    ("dodecane53_64_mix", "dodecane53", 48, 1, 0),
    ("dodecane56_64_mix", "dodecane56", 24, 1, 0),
    ("heptane88_64_mix", "heptane88", 16, 0, 0),
    ("heptane88_64_mix", "heptane88", 16, 1, 0),
])
def test_sanitised_bounds_checked_emulation_matches_oracle(exe, tmp_path, wl, mech, n, lu, gmres):
    res, ref = _run(exe, tmp_path, wl, mech, n, lu, gmres)
    assert res["cnt"][0] == n and res["cnt"][8] == 0       # every cell integrated, none failed
    np.testing.assert_array_equal(res["FC"], ref["FC"])
    tol = 1e-6 * np.abs(ref["rY"]) + 1e-10
    # FP32 factors with refinement: the solves differ from FP64 below the refinement test
    assert np.max(np.abs(res["rY"] - ref["rY"]) / tol) < (1e-3 if lu == 0 else 1e-1)
    assert np.max(np.abs(res["T"] - ref["T"])) < (1e-8 if lu == 0 else 1e-6)
