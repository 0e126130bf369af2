"""The reactor's device code, run serially on the CPU, against the oracle.

tests/cpu_emulation builds reactor_bdf.hip with the host compiler through a small
stand-in for the HIP runtime (one lane at a time), so the trip pipeline -- k_init,
k_eval, k_setup, k_consume and the host loop -- is checked on every CPU test run.
The device code follows the same algorithm as the oracle, so every cell must take the
same number of right-hand-side evaluations and agree to rounding.  The workspace starts
filled with NaNs: device memory holds whatever was there before, and a word read before
it is written (even multiplied by zero) must not reach a result.
"""
import ctypes as C
import glob
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from pelelm_amd import cells

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "cpu_emulation")
LIB = os.path.join(EMU, "_build", "libemu.so")
SRC = [os.path.join(EMU, "emu.cpp"), os.path.join(EMU, "hip", "hip_runtime.h"),
       os.path.join(os.path.dirname(HERE), "pelelm_amd", "csrc", "reactor_bdf.hip"),
       os.path.join(os.path.dirname(HERE), "pelelm_amd", "csrc", "reactor_internal.h"),
       os.path.join(os.path.dirname(HERE), "pelelm_amd", "csrc", "mlp_internal.h")]
SRC += sorted(glob.glob(os.path.join(os.path.dirname(HERE), "pelelm_amd", "csrc", "mech", "*.inc")))
CLANG = "/opt/rocm/llvm/bin/clang++"   # host compile; the device code uses clang vector types


@pytest.fixture(scope="module")
def emu():
    stale = not os.path.exists(LIB) or any(os.path.getmtime(s) > os.path.getmtime(LIB) for s in SRC)
    if stale:
        os.makedirs(os.path.dirname(LIB), exist_ok=True)
        subprocess.run([CLANG, "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + EMU, "-o", LIB, SRC[0]], check=True)
    L = C.CDLL(LIB)
    L.emu_react.restype = C.c_int
    # every workspace word starts as a NaN (and every list entry as -1), so a read of a
    # word the integrator has not written poisons the cell instead of passing unseen
    L.emu_set_fill(0xFF)
    return L


@pytest.mark.parametrize("wl,mech,n", [("lidryer_16", "LiDryer", 256), ("drm19_64", "drm19", 256),
                                       ("drm19_128_mix", "drm19", 256), ("dodecane53_64_mix", "dodecane53", 128),
                                       ("dodecane56_64_mix", "dodecane56", 64), ("heptane88_64_mix", "heptane88", 64)])
@pytest.mark.parametrize("compact", [1, 0])
def test_emulated_kernels_match_oracle(emu, wl, mech, n, compact):
    d = cells.generate_workload(wl, ncell=n)
    m = O.Mech(mech)
    ref = m.react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10, max_steps=20000)
    rY, T, rE = d["rY"].copy(), d["T"].copy(), d["rE"].copy()
    FC = np.zeros(n)
    cnt = np.zeros(16, dtype=np.uint64)
    trips = C.c_long(0)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    rc = emu.emu_react(mech.encode(), C.c_long(n), p(rY), p(T), p(rE), p(FC), C.c_double(d["dt"]),
                       C.c_double(1e-6), C.c_double(1e-10), 4, compact, p(cnt), C.byref(trips))
    assert rc == 0
    assert cnt[0] == n and cnt[8] == 0          # cells integrated, none failed
    np.testing.assert_array_equal(FC, ref["FC"])  # identical step sequences
    tol = 1e-6 * np.abs(ref["rY"]) + 1e-10
    assert np.max(np.abs(rY - ref["rY"]) / tol) < 1e-3
    assert np.max(np.abs(T - ref["T"])) < 1e-8
    assert trips.value >= int(ref["FC"].max())


def test_predictor_plumbing_with_zero_network(emu):
    """The learned-predictor path (iterate from the network, Nordsieck reference for the
    error test, fallback on Newton failure) with a network of zeros, whose iterate is
    the state at tn: every cell must finish, predictions and fallbacks are counted, and
    the accuracy against the tight oracle stays at the plain form's level."""
    d = cells.generate_workload("drm19_128_mix", ncell=256)
    m = O.Mech("drm19")
    ref = m.react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10)
    tight = m.react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-12, atol=1e-20, max_steps=200000)
    n = 256
    rY, T, rE = d["rY"].copy(), d["T"].copy(), d["rE"].copy()
    FC = np.zeros(n)
    cnt = np.zeros(16, dtype=np.uint64)
    trips = C.c_long(0)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    emu.emu_set_pred(1)
    try:
        rc = emu.emu_react(b"drm19", C.c_long(n), p(rY), p(T), p(rE), p(FC), C.c_double(d["dt"]),
                           C.c_double(1e-6), C.c_double(1e-10), 4, 1, p(cnt), C.byref(trips))
    finally:
        emu.emu_set_pred(0)
    assert rc == 0
    assert cnt[8] == 0                      # no failed cell
    assert cnt[13] > 0 and cnt[14] > 0      # predictions used, fallbacks taken
    e_plain = np.abs(ref["T"] - tight["T"]).max()
    assert np.abs(T - tight["T"]).max() <= 5.0 * e_plain + 0.05
    assert np.all(FC > 0)


@pytest.mark.parametrize("wl,mech,n", [("drm19_64", "drm19", 256), ("dodecane53_64_mix", "dodecane53", 64)])
def test_emulated_fp32_factors_take_the_fp64_steps(emu, wl, mech, n):
    """FP32 factors with FP64-residual refinement: on these cells the refinement brings
    every solve to the FP64 solve's accuracy, so the step sequence (FC) is the FP64
    form's and the states agree far inside the tolerance."""
    d = cells.generate_workload(wl, ncell=n)
    ref = O.Mech(mech).react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10, max_steps=20000)
    rY, T, rE = d["rY"].copy(), d["T"].copy(), d["rE"].copy()
    FC = np.zeros(n)
    cnt = np.zeros(16, dtype=np.uint64)
    trips = C.c_long(0)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    emu.emu_set_lu(1, 3)
    try:
        rc = emu.emu_react(mech.encode(), C.c_long(n), p(rY), p(T), p(rE), p(FC), C.c_double(d["dt"]),
                           C.c_double(1e-6), C.c_double(1e-10), 4, 0, p(cnt), C.byref(trips))
    finally:
        emu.emu_set_lu(0, 3)
    assert rc == 0 and cnt[8] == 0
    assert cnt[10] > 0                          # refinement sweeps ran
    assert np.mean(FC == ref["FC"]) >= 0.99
    tol = 1e-6 * np.abs(ref["rY"]) + 1e-10
    assert np.max(np.abs(rY - ref["rY"]) / tol) < 0.05
    assert np.max(np.abs(T - ref["T"])) < 1e-6


@pytest.mark.parametrize("compact", [1, 0])
def test_emulated_cost_ordering_changes_nothing_but_the_slots(emu, compact):
    """With a slot -> cell map (mi.order), every cell's result and RHS count are those of
    the unordered run: only the workspace slot a cell occupies changes."""
    n = 192
    d = cells.generate_workload("drm19_128_mix", ncell=n)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    out = []
    perm = np.random.default_rng(3).permutation(n).astype(np.int32)
    for order in (None, perm):
        rY, T, rE = d["rY"].copy(), d["T"].copy(), d["rE"].copy()
        FC = np.zeros(n)
        cnt = np.zeros(16, dtype=np.uint64)
        trips = C.c_long(0)
        emu.emu_set_order(None if order is None else p(order))
        try:
            rc = emu.emu_react(b"drm19", C.c_long(n), p(rY), p(T), p(rE), p(FC), C.c_double(d["dt"]),
                               C.c_double(1e-6), C.c_double(1e-10), 4, compact, p(cnt), C.byref(trips))
        finally:
            emu.emu_set_order(None)
        assert rc == 0
        out.append((rY, T, rE, FC))
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)


def test_emulated_cost_keys(emu):
    """The ordering keys: ORDER_RATE is max_k |W_k wdot_k| dt / (rtol rhoY_k + atol) at the
    input state (T from HY2T), computed here from the oracle's rates; ORDER_FC is |FC|."""
    n = 64
    d = cells.generate_workload("drm19_128_mix", ncell=n)
    m = O.Mech("drm19")
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    rY, T, rE = d["rY"].copy(), d["T"].copy() + 5.0, d["rE"].copy()
    FC = -np.arange(n, dtype=np.float64)
    keys = np.zeros(n, dtype=np.uint32)
    vals = np.zeros(n, dtype=np.int32)
    rtol, atol = 1e-6, 1e-10
    assert emu.emu_cost_keys(b"drm19", C.c_long(n), p(rY), p(T), p(rE), p(FC), C.c_double(d["dt"]),
                             C.c_double(rtol), C.c_double(atol), 1, p(keys), p(vals)) == 0
    got = keys.view(np.float32)
    np.testing.assert_array_equal(vals, np.arange(n))
    for c in range(n):
        rho = rY[:, c].sum()
        Tc, _ = m.hy2t(rE[c] / rho, rY[:, c] / rho, T[c])
        w = m.wdot(rY[:, c] / m.W, Tc) * m.W
        ref = np.max(np.abs(w) * d["dt"] / (rtol * np.abs(rY[:, c]) + atol))
        assert abs(got[c] - ref) <= 1e-5 * ref
    assert emu.emu_cost_keys(b"drm19", C.c_long(n), p(rY), p(T), p(rE), p(FC), C.c_double(d["dt"]),
                             C.c_double(rtol), C.c_double(atol), 2, p(keys), p(vals)) == 0
    np.testing.assert_array_equal(keys.view(np.float32), np.arange(n, dtype=np.float32))


@pytest.mark.parametrize("wl,mech,n", [("lidryer_16", "LiDryer", 128), ("drm19_64", "drm19", 64),
                                       ("drm19_128_mix", "drm19", 64)])
def test_emulated_rk64_matches_oracle(emu, wl, mech, n):
    """ReactorRK64's kernel against the oracle's restatement: the same sub-steps (RHS count
    per cell) and states to rounding where the chemistry is not stiff.  Stiff flame-front
    cells (dt = 1e-5 s) run at the explicit stability limit, where the accept/reject
    sequence of a few hundred sub-steps follows rounding (the generated device RHS and the
    oracle's table-driven one differ in the last bits): there the same cells fail at the
    budget of 500 sub-steps, the counts agree to 5 % and the states to the method's
    accuracy."""
    d = cells.generate_workload(wl, ncell=n)
    m = O.Mech(mech)
    rng = np.random.default_rng(8)
    fY = np.ascontiguousarray(rng.normal(size=d["rY"].shape) * 1e-3 * d["rY"] / d["dt"])
    fH = np.ascontiguousarray(rng.normal(size=n) * 1e-3 * np.abs(d["rE"]) / d["dt"])
    mask = np.where(rng.random(n) < 0.1, -1, 1).astype(np.int32)
    ref = m.rk64_react(d["rY"], d["T"], d["rE"], d["dt"], rYsrc=fY, rEsrc=fH, mask=mask, nsub=(10, 5, 500))
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    rY, T, rE, FC = d["rY"].copy(), d["T"].copy(), d["rE"].copy(), np.zeros(n)
    cnt = np.zeros(16, dtype=np.uint64)
    emu.emu_set_forcing(p(fY), p(fH), p(mask))
    try:
        rc = emu.emu_rk64(mech.encode(), C.c_long(n), p(rY), p(T), p(rE), p(FC), C.c_double(d["dt"]),
                          C.c_double(1e-6), C.c_double(1e-10), 10, 5, 500, p(cnt))
    finally:
        emu.emu_set_forcing(None, None, None)
    assert rc == 0
    assert cnt[8] == ref["nfail"] and cnt[0] == (mask >= 0).sum()
    np.testing.assert_array_equal(FC < 0, ref["FC"] < 0)
    tol = 1e-6 * np.abs(ref["rY"]) + 1e-10
    if wl != "drm19_128_mix":
        np.testing.assert_array_equal(FC, ref["FC"])
        assert np.max(np.abs(rY - ref["rY"]) / tol) < 1e-6
        assert np.max(np.abs(T - ref["T"])) < 1e-9
    else:
        assert 0 < (FC < 0).sum() < n
        ok = FC > 0
        assert np.max(np.abs(FC[ok] - ref["FC"][ok]) / ref["FC"][ok]) < 0.05
        assert np.max(np.abs(rY - ref["rY"])[:, ok] / tol[:, ok]) < 10.0


@pytest.mark.parametrize("wl,mech,n,maxl", [("lidryer_16", "LiDryer", 256, 5), ("drm19_128_mix", "drm19", 256, 5),
                                            ("dodecane53_64_mix", "dodecane53", 64, 53)])
def test_emulated_gmres_kernels_match_oracle(emu, wl, mech, n, maxl):
    """cvode.solve_type = GMRES: the trip kernels with the GMRES solve kernel (Krylov
    vectors in the workspace, J v by difference quotients of the right-hand side)
    against the GMRES oracle.  LiDryer and drm19 take the oracle's steps in every cell
    with its Krylov iteration count.  On the 53-species stiff mix the difference
    quotients make the method itself sensitive to last-bit differences of the right-
    hand side: the oracle against itself with sig scaled by 1 + 1e-13 keeps the step
    sequence in 53 of these 64 cells and differs by up to 3.7 tolerance units and
    0.008 K; the kernels are held to that level, and to the oracle's accuracy against
    the tight reference."""
    d = cells.generate_workload(wl, ncell=n)
    m = O.Mech(mech)
    ref = m.react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10, max_steps=20000, stats=True,
                  solve="gmres", maxl=maxl)
    assert ref["nfail"] == 0
    rY, T, rE = d["rY"].copy(), d["T"].copy(), d["rE"].copy()
    FC = np.zeros(n)
    cnt = np.zeros(16, dtype=np.uint64)
    trips = C.c_long(0)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    emu.emu_set_gmres(1, maxl)
    try:
        rc = emu.emu_react(mech.encode(), C.c_long(n), p(rY), p(T), p(rE), p(FC), C.c_double(d["dt"]),
                           C.c_double(1e-6), C.c_double(1e-10), 4, 0, p(cnt), C.byref(trips))
    finally:
        emu.emu_set_gmres(0, 5)
    assert rc == 0
    assert cnt[0] == n and cnt[8] == 0
    assert cnt[3] == 0 and cnt[4] == 0                # no Jacobian, no set-up
    tol = 1e-6 * np.abs(ref["rY"]) + 1e-10
    ey, et = np.max(np.abs(rY - ref["rY"]) / tol), np.max(np.abs(T - ref["T"]))
    same = np.mean(FC == ref["FC"])
    print(f"{wl} GMRES({maxl}): identical FC {same:.3f}, max |dY|/tol {ey:.3g}, max |dT| {et:.3g} K")
    if mech != "dodecane53":
        np.testing.assert_array_equal(FC, ref["FC"])
        assert cnt[15] == ref["stats"]["nli"].sum()   # Krylov iterations
        # the difference quotients divide the right-hand sides' last-bit differences
        # (the generated rates against the oracle's) by sig ~ 1 / ||v||: T agrees to
        # 1e-6 K here, against 1e-8 K with the direct solve
        assert ey < 1e-3 and et < 1e-6
    else:
        tight = m.react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-12, atol=1e-20, max_steps=200000)
        assert same >= 0.75 and ey < 20.0 and et < 0.05
        e_k, e_o = np.abs(T - tight["T"]).max(), np.abs(ref["T"] - tight["T"]).max()
        assert e_k <= 2.0 * e_o + 1e-3, (e_k, e_o)
