"""Device components of the reactor on the GPU against the same code on the CPU.

tests/gpu_unit/unit_kernels.hip wraps the reactor's device functions (right-hand side
with HY2T, analytic Jacobian, LU factor/solve of I - gamma J -- one lane per cell and
the cooperative kernels (32 lanes per cell up to 32 unknowns, a whole wave for the
53-species mechanism) -- BDF coefficients and Nordsieck order
changes) in bounded kernels; libunit_gpu.so is that file built for
gfx950, libunit_cpu.so the same file built for the host through the serial stand-in.
The CPU twin is itself checked against the oracle in test_kernel_emulation.py.
Jacobian/LU calls run in divergent lanes (lane % 3 != 0) while every lane carries a
live vector across the branch, which must come back intact.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as O
from pelelm_amd import cells

HERE = os.path.dirname(os.path.abspath(__file__))
UNIT = os.path.join(HERE, "gpu_unit")


def _load(name):
    path = os.path.join(UNIT, name)
    if not os.path.exists(path):
        pytest.skip(f"{path} not built")
    return C.CDLL(path)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _inputs(mech, wl, n, seed=3):
    d = cells.generate_workload(wl, ncell=n)
    m = O.Mech(mech)
    N = m.nspec
    N1 = N + 1
    rng = np.random.default_rng(seed)
    y = np.ascontiguousarray(np.concatenate([d["rY"].T, d["T"][:, None] + rng.normal(0, 5, n)[:, None]], axis=1))
    t = rng.uniform(0, 1e-6, n)
    rE0 = d["rE"].copy()
    fH = rng.normal(0, 1e7, n)
    fy = np.ascontiguousarray(rng.normal(0, 1e-4, (n, N)) * d["rY"].T)
    return m, N1, y, t, rE0, fH, fy


def run_rhs(L, mech, y, t, rE0, fH, fy):
    n, N1 = y.shape
    f = np.zeros((n, N1))
    T = np.zeros(n)
    assert L.unit_rhs(mech.encode(), C.c_long(n), _p(y), _p(t), _p(rE0), _p(fH), _p(fy), _p(f), _p(T)) == 0
    return f, T


def run_jac_lu(L, mech, y, T, fT, gamma, b):
    n, N1 = y.shape
    J = np.zeros((n, N1, N1))
    x = np.zeros((n, N1))
    keep = np.zeros(n * N1 + 4 * n)
    sing = np.zeros(n, dtype=np.int32)
    assert L.unit_jac_lu(mech.encode(), C.c_long(n), _p(y), _p(T), _p(fT), _p(gamma), _p(b), _p(J), _p(x),
                         _p(keep), _p(sing)) == 0
    return J, x, keep, sing


def _dump(name, **arrays):
    """Keep the compared arrays for inspection when PELE_UNIT_DUMP names a directory."""
    out = os.environ.get("PELE_UNIT_DUMP")
    if out and os.path.isdir(out):
        np.savez(os.path.join(out, f"unit_{name}.npz"), **arrays)


def coeff_inputs(n, N1, seed=5):
    rng = np.random.default_rng(seed)
    scal = np.zeros((n, 13))
    scal[:, 0] = rng.integers(1, 5, n)          # q in 1..4
    scal[:, 1] = rng.integers(0, 4, n)          # qwait
    scal[:, 2] = rng.choice([0, 3, 25], n)      # nst
    scal[:, 3] = 4                              # qmax
    h = 10 ** rng.uniform(-12, -7, n)
    scal[:, 4] = h
    scal[:, 5] = h * rng.uniform(0.8, 1.2, n)
    scal[:, 6] = h * rng.uniform(0.5, 2.0, n)
    scal[:, 7:13] = h[:, None] * rng.uniform(0.3, 3.0, (n, 6))
    zn = rng.normal(size=(n, 6, N1))
    return scal, zn


def run_coeffs(L, mech, scal, zn):
    n = scal.shape[0]
    N1 = zn.shape[2]
    out = np.zeros((n, 14 + 12 * N1))
    assert L.unit_coeffs(mech.encode(), C.c_long(n), _p(scal), _p(np.ascontiguousarray(zn)), _p(out)) == 0
    return out


CASES = [("LiDryer", "lidryer_16"), ("drm19", "drm19_64"), ("dodecane53", "dodecane53_64_mix")]
N1_OF = {"LiDryer": 10, "drm19": 22, "dodecane53": 54}


@pytest.mark.gpu
@pytest.mark.parametrize("mech,wl", CASES)
def test_rhs_gpu_matches_cpu(mech, wl):
    G, Cl = _load("libunit_gpu.so"), _load("libunit_cpu.so")
    m, N1, y, t, rE0, fH, fy = _inputs(mech, wl, 640)
    fg, Tg = run_rhs(G, mech, y, t, rE0, fH, fy)
    fc, Tc = run_rhs(Cl, mech, y, t, rE0, fH, fy)
    np.testing.assert_allclose(Tg, Tc, rtol=1e-13)
    scale = np.abs(fc).max(axis=1, keepdims=True)
    assert np.max(np.abs(fg - fc) / scale) < 1e-11


@pytest.mark.gpu
@pytest.mark.parametrize("mech,wl", CASES)
def test_jac_lu_gpu_matches_cpu_in_divergent_lanes(mech, wl):
    G, Cl = _load("libunit_gpu.so"), _load("libunit_cpu.so")
    m, N1, y, t, rE0, fH, fy = _inputs(mech, wl, 640)
    n = y.shape[0]
    _, T = run_rhs(Cl, mech, y, t, rE0, fH, fy)
    f, _ = run_rhs(Cl, mech, y, t, rE0, fH, fy)
    fT = f[:, -1].copy()
    rng = np.random.default_rng(11)
    gamma = 10 ** rng.uniform(-10, -6, n)
    b = rng.normal(size=(n, N1))
    Jg, xg, kg, sg = run_jac_lu(G, mech, y, T, fT, gamma, b)
    Jc, xc, kc, sc = run_jac_lu(Cl, mech, y, T, fT, gamma, b)
    _dump(f"jac_lu_{mech}", Jg=Jg, Jc=Jc, xg=xg, xc=xc, kg=kg, kc=kc, sg=sg, sc=sc)
    active = (np.arange(n) % 64) % 3 != 0
    # live state across the divergent calls: registers (1-ulp FMA contraction allowed) and
    # private arrays indexed at run time, which must be untouched
    np.testing.assert_allclose(kg[: n * N1], kc[: n * N1], rtol=1e-15)
    np.testing.assert_array_equal(kg[n * N1:], kc[n * N1:])
    np.testing.assert_array_equal(sg, sc)
    assert np.all(sg[~active] == -1) and np.all(sg[active] == 0)
    scale = np.abs(Jc[active]).max(axis=2, keepdims=True) + 1e-300
    assert np.max(np.abs(Jg[active] - Jc[active]) / scale) < 1e-11
    # near-singular systems (|x| up to 1e7) amplify rounding: compare against each solution's scale
    xs = np.abs(xc[active]).max(axis=1, keepdims=True)
    assert np.max(np.abs(xg[active] - xc[active]) / xs) < 1e-10


def run_lu(L, mech, coop, J, gamma, pend, b, fp32=0, sweeps=3, ewt=None):
    """LU of I - gamma J then the Newton solve for b (pending, regular cells) by the
    reactor's kernels; returns x, singularity word, factors, permutation and the
    refinement (sweeps, failed) pair of every cell."""
    n, N1 = b.shape
    if ewt is None:
        ewt = np.ones((n, N1))
    ewt = np.ascontiguousarray(ewt)
    x = np.zeros((n, N1))
    sing = np.zeros(n)
    lu = np.zeros((n, N1, N1))
    perm = np.zeros((n, N1), dtype=np.int32)
    ir = np.zeros((n, 2))
    rc = L.unit_lu(mech.encode(), coop, fp32, sweeps, C.c_long(n), _p(J), _p(gamma), _p(pend), _p(b), _p(ewt),
                   _p(x), _p(sing), _p(lu), _p(perm), _p(ir))
    assert rc == 0
    return x, sing, lu, perm, ir


def _lu_case(mech, wl):
    Cl = _load("libunit_cpu.so")
    m, N1, y, t, rE0, fH, fy = _inputs(mech, wl, 645)
    n = y.shape[0]
    _, T = run_rhs(Cl, mech, y, t, rE0, fH, fy)
    f, _ = run_rhs(Cl, mech, y, t, rE0, fH, fy)
    rng = np.random.default_rng(12)
    gamma = 10 ** rng.uniform(-10, -5, n)
    b = rng.normal(size=(n, N1))
    J, _, _, _ = run_jac_lu(Cl, mech, y, T, f[:, -1].copy(), gamma, b)
    J = np.ascontiguousarray(J)
    pend = (rng.random(n) < 0.8).astype(np.int32)
    return N1, n, J, gamma, pend, b


@pytest.mark.gpu
@pytest.mark.parametrize("mech,wl", CASES)
def test_cooperative_fp32_lu_matches_lane_fp32_lu(mech, wl):
    """FP32 factors from the cooperative kernel against the one-lane FP32 factorisation,
    on the GPU and on the CPU twin (each reproduces P M to FP32 accuracy), and the
    Newton solve with them: unrefined it carries the FP32 error, with up to 3 FP64-
    residual refinement sweeps it reaches the FP64 solution, in both solve kernels."""
    G, Cl = _load("libunit_gpu.so"), _load("libunit_cpu.so")
    N1, n, J, gamma, pend, b = _lu_case(mech, wl)
    M = np.eye(N1)[None] - gamma[:, None, None] * J
    xe = np.linalg.solve(M, b[..., None])[..., 0]
    xs_all = np.abs(xe).max(axis=1, keepdims=True)
    ewt = 1.0 / (1e-6 * np.abs(xe) + 1e-10 * xs_all)
    xc, sc, luc, pc, irc = run_lu(G, mech, 1, J, gamma, pend, b, fp32=1, sweeps=3, ewt=ewt)
    xl, sl, lul, pl, irl = run_lu(G, mech, 0, J, gamma, pend, b, fp32=1, sweeps=3, ewt=ewt)
    xh, sh, luh, ph, irh = run_lu(Cl, mech, 0, J, gamma, pend, b, fp32=1, sweeps=3, ewt=ewt)
    x0, _, _, _, ir0 = run_lu(G, mech, 1, J, gamma, pend, b, fp32=1, sweeps=0, ewt=ewt)
    _dump(f"lu32_{mech}", xc=xc, irc=irc, xl=xl, irl=irl, xh=xh, irh=irh, x0=x0, xe=xe, luc=luc, lul=lul, pc=pc,
          pl=pl, pend=pend)
    idle = pend == 0
    assert np.all(sc[idle] == -3.0) and np.all(luc[idle] == 7.0) and np.all(pc[idle] == -1)
    ok = pend == 1
    assert np.all(sc[ok] == 0) and np.all(sl[ok] == 0) and np.all(sh[ok] == 0)
    # FP32 rounding breaks near-ties of the pivot search differently in the two kernels
    # (more often the more columns there are); the residual checks below hold regardless
    assert np.mean(np.all(pc[ok] == pl[ok], axis=1)) > 0.97
    assert np.mean(np.all(pl[ok] == ph[ok], axis=1)) > 0.97
    same = np.all(pc == pl, axis=1) & ok
    scale = np.abs(lul[same]).max(axis=2, keepdims=True)
    assert np.max(np.abs(luc[same] - lul[same]) / scale) < 1e-5
    Mo = M[ok]
    # each factorisation reproduces P M to FP32 accuracy (the GPU contracts to FMA, the
    # CPU twin does not, so their factors differ where elimination cancels)
    for lu, pm in ((luc, pc), (lul, pl), (luh, ph)):
        L = np.tril(lu[ok], -1) + np.eye(N1)[None]
        U = np.triu(lu[ok])
        PM = np.take_along_axis(Mo, pm[ok][:, :, None].astype(np.int64), axis=1)
        res = np.abs(PM - L @ U).max(axis=(1, 2)) / np.abs(Mo).max(axis=(1, 2))
        assert res.max() < 1e-5, res.max()
    xs = xs_all[ok]
    # unrefined: the FP32 error, at most a modest multiple of kappa(M) u_32 where the
    # system is ill-conditioned
    assert np.all(ir0[ok, 0] == 0) and np.all(ir0[ok, 1] == 0)
    rel0 = np.abs(x0[ok] - xe[ok]).max(axis=1) / xs[:, 0]
    assert np.all(rel0 < np.maximum(1e-3, 100.0 * np.linalg.cond(Mo) * 2.0 ** -24))
    # refined: converged (flag 0) to the FP64 solution; a cell whose refinement does not
    # converge in 3 sweeps is flagged (the reactor refactors it in FP64) -- rare, and
    # only in the 53-species systems, whose conditioning reaches the FP32 limit
    for x, ir in ((xc, irc), (xl, irl), (xh, irh)):
        conv = ir[ok, 1] == 0
        assert conv.mean() > (0.99 if N1 > 32 else 1.0 - 1e-12)
        assert np.all(ir[ok, 0][conv] >= 1) and np.all(ir[ok, 0] <= 3) and np.all(ir[ok, 0][~conv] == 3)
        assert np.max(np.abs(x[ok][conv] - xe[ok][conv]) * ewt[ok][conv]) < 1e-3
    assert np.all(irc[idle] == -1.0)


@pytest.mark.gpu
@pytest.mark.parametrize("mech,wl", CASES)
def test_cooperative_lu_matches_lane_lu(mech, wl):
    """The cooperative LU (32 lanes per cell and two cells per wave, or one cell per wave
    for the 53-species mechanism) against the one-lane
    factorisation on reactor Jacobians: same pivot order, same factors to rounding, a
    valid permutation, the first zero pivot reported for singular matrices, and cells
    with no set-up pending left untouched.  An odd cell count leaves the last wave half
    empty."""
    G, Cl = _load("libunit_gpu.so"), _load("libunit_cpu.so")
    m, N1, y, t, rE0, fH, fy = _inputs(mech, wl, 645)
    n = y.shape[0]
    _, T = run_rhs(Cl, mech, y, t, rE0, fH, fy)
    f, _ = run_rhs(Cl, mech, y, t, rE0, fH, fy)
    rng = np.random.default_rng(12)
    gamma = 10 ** rng.uniform(-10, -5, n)
    b = rng.normal(size=(n, N1))
    J, _, _, _ = run_jac_lu(Cl, mech, y, T, f[:, -1].copy(), gamma, b)
    J = np.ascontiguousarray(J)
    pend = (rng.random(n) < 0.8).astype(np.int32)
    # singular matrices: column k of I - gamma J vanishes exactly (gamma a power of two)
    sing_cells = np.arange(5, n, 97)
    for i, c in enumerate(sing_cells):
        k = (3 * i) % N1
        gamma[c] = 2.0 ** -20
        J[c, :, k] = 0.0
        J[c, k, k] = 2.0 ** 20
        pend[c] = 1
    xc, sc, luc, pc, _ = run_lu(G, mech, 1, J, gamma, pend, b)
    xl, sl, lul, pl, _ = run_lu(G, mech, 0, J, gamma, pend, b)
    xh, sh, luh, ph, _ = run_lu(Cl, mech, 0, J, gamma, pend, b)
    _dump(f"lu_{mech}", xc=xc, xl=xl, sc=sc, sl=sl, pc=pc, pl=pl)
    idle = pend == 0
    assert np.all(sc[idle] == -3.0) and np.all(luc[idle] == 7.0) and np.all(pc[idle] == -1)
    ok = (pend == 1) & (sl == 0)
    is_sing = np.zeros(n, bool)
    is_sing[sing_cells] = True
    assert np.all(sl[is_sing] > 0) and np.all(sl[~is_sing & (pend == 1)] == 0)
    np.testing.assert_array_equal(sc[pend == 1], sl[pend == 1])
    np.testing.assert_array_equal(sl, sh)
    for c in np.where(pend == 1)[0]:
        assert sorted(pc[c]) == list(range(N1))
    np.testing.assert_array_equal(pc[ok], pl[ok])
    # the CPU twin does not contract to FMA: a near-tie may pivot differently there
    assert np.mean(np.all(pl[ok] == ph[ok], axis=1)) > 0.99
    scale = np.abs(lul[ok]).max(axis=2, keepdims=True)
    assert np.max(np.abs(luc[ok] - lul[ok]) / scale) < 1e-13
    xs = np.abs(xl[ok]).max(axis=1, keepdims=True)
    assert np.max(np.abs(xc[ok] - xl[ok]) / xs) < 1e-12
    assert np.max(np.abs(xl[ok] - xh[ok]) / xs) < 1e-10
    # and the solution solves the system
    M = np.eye(N1)[None] - gamma[ok, None, None] * J[ok]
    res = np.einsum("cij,cj->ci", M, xc[ok]) - b[ok]
    assert np.max(np.abs(res) / (np.abs(M).max(axis=2) * xs + np.abs(b[ok]))) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("mech", list(N1_OF))
def test_bdf_coefficients_gpu_matches_cpu(mech):
    G, Cl = _load("libunit_gpu.so"), _load("libunit_cpu.so")
    N1 = N1_OF[mech]
    scal, zn = coeff_inputs(256, N1)
    og = run_coeffs(G, mech, scal, zn)
    oc = run_coeffs(Cl, mech, scal, zn)
    _dump(f"coeffs_{mech}", og=og, oc=oc, scal=scal, zn=zn)
    np.testing.assert_allclose(og, oc, rtol=1e-12, atol=1e-300)


def stepctl_inputs(n, N1, seed=9):
    rng = np.random.default_rng(seed)
    scal = np.zeros((n, 18))
    scal[:, 0] = rng.integers(1, 5, n)                  # q
    scal[:, 1] = 4                                      # qmax
    scal[:, 2] = rng.integers(0, 3, n)                  # qwait
    scal[:, 3] = rng.choice([1.0, 10.0, 1e4], n)        # etamax
    h = 10 ** rng.uniform(-12, -7, n)
    scal[:, 4] = h
    scal[:, 5] = 10 ** rng.uniform(-6, 0, n)            # dsm
    scal[:, 6:11] = rng.uniform(0.05, 2.0, (n, 5))      # tq1..5
    scal[:, 11] = rng.choice([0.0, 1.0], n) * rng.uniform(0.1, 2, n)   # saved_tq5
    scal[:, 12:18] = h[:, None] * rng.uniform(0.3, 3.0, (n, 6))
    zn = rng.normal(size=(n, 6, N1)) * 1e-3
    acor = rng.normal(size=(n, N1)) * 1e-4
    ewt = 10 ** rng.uniform(2, 8, (n, N1))
    px = 10 ** rng.uniform(-6, 1.5, n)
    py = rng.choice([1 / 2, 1 / 3, 1 / 4, 1 / 5, 2.0, 3.0, 5.0], n)
    return scal, zn, acor, ewt, px, py


def run_stepctl(L, mech, scal, zn, acor, ewt, px, py):
    n, N1 = acor.shape
    out = np.zeros((n, 4 + N1))
    pz = np.zeros(n)
    assert L.unit_stepctl(mech.encode(), C.c_long(n), _p(scal), _p(np.ascontiguousarray(zn)), _p(acor), _p(ewt),
                          _p(out), _p(px), _p(py), _p(pz)) == 0
    return out, pz


@pytest.mark.gpu
@pytest.mark.parametrize("mech", list(N1_OF))
def test_step_control_gpu_matches_cpu(mech):
    G, Cl = _load("libunit_gpu.so"), _load("libunit_cpu.so")
    N1 = N1_OF[mech]
    args = stepctl_inputs(512, N1)
    og, pg = run_stepctl(G, mech, *args)
    oc, pc = run_stepctl(Cl, mech, *args)
    _dump(f"stepctl_{mech}", og=og, oc=oc, pg=pg, pc=pc)
    np.testing.assert_allclose(pg, pc, rtol=1e-15)
    np.testing.assert_allclose(og, oc, rtol=1e-12, atol=1e-300)


def test_matrix_core_probe_needs_a_wave_on_the_cpu_twin():
    Cl = _load("libunit_cpu.so")
    z = np.zeros(256)
    assert Cl.unit_mfma_probe(1, _p(z), _p(z), _p(z), _p(z)) == -2


def _mfma_pack(A, B, Cm, f64):
    """Lane registers of one 16 x 16 x 4 product: lane l holds A[l % 16][l / 16] and
    B[l / 16][l % 16]; register v of lane l holds D[l / 16 + 4 v][l % 16] in the FP64
    product and D[4 (l / 16) + v][l % 16] in the FP32 one (the layout the blocked LU uses)."""
    lane = np.arange(64)
    a, b = A[lane % 16, lane // 16], B[lane // 16, lane % 16]
    rows = np.array([[(l // 16 + 4 * v) if f64 else (4 * (l // 16) + v) for v in range(4)] for l in lane])
    cols = np.repeat((lane % 16)[:, None], 4, axis=1)
    return np.ascontiguousarray(a), np.ascontiguousarray(b), rows.reshape(-1), cols.reshape(-1)


@pytest.mark.gpu
@pytest.mark.parametrize("f64", [1, 0])
def test_matrix_core_operand_layout(f64):
    """One v_mfma_{f64,f32}_16x16x4 product on raw lane registers reproduces C + A B with
    the operand layout the blocked LU assumes; a transposed or shifted operand fails by
    O(1)."""
    G = _load("libunit_gpu.so")
    rng = np.random.default_rng(5 + f64)
    A, B, Cm = rng.normal(size=(16, 4)), rng.normal(size=(4, 16)), rng.normal(size=(16, 16))
    if not f64:
        A, B, Cm = (x.astype(np.float32).astype(np.float64) for x in (A, B, Cm))
    a, b, rows, cols = _mfma_pack(A, B, Cm, f64)
    c = np.ascontiguousarray(Cm[rows, cols])
    d = np.zeros(256)
    assert G.unit_mfma_probe(f64, _p(a), _p(b), _p(c), _p(d)) == 0
    D = np.full((16, 16), np.nan)
    D[rows, cols] = d
    ref = Cm + A @ B
    eps = 2.0 ** -52 if f64 else 2.0 ** -23
    bound = 8 * eps * (np.abs(Cm) + np.abs(A) @ np.abs(B))
    assert np.all(np.abs(D - ref) <= bound), np.abs(D - ref).max()
    assert np.abs(D - (Cm + A @ B[:, ::-1])).max() > 1e-3   # the check can fail
    wrong = np.full((16, 16), np.nan)
    _, _, rows_other, _ = _mfma_pack(A, B, Cm, 1 - f64)
    wrong[rows_other, cols] = d
    assert np.abs(wrong - ref).max() > 1e-3                 # the two row maps differ


@pytest.mark.gpu
def test_dpp_group_maximum_of_pivot_keys():
    """The pivot searches' 64-bit maximum through DPP moves and v_readlane, on raw lane
    keys: every lane must return its half-wave's and the wave's maximum exactly.  The
    keys mix random 64-bit values (high words decide), keys equal in the high word (the
    low word decides), ties, zeros and a maximum in each of the four 16-lane rows."""
    G = _load("libunit_gpu.so")
    rng = np.random.default_rng(11)
    cases = [rng.integers(0, 2**63, size=64, dtype=np.uint64) for _ in range(3)]
    hi = np.uint64(0x3FF0000000000000)
    cases.append(hi + rng.integers(0, 2**20, size=64, dtype=np.uint64))   # equal high words
    cases.append(np.zeros(64, dtype=np.uint64))
    tie = np.ones(64, dtype=np.uint64)
    tie[[5, 40]] = np.uint64(2**40 + 1)
    cases.append(tie)
    for row in range(4):
        k = rng.integers(0, 2**32, size=64, dtype=np.uint64)
        k[16 * row + 13] = np.uint64(2**62 + 7)
        cases.append(k)
    for key in cases:
        key = np.ascontiguousarray(key, dtype=np.uint64)
        out = np.zeros(128, dtype=np.uint64)
        assert G.unit_group_max(_p(key), _p(out)) == 0
        halves = np.repeat([key[:32].max(), key[32:].max()], 32)
        np.testing.assert_array_equal(out[:64], halves)
        np.testing.assert_array_equal(out[64:], np.full(64, key.max(), dtype=np.uint64))


@pytest.mark.gpu
def test_bf16_matrix_core_operand_layout():
    """One v_mfma_f32_16x16x32_bf16 (the learned predictor's product) on raw lane
    registers with the layout the network kernels assume: lane l holds A[l % 16][8 (l / 16)
    + i] and B[8 (l / 16) + i][l % 16] (i = 0..7) and D[4 (l / 16) + v][l % 16].  Small
    integer quarters are exact in bf16 and their sums exact in FP32, so D must equal
    C + A B bit for bit; transposed operands fail."""
    G = _load("libunit_gpu.so")
    rng = np.random.default_rng(8)
    A = rng.integers(-8, 9, size=(16, 32)) / 4.0
    B = rng.integers(-8, 9, size=(32, 16)) / 4.0
    Cm = rng.integers(-8, 9, size=(16, 16)) / 4.0
    lane = np.arange(64)
    a = np.array([A[l % 16, 8 * (l // 16) + i] for l in lane for i in range(8)], dtype=np.float32)
    b = np.array([B[8 * (l // 16) + i, l % 16] for l in lane for i in range(8)], dtype=np.float32)
    rows = np.array([4 * (l // 16) + v for l in lane for v in range(4)])
    cols = np.array([l % 16 for l in lane for v in range(4)])
    c = np.ascontiguousarray(Cm[rows, cols], dtype=np.float32)
    d = np.zeros(256, dtype=np.float32)
    assert G.unit_mfma_bf16_probe(_p(a), _p(b), _p(c), _p(d)) == 0
    D = np.full((16, 16), np.nan)
    D[rows, cols] = d
    np.testing.assert_array_equal(D, Cm + A @ B)
    assert np.abs(D - (Cm + A @ B[::-1])).max() > 0.1   # the check can fail


@pytest.mark.gpu
@pytest.mark.parametrize("mech,wl", CASES)
@pytest.mark.parametrize("fp32", [0, 1])
def test_matrix_core_lu_matches_cooperative_lu(mech, wl, fp32):
    """The blocked LU (panels of four columns, the trailing update as 16 x 16 x 4 matrix-
    core tiles) against the cooperative factorisation on reactor Jacobians, FP64 and
    FP32 factors: the same singular columns reported, valid permutations that agree
    with the cooperative pivots (a near-tie may break differently where the two round
    differently), factors that reproduce P M, solves that reach the same solution (with
    FP32 factors after refinement), and cells with no set-up pending left untouched."""
    G = _load("libunit_gpu.so")
    m, N1, y, t, rE0, fH, fy = _inputs(mech, wl, 645)
    N1, n, J, gamma, pend, b = _lu_case(mech, wl)
    sing_cells = np.arange(5, n, 97)
    for i, c in enumerate(sing_cells):
        k = (3 * i) % N1
        gamma[c] = 2.0 ** -20
        J[c, :, k] = 0.0
        J[c, k, k] = 2.0 ** 20
        pend[c] = 1
    M = np.eye(N1)[None] - gamma[:, None, None] * J
    ewt = None
    if fp32:
        xe = np.linalg.solve(M + np.where(np.isin(np.arange(n), sing_cells), 1.0, 0.0)[:, None, None] * np.eye(N1),
                             b[..., None])[..., 0]
        ewt = 1.0 / (1e-6 * np.abs(xe) + 1e-10 * np.abs(xe).max(axis=1, keepdims=True))
    xm, sm, lum, pm, irm = run_lu(G, mech, 2, J, gamma, pend, b, fp32=fp32, ewt=ewt)
    xc, sc, luc, pc, irc = run_lu(G, mech, 1, J, gamma, pend, b, fp32=fp32, ewt=ewt)
    _dump(f"lu_mfma_{mech}_{fp32}", xm=xm, xc=xc, sm=sm, sc=sc, pm=pm, pc=pc, lum=lum, luc=luc)
    idle = pend == 0
    assert np.all(sm[idle] == -3.0) and np.all(lum[idle] == 7.0) and np.all(pm[idle] == -1)
    np.testing.assert_array_equal(sm[pend == 1], sc[pend == 1])
    is_sing = np.zeros(n, bool)
    is_sing[sing_cells] = True
    assert np.all(sm[is_sing] > 0)
    ok = (pend == 1) & (sc == 0)
    for c in np.where(pend == 1)[0]:
        assert sorted(pm[c]) == list(range(N1))
    agree = np.all(pm[ok] == pc[ok], axis=1)
    assert agree.mean() > (0.97 if fp32 else 0.99), agree.mean()
    Mo = M[ok]
    Lf = np.tril(lum[ok], -1) + np.eye(N1)[None]
    PM = np.take_along_axis(Mo, pm[ok][:, :, None].astype(np.int64), axis=1)
    res = np.abs(PM - Lf @ np.triu(lum[ok])).max(axis=(1, 2)) / np.abs(Mo).max(axis=(1, 2))
    assert res.max() < (1e-5 if fp32 else 1e-13), res.max()
    same = ok.copy()
    same[ok] = agree
    scale = np.abs(luc[same]).max(axis=2, keepdims=True)
    assert np.max(np.abs(lum[same] - luc[same]) / scale) < (1e-5 if fp32 else 1e-13)
    xs = np.abs(xc[ok]).max(axis=1, keepdims=True)
    if fp32:
        conv = (irm[ok, 1] == 0) & (irc[ok, 1] == 0)
        assert conv.mean() > (0.99 if N1 > 32 else 1.0 - 1e-12)
        assert np.max(np.abs(xm[ok][conv] - xc[ok][conv]) * ewt[ok][conv]) < 2e-3
    else:
        assert np.max(np.abs(xm[ok] - xc[ok]) / xs) < 1e-12
