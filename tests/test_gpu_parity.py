"""GPU parity: the HIP reactor against the CPU oracle on identical seeded cells.

Every call goes through the C ABI (pelelm_amd/reactor.py -> libpelelm_chem.so).
Tolerances: production rates are deterministic FP64 arithmetic and agree to 1e-11
relative.  The FP64 integrations run the oracle's algorithm step for step, so they are
held to the agreement that identical step sequences give: every component within 1e-2
of its tolerance (rtol |y| + atol), T within 1e-7 K, and the same RHS count (FC) in at
least 99 % of the cells.  Forms with FP32 factors are held to the tolerance itself (their
rounding can move a step decision in a cell, which then differs by O(rtol)).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from pelelm_amd import cells, reactor  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _mk(mech, **keys):
    r = reactor.ReactorBase.create("ReactorCvode")
    r.set("mi.mechanism", mech)
    for k, v in keys.items():
        r.set(k.replace("__", "."), v)
    r.init(2, 1)
    return r


def _dev(a, dtype=torch.float64):
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def _parity(label, gY, gT, gFC, ref, tol=None, y_frac=1e-2, dT=1e-7, fc=0.99):
    """The GPU states against the oracle's: max |dY| / tol, max |dT| and the share of
    cells with the oracle's RHS count, printed and checked against the bounds."""
    tol = 1e-6 * np.abs(ref["rY"]) + 1e-10 if tol is None else tol
    ey = float((np.abs(gY - ref["rY"]) / tol).max())
    et = float(np.abs(gT - ref["T"]).max())
    same = float(np.mean(gFC == ref["FC"])) if gFC is not None else 1.0
    print(f"{label}: max |dY|/tol {ey:.3g}, max |dT| {et:.3g} K, identical FC {same:.4f}")
    assert ey <= y_frac, (label, ey)
    assert et <= dT, (label, et)
    assert same >= fc, (label, same)


@pytest.mark.parametrize("wl,mech", [("lidryer_16", "LiDryer"), ("drm19_64", "drm19"),
                                     ("dodecane53_64_mix", "dodecane53"), ("dodecane56_64_mix", "dodecane56"),
                                     ("heptane88_64_mix", "heptane88")])
def test_rates_match_oracle(wl, mech):
    d = cells.generate_workload(wl, ncell=2048)
    m = O.Mech(mech)
    r = _mk(mech)
    T_in = d["T"] + 15.0   # perturbed guess for HY2T
    rY, rE, T = _dev(d["rY"]), _dev(d["rE"]), _dev(T_in)
    out = torch.zeros_like(rY)
    T_out = torch.zeros_like(T)
    r.reaction_rates(rY, rE, T, None, out, T_out=T_out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    Tg = T_out.cpu().numpy()
    np.testing.assert_array_equal(T.cpu().numpy(), T_in)   # T is read only
    for c in range(0, 2048, 7):
        rho = d["rY"][:, c].sum()
        Y = d["rY"][:, c] / rho
        Tc, _ = m.hy2t(d["rE"][c] / rho, Y, d["T"][c] + 15.0)
        assert abs(Tg[c] - Tc) <= 1e-9 * Tc
        w = m.wdot(d["rY"][:, c] / m.W, Tc) * m.W
        scale = np.abs(w).max()
        assert np.max(np.abs(got[:, c] - w)) <= 1e-11 * scale + 1e-300, c
    r.close()


@pytest.mark.parametrize("wl,mech", [("drm19_64", "drm19"), ("dodecane53_64_mix", "dodecane53")])
def test_reaction_rates_box_in_pelelm_units(wl, mech):
    """compute_instantaneous_reaction_rates on an Array4 box inside a larger fab with
    ghost cells and an EB mask, MKS in and out (reactionRateRhoY, Source/PeleLM_K.H:57-98):
    the oracle's HY2T and wdot with PeleLM's unit factors restated around them, T left
    unchanged, covered cells zero, ghost cells untouched."""
    n_side, g = 8, 2
    n = n_side ** 3
    d = cells.generate_workload(wl, ncell=n)
    m = O.Mech(mech)
    N = d["rY"].shape[0]
    F = n_side + 2 * g
    lo, hi = (3, -2, 5), (3 + n_side - 1, -2 + n_side - 1, 5 + n_side - 1)
    flo = tuple(x - g for x in lo)
    inner = (slice(None), slice(g, g + n_side), slice(g, g + n_side), slice(g, g + n_side))
    # MKS state: rhoY x 1e3 (kg/m^3), rhoH x 0.1 (J/m^3), T + 10 K as the HY2T guess
    rhoY = np.full((N, F, F, F), 7.0)
    rhoH = np.full((1, F, F, F), 7.0)
    Tf = np.full((1, F, F, F), 7.0)
    rhoY[inner] = (d["rY"] * 1e3).reshape(N, n_side, n_side, n_side)
    rhoH[inner] = (d["rE"] * 0.1).reshape(1, n_side, n_side, n_side)
    Tf[inner] = (d["T"] + 10.0).reshape(1, n_side, n_side, n_side)
    mask = np.ones((1, F, F, F), dtype=np.int32)
    rng = np.random.default_rng(7)
    covered = rng.random((n_side, n_side, n_side)) < 0.2
    mask[inner][0][covered] = -1
    mflat = np.where(covered.reshape(-1), -1, 1)
    gY, gH, gT, gM = _dev(rhoY), _dev(rhoH), _dev(Tf), _dev(mask, torch.int32)
    gR = _dev(np.full((N, F, F, F), -3.0))
    r = _mk(mech)
    r.reaction_rates_box(lo, hi, reactor.array4(gY, flo, None, N), reactor.array4(gH, flo, None, 1),
                         reactor.array4(gT, flo, None, 1), reactor.int4(gM, flo),
                         reactor.array4(gR, flo, None, N))
    torch.cuda.synchronize()
    r.close()
    got = gR.cpu().numpy()
    np.testing.assert_array_equal(gT.cpu().numpy(), Tf)            # T read only
    outside = np.ones(got.shape, dtype=bool)
    outside[inner] = False
    assert np.all(got[outside] == -3.0)                            # ghost cells untouched
    gin = got[inner].reshape(N, -1)
    assert np.all(gin[:, mflat < 0] == 0.0)
    for c in np.nonzero(mflat > 0)[0][::3]:
        rho_mks = d["rY"][:, c].sum() * 1e3
        Y = d["rY"][:, c] * 1e3 / rho_mks
        Tc, _ = m.hy2t(d["rE"][c] * 0.1 / rho_mks * 1e4, Y, d["T"][c] + 10.0)
        w = m.wdot(rho_mks * 1e-3 * Y / m.W, Tc) * m.W * 1e3
        scale = np.abs(w).max()
        assert np.max(np.abs(gin[:, c] - w)) <= 1e-10 * scale + 1e-300, c


@pytest.mark.parametrize("wl,mech,n", [("lidryer_16", "LiDryer", 4096), ("drm19_64", "drm19", 4096),
                                       ("drm19_128_mix", "drm19", 4096), ("dodecane53_64_mix", "dodecane53", 2048),
                                       ("dodecane56_64_mix", "dodecane56", 2048), ("heptane88_64_mix", "heptane88", 1024)])
@pytest.mark.parametrize("schedule", ["plain", "queue"])
def test_react_matches_oracle(wl, mech, n, schedule):
    """The reactor against the oracle on every mechanism; the 88-species one factors and
    solves one cell per lane (its 89 unknowns exceed a wave's rows)."""
    d = cells.generate_workload(wl, ncell=n)
    m = O.Mech(mech)
    ref = m.react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10, stats=True)
    assert ref["nfail"] == 0
    r = _mk(mech, ode__rtol=1e-6, ode__atol=1e-10, mi__schedule=schedule)
    rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
    FC = torch.zeros(n, dtype=torch.float64, device=DEV)
    dt, tm = r.react(rY, None, T, rE, None, FC, None, d["dt"], 0.0)
    torch.cuda.synchronize()
    st = r.stats()
    r.close()
    assert st["nfail"] == 0 and st["ncells"] == n
    assert tm == pytest.approx(d["dt"])
    gY, gT, gE, gFC = rY.cpu().numpy(), T.cpu().numpy(), rE.cpu().numpy(), FC.cpu().numpy()
    _parity(f"{wl} {schedule}", gY, gT, gFC, ref)
    np.testing.assert_array_equal(gE, ref["rE"])
    # mass conservation is exact up to rounding
    np.testing.assert_allclose(gY.sum(0), d["rY"].sum(0), rtol=1e-12)


@pytest.mark.parametrize("wl,mech", [("drm19_64", "drm19"), ("lidryer_16", "LiDryer"),
                                     ("dodecane53_64_mix", "dodecane53")])
def test_fp32_ir_matches_fp64_and_tight_oracle(wl, mech):
    """Mixed precision (FP32 factors, up to 3 FP64-residual refinement sweeps per solve)
    against the FP64 oracle at the same rtol and against the tight oracle (rtol 1e-12):
    it must add no error beyond what the rtol-1e-6 integration itself makes."""
    n = 2048
    d = cells.generate_workload(wl, ncell=n)
    m = O.Mech(mech)
    ref = m.react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10)
    tight = m.react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-12, atol=1e-20, max_steps=200000)
    r = _mk(mech, ode__rtol=1e-6, ode__atol=1e-10, mi__lu="fp32_ir", mi__ir_sweeps=3)
    rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
    FC = torch.zeros(n, dtype=torch.float64, device=DEV)
    r.react(rY, None, T, rE, None, FC, None, d["dt"], 0.0)
    torch.cuda.synchronize()
    st = r.stats()
    r.close()
    assert st["nfail"] == 0 and st["nlu_refine"] > 0
    gY, gT = rY.cpu().numpy(), T.cpu().numpy()
    _parity(f"{wl} fp32_ir", gY, gT, FC.cpu().numpy(), ref, y_frac=1.0, dT=1e-3, fc=0.95)
    e_ir = np.abs(gT - tight["T"]).max()
    e_64 = np.abs(ref["T"] - tight["T"]).max()
    print(f"{wl}: max|dT| vs tight oracle: fp32_ir {e_ir:.3g} K, fp64 {e_64:.3g} K; "
          f"refinement sweeps {st['nlu_refine']}, FP64 fallbacks {st['nlu_fallback']}")
    assert e_ir <= 1.01 * e_64 + 1e-6


def test_forcing_matches_oracle():
    n = 1024
    d = cells.generate_workload("drm19_64", ncell=n)
    m = O.Mech("drm19")
    rng = np.random.default_rng(5)
    fY = rng.normal(size=d["rY"].shape) * 1e-3 * d["rY"]
    fH = rng.normal(size=n) * 1e6
    ref = m.react(d["rY"], d["T"], d["rE"], d["dt"], rYsrc=fY, rEsrc=fH, rtol=1e-6, atol=1e-10)
    r = _mk("drm19", ode__rtol=1e-6, ode__atol=1e-10)
    rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
    FC = torch.zeros(n, dtype=torch.float64, device=DEV)
    r.react(rY, _dev(fY), T, rE, _dev(fH), FC, None, d["dt"], 0.0)
    torch.cuda.synchronize()
    assert r.stats()["nfail"] == 0
    r.close()
    gY, gT, gE = rY.cpu().numpy(), T.cpu().numpy(), rE.cpu().numpy()
    _parity("drm19_64 forced", gY, gT, FC.cpu().numpy(), ref)
    np.testing.assert_allclose(gE, d["rE"] + d["dt"] * fH, rtol=1e-15)
    np.testing.assert_allclose(gY.sum(0), d["rY"].sum(0) + d["dt"] * fY.sum(0), rtol=1e-12)


@pytest.mark.parametrize("wl,mech", [("drm19_64", "drm19"), ("dodecane53_64_mix", "dodecane53")])
def test_forced_T_stays_on_its_enthalpy(wl, mech):
    """rho h = sum_k rhoY_k h_k(T): after a forced advance the temperature the kernels
    integrate must be the one the advanced (rho h, rhoY) define, T_out = HY2T(rhoh_out /
    rho_out, Y_out), to 1e-6 K at rtol 1e-8 (an energy equation without the enthalpy the
    species forcing carries misses by about 0.2 K here)."""
    n = 256
    d = cells.generate_workload(wl, ncell=n)
    m = O.Mech(mech)
    dt = 1e-6
    rng = np.random.default_rng(11)
    fY = rng.uniform(-1.0, 1.0, size=d["rY"].shape) * 1e-4 * d["rY"] / dt
    fH = rng.uniform(-1.0, 1.0, size=n) * 1e-4 * np.abs(d["rE"]) / dt
    r = _mk(mech, ode__rtol=1e-8, ode__atol=1e-14, cvode__max_num_steps=200000)
    rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
    FC = torch.zeros(n, dtype=torch.float64, device=DEV)
    r.react(rY, _dev(fY), T, rE, _dev(fH), FC, None, dt, 0.0)
    torch.cuda.synchronize()
    assert r.stats()["nfail"] == 0
    r.close()
    gY, gT, gE = rY.cpu().numpy(), T.cpu().numpy(), rE.cpu().numpy()
    rho = gY.sum(axis=0)
    worst = 0.0
    for c in range(n):
        Tc, _ = m.hy2t(gE[c] / rho[c], gY[:, c] / rho[c], gT[c])
        worst = max(worst, abs(gT[c] - Tc))
    print(f"{wl}: max |T_out - HY2T(rhoh_out, Y_out)| = {worst:.3g} K under forcing at rtol 1e-8")
    assert worst <= 1e-6


def test_mask_skips_covered_cells():
    n = 1024
    d = cells.generate_workload("lidryer_16", ncell=n)
    mask = np.zeros(n, dtype=np.int32)
    mask[::3] = -1
    ref = O.Mech("LiDryer").react(d["rY"], d["T"], d["rE"], d["dt"], mask=mask)
    r = _mk("LiDryer", ode__rtol=1e-6, ode__atol=1e-10)
    rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
    FC = torch.full((n,), 123.0, dtype=torch.float64, device=DEV)
    r.react(rY, None, T, rE, None, FC, _dev(mask, torch.int32), d["dt"], 0.0)
    torch.cuda.synchronize()
    st = r.stats()
    r.close()
    assert st["ncells"] == n - len(mask[::3])
    gY, gT, gFC = rY.cpu().numpy(), T.cpu().numpy(), FC.cpu().numpy()
    np.testing.assert_array_equal(gY[:, ::3], d["rY"][:, ::3])
    np.testing.assert_array_equal(gT[::3], d["T"][::3])
    act = mask >= 0
    sub = {"rY": ref["rY"][:, act], "T": ref["T"][act], "FC": ref["FC"][act]}
    _parity("lidryer_16 masked", gY[:, act], gT[act], gFC[act], sub)
    np.testing.assert_array_equal(gFC[~act], 0.0)


def test_react_box_on_array4_views_with_ghost_cells():
    """A 12^3 valid box inside a 16^3 fab (2 ghost cells each side), components in planes
    like PeleLM's STemp (rhoY, rhoH, T, Source/PeleLM.cpp:5728-5730): only valid cells
    change, and they match the flat call."""
    nv, ng = 12, 2
    nf = nv + 2 * ng
    d = cells.generate_workload("drm19_64", ncell=nf ** 3)
    N = d["rY"].shape[0]
    fab = np.concatenate([d["rY"], d["rE"][None], d["T"][None]], axis=0).reshape(N + 2, nf, nf, nf)
    S = _dev(fab)
    FCf = torch.zeros((1, nf, nf, nf), dtype=torch.float64, device=DEV)
    lo_fab = (-ng, -ng, -ng)
    r = _mk("drm19", ode__rtol=1e-6, ode__atol=1e-10)
    v = lambda t, nc: reactor.array4(t, lo_fab, None, nc)  # noqa: E731
    rYv, rEv, Tv = v(S[:N], N), v(S[N:N + 1], 1), v(S[N + 1:N + 2], 1)
    z = reactor.null_real4()
    r.react_box((0, 0, 0), (nv - 1, nv - 1, nv - 1), rYv, z, Tv, rEv, z, v(FCf, 1), reactor.int4(None, lo_fab),
                d["dt"], 0.0)
    torch.cuda.synchronize()
    assert r.stats()["ncells"] == nv ** 3
    r.close()
    out = S.cpu().numpy()
    inner = (slice(None), slice(ng, ng + nv), slice(ng, ng + nv), slice(ng, ng + nv))
    ghost = np.ones((nf, nf, nf), bool)
    ghost[ng:ng + nv, ng:ng + nv, ng:ng + nv] = False
    np.testing.assert_array_equal(out[:, ghost], fab[:, ghost])
    flat_in = fab[inner].reshape(N + 2, -1)
    ref = O.Mech("drm19").react(flat_in[:N], flat_in[N + 1], flat_in[N], d["dt"], rtol=1e-6, atol=1e-10)
    got = out[inner].reshape(N + 2, -1)
    _parity("drm19 Array4 box", got[:N], got[N + 1], None, ref)


@pytest.mark.parametrize("wl,mech", [("drm19_128_mix", "drm19"), ("dodecane53_64_mix", "dodecane53")])
@pytest.mark.parametrize("schedule", ["plain", "queue"])
def test_graph_captured_trips_match_oracle(schedule, wl, mech):
    """Trips captured into HIP graphs of 8 (grids sized for every cell, kernels skipping
    beyond their list counts) give the host-driven results."""
    n = 4096
    d = cells.generate_workload(wl, ncell=n)
    ref = O.Mech(mech).react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10)
    r = _mk(mech, ode__rtol=1e-6, ode__atol=1e-10, mi__schedule=schedule, mi__graph_trips=8)
    rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
    FC = torch.zeros(n, dtype=torch.float64, device=DEV)
    r.react(rY, None, T, rE, None, FC, None, d["dt"], 0.0)
    torch.cuda.synchronize()
    st = r.stats()
    r.close()
    assert st["nfail"] == 0 and st["graph_trips_used"] == 8 and st["ntrips"] % 8 == 0
    _parity(f"{wl} {schedule} graphs", rY.cpu().numpy(), T.cpu().numpy(), FC.cpu().numpy(), ref)


@pytest.mark.parametrize("wl,mech,n", [("lidryer_16", "LiDryer", 4096), ("drm19_128_mix", "drm19", 16384)])
def test_auto_schedule_tunes_once_per_box_size_and_caches(wl, mech, n, tmp_path):
    """mi.schedule = auto: the first call of a box size runs the successive-halving search
    (9 candidates), later calls reuse the choice, the results match the oracle whatever
    was chosen; with mi.tune_cache the choice goes to a JSON file, and a new reactor with
    the same file takes it from there without searching."""
    d = cells.generate_workload(wl, ncell=n)
    ref = O.Mech(mech).react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10)
    cache = str(tmp_path / "tune.json")
    chosen = None
    for reactor_no in range(2):
        r = _mk(mech, ode__rtol=1e-6, ode__atol=1e-10, mi__schedule="auto", mi__tune_cache=cache)
        tunes = []
        for call in range(2):
            rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
            FC = torch.zeros(n, dtype=torch.float64, device=DEV)
            r.react(rY, None, T, rE, None, FC, None, d["dt"], 0.0)
            torch.cuda.synchronize()
            st = r.stats()
            assert st["nfail"] == 0
            _parity(f"{wl} auto reactor {reactor_no} call {call}", rY.cpu().numpy(), T.cpu().numpy(),
                    FC.cpu().numpy(), ref)
            tunes.append(list(st["tune_ms"]))
            used = (st["schedule_used"], st["graph_trips_used"])
            chosen = used if chosen is None else chosen
            assert used == chosen
        r.close()
        assert tunes[0] == tunes[1]   # searched (or read) on the first call only
        if reactor_no == 0:
            assert tunes[0][2] == 9 and tunes[0][3] == 0 and tunes[0][1] > tunes[0][0] > 0
            print(f"auto, {n} {wl} cells: chose schedule {chosen[0]} graphs {chosen[1]}, "
                  f"{tunes[0][0]:.2f} ms, search {tunes[0][1]:.0f} ms")
        else:
            assert tunes[0][3] == 1 and tunes[0][2] == 0   # from the cache file
    txt = open(cache).read()
    assert mech in txt and str(n) in txt and '"schedule"' in txt


def test_auto_schedule_on_the_88_species_mechanism():
    """mi.schedule = auto above 64 unknowns searches the trip schedules only (5 candidates:
    the persistent kernel does not run these mechanisms) and matches the oracle."""
    n = 1024
    d = cells.generate_workload("heptane88_64_mix", ncell=n)
    ref = O.Mech("heptane88").react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10)
    r = _mk("heptane88", ode__rtol=1e-6, ode__atol=1e-10, mi__schedule="auto")
    rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
    FC = torch.zeros(n, dtype=torch.float64, device=DEV)
    r.react(rY, None, T, rE, None, FC, None, d["dt"], 0.0)
    torch.cuda.synchronize()
    st = r.stats()
    r.close()
    assert st["nfail"] == 0 and st["tune_ms"][2] == 5 and st["schedule_used"] in (0, 1)
    _parity("heptane88 auto", rY.cpu().numpy(), T.cpu().numpy(), FC.cpu().numpy(), ref)


def test_typical_values_scale_the_absolute_tolerances():
    """peleLM.use_typ_vals_chem = 1 with PeleLM's typical values
    (update_typical_values_chem, Source/PeleLM.cpp:1497-1516: max(typical_Y_val_min,
    typY * typrho * 1e-3) per species, typT for T): component i's absolute tolerance is
    ode.atol * typ[i], checked against the oracle run with that tolerance vector."""
    n = 2048
    d = cells.generate_workload("drm19_64", ncell=n)
    N = d["rY"].shape[0]
    rho = d["rY"].sum(axis=0)
    typ_rho = rho.max() * 1e3                       # MKS, as PeleLM's typical values are
    typ_Y = (d["rY"] / rho).max(axis=1)
    typ = [max(1e-10, typ_Y[k] * typ_rho * 1e-3) for k in range(N)] + [float(d["T"].max())]
    atol = 1e-8
    ref = O.Mech("drm19").react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6,
                                atol_vec=np.array(typ) * atol, stats=True)
    plain = O.Mech("drm19").react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=atol, stats=True)
    assert ref["nfail"] == 0
    # the vector changes the integration (the scalar tolerance takes other steps)
    assert not np.array_equal(ref["FC"], plain["FC"])
    r = _mk("drm19", ode__rtol=1e-6, ode__atol=atol, peleLM__use_typ_vals_chem=1)
    r.set_typ_vals_ode(typ)
    rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
    FC = torch.zeros(n, dtype=torch.float64, device=DEV)
    r.react(rY, None, T, rE, None, FC, None, d["dt"], 0.0)
    torch.cuda.synchronize()
    assert r.stats()["nfail"] == 0
    r.close()
    tol = 1e-6 * np.abs(ref["rY"]) + atol * np.array(typ[:N])[:, None]
    _parity("drm19_64 typical values", rY.cpu().numpy(), T.cpu().numpy(), FC.cpu().numpy(), ref, tol=tol)


@pytest.mark.parametrize("wl,mech", [("drm19_64", "drm19"), ("dodecane53_64_mix", "dodecane53")])
@pytest.mark.parametrize("lu", ["fp64", "fp32_ir"])
def test_one_lane_linear_algebra_matches_oracle(wl, mech, lu):
    """mi.lu_form = lane: one lane per cell factors and solves (the kernels the CPU
    emulation runs), in FP64 and with FP32 factors plus refinement, in the reactor."""
    n = 1024
    d = cells.generate_workload(wl, ncell=n)
    ref = O.Mech(mech).react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10)
    r = _mk(mech, ode__rtol=1e-6, ode__atol=1e-10, mi__lu_form="lane", mi__lu=lu)
    rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
    FC = torch.zeros(n, dtype=torch.float64, device=DEV)
    r.react(rY, None, T, rE, None, FC, None, d["dt"], 0.0)
    torch.cuda.synchronize()
    st = r.stats()
    r.close()
    assert st["nfail"] == 0 and st["ncells"] == n
    if lu == "fp64":
        _parity(f"{wl} lane fp64", rY.cpu().numpy(), T.cpu().numpy(), FC.cpu().numpy(), ref)
    else:
        _parity(f"{wl} lane fp32_ir", rY.cpu().numpy(), T.cpu().numpy(), FC.cpu().numpy(), ref,
                y_frac=1.0, dT=1e-3, fc=0.95)


@pytest.mark.parametrize("wl,mech", [("drm19_128_mix", "drm19"), ("dodecane53_64_mix", "dodecane53")])
@pytest.mark.parametrize("lu,graphs", [("fp64", 0), ("fp32_ir", 0), ("fp64", 8)])
def test_matrix_core_lu_matches_oracle(wl, mech, lu, graphs):
    """mi.lu_form = mfma: the blocked LU on 16 x 16 x 4 matrix-core tiles in the reactor,
    FP64 factors (held to the FP64 step-for-step bounds) and FP32 factors with refinement,
    host-driven and in graph-captured trips."""
    n = 4096
    d = cells.generate_workload(wl, ncell=n)
    ref = O.Mech(mech).react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10)
    r = _mk(mech, ode__rtol=1e-6, ode__atol=1e-10, mi__lu_form="mfma", mi__lu=lu, mi__graph_trips=graphs)
    rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
    FC = torch.zeros(n, dtype=torch.float64, device=DEV)
    r.react(rY, None, T, rE, None, FC, None, d["dt"], 0.0)
    torch.cuda.synchronize()
    st = r.stats()
    r.close()
    assert st["nfail"] == 0 and st["ncells"] == n
    label = f"{wl} mfma {lu} graphs={graphs}"
    if lu == "fp64":
        _parity(label, rY.cpu().numpy(), T.cpu().numpy(), FC.cpu().numpy(), ref)
    else:
        _parity(label, rY.cpu().numpy(), T.cpu().numpy(), FC.cpu().numpy(), ref, y_frac=1.0, dT=1e-3, fc=0.95)


@pytest.mark.parametrize("wl,mech,schedule,graphs", [("lidryer_16", "LiDryer", "plain", 0),
                                                     ("drm19_128_mix", "drm19", "plain", 0),
                                                     ("drm19_128_mix", "drm19", "queue", 0),
                                                     ("drm19_128_mix", "drm19", "plain", 8)])
def test_gmres_matches_oracle(wl, mech, schedule, graphs):
    """cvode.solve_type = GMRES (the reference's GPU decks): no Jacobian, no set-up, the
    Newton systems by GMRES(5) with difference-quotient J v, against the GMRES oracle
    step for step.  The difference quotients divide last-bit differences of the right-
    hand side (the GPU contracts to FMA, the oracle does not) by the perturbation's size,
    about rtol: T is held to 1e-4 K, 5 % of its tolerance (measured 8.5e-6 K on drm19;
    1e-7 K with the direct solve)."""
    n = 4096
    d = cells.generate_workload(wl, ncell=n)
    ref = O.Mech(mech).react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10, stats=True, solve="gmres")
    assert ref["nfail"] == 0
    r = _mk(mech, ode__rtol=1e-6, ode__atol=1e-10, cvode__solve_type="GMRES", mi__schedule=schedule,
            mi__graph_trips=graphs)
    rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
    FC = torch.zeros(n, dtype=torch.float64, device=DEV)
    r.react(rY, None, T, rE, None, FC, None, d["dt"], 0.0)
    torch.cuda.synchronize()
    st = r.stats()
    r.close()
    assert st["nfail"] == 0 and st["ncells"] == n
    assert st["nje"] == 0 and st["nsetups"] == 0
    assert st["nli"] == pytest.approx(ref["stats"]["nli"].sum(), rel=1e-2)
    _parity(f"{wl} GMRES {schedule} graphs={graphs}", rY.cpu().numpy(), T.cpu().numpy(), FC.cpu().numpy(), ref,
            dT=1e-4)


def test_gmres_on_the_stiff_53_species_mix():
    """GMRES with a Krylov space of every unknown but one on the 53-species stiffness
    mix, where the difference quotients make the method itself sensitive to the right-
    hand side's last bits (the oracle against itself with sig scaled by 1 + 1e-13 keeps
    the step sequence in 83 % of the cells and moves T by up to 0.008 K): the GPU is held
    to that level against the oracle and to the oracle's accuracy against the tight
    reference."""
    n = 1024
    d = cells.generate_workload("dodecane53_64_mix", ncell=n)
    m = O.Mech("dodecane53")
    ref = m.react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10, solve="gmres", maxl=53)
    tight = m.react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-12, atol=1e-20, max_steps=200000)
    r = _mk("dodecane53", ode__rtol=1e-6, ode__atol=1e-10, cvode__solve_type="GMRES", mi__gmres_maxl=53)
    rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
    FC = torch.zeros(n, dtype=torch.float64, device=DEV)
    r.react(rY, None, T, rE, None, FC, None, d["dt"], 0.0)
    torch.cuda.synchronize()
    st = r.stats()
    r.close()
    assert st["nfail"] == 0 and st["nje"] == 0
    gT = T.cpu().numpy()
    _parity("dodecane53 GMRES(53)", rY.cpu().numpy(), gT, FC.cpu().numpy(), ref, y_frac=20.0, dT=0.05, fc=0.75)
    e_g, e_o = np.abs(gT - tight["T"]).max(), np.abs(ref["T"] - tight["T"]).max()
    print(f"max |dT| vs tight: GPU {e_g:.3g} K, oracle {e_o:.3g} K")
    assert e_g <= 2.0 * e_o + 1e-3


def test_gmres_rejects_what_it_cannot_run():
    """GMRES runs on the trip schedules: the persistent kernel, a Krylov space larger than
    the species count and the learned predictor are refused with an error."""
    d = cells.generate_workload("lidryer_16", ncell=64)
    for keys, msg in (({"mi__schedule": "persistent"}, "persistent"), ({"mi__gmres_maxl": 10}, "gmres_maxl")):
        r = _mk("LiDryer", cvode__solve_type="GMRES", **keys)
        rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
        FC = torch.zeros(64, dtype=torch.float64, device=DEV)
        with pytest.raises(reactor.ReactorError, match=msg):
            r.react(rY, None, T, rE, None, FC, None, d["dt"], 0.0)
        r.close()


@pytest.mark.parametrize("lu_form", ["mfma", "lane"])
def test_88_species_factorisation_forms_match_oracle(lu_form):
    """Above 64 unknowns: the LU staged in LDS (mi.lu_form = mfma selects it there) and the
    one-lane LU and solves, against the oracle (the default keeps the rows in registers,
    covered by test_react_matches_oracle)."""
    n = 512
    d = cells.generate_workload("heptane88_64_mix", ncell=n)
    ref = O.Mech("heptane88").react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10)
    r = _mk("heptane88", ode__rtol=1e-6, ode__atol=1e-10, mi__lu_form=lu_form)
    rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
    FC = torch.zeros(n, dtype=torch.float64, device=DEV)
    r.react(rY, None, T, rE, None, FC, None, d["dt"], 0.0)
    torch.cuda.synchronize()
    st = r.stats()
    r.close()
    assert st["nfail"] == 0 and st["ncells"] == n
    _parity(f"heptane88 lu_form={lu_form}", rY.cpu().numpy(), T.cpu().numpy(), FC.cpu().numpy(), ref)


def test_one_lane_solves_in_graph_trips_cover_every_cell():
    """mi.lu_form = lane with graph-captured trips on more cells than one fixed solve grid
    of 16,384 lanes holds: the one-lane solve's grid is sized for every cell, so cells
    past that index get their Newton solves (a grid of 256 blocks left them with stale
    solutions)."""
    n = 32768
    d = cells.generate_workload("lidryer_16", ncell=n)
    ref = O.Mech("LiDryer").react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10)
    r = _mk("LiDryer", ode__rtol=1e-6, ode__atol=1e-10, mi__lu_form="lane", mi__graph_trips=8)
    rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
    FC = torch.zeros(n, dtype=torch.float64, device=DEV)
    r.react(rY, None, T, rE, None, FC, None, d["dt"], 0.0)
    torch.cuda.synchronize()
    st = r.stats()
    r.close()
    assert st["nfail"] == 0 and st["graph_trips_used"] == 8
    _parity("lidryer 32768 lane graphs", rY.cpu().numpy(), T.cpu().numpy(), FC.cpu().numpy(), ref)


@pytest.mark.parametrize("schedule,graphs", [("plain", 0), ("queue", 0), ("plain", 8)])
def test_cost_ordering_changes_nothing_but_the_lanes(schedule, graphs):
    """mi.order = rate and mi.order = fc (the RHS counts of the previous call, which the FC
    array holds on entry as PeleLM's FuncCount does) put cells of similar cost in the same
    waves; every cell's result and RHS count stay bitwise those of the unordered call, with
    masked cells and forcing, and match the oracle."""
    n = 8192
    d = cells.generate_workload("drm19_128_mix", ncell=n)
    rng = np.random.default_rng(21)
    mask = np.where(rng.random(n) < 0.1, -1, 1).astype(np.int32)
    fY = rng.normal(size=d["rY"].shape) * 1e-3 * d["rY"] / d["dt"]
    fH = rng.normal(size=n) * 1e-3 * np.abs(d["rE"]) / d["dt"]
    ref = O.Mech("drm19").react(d["rY"], d["T"], d["rE"], d["dt"], rYsrc=fY, rEsrc=fH, mask=mask,
                                rtol=1e-6, atol=1e-10)
    out = {}
    prev_fc = None
    for order in ("none", "rate", "fc"):
        r = _mk("drm19", ode__rtol=1e-6, ode__atol=1e-10, mi__schedule=schedule, mi__graph_trips=graphs,
                mi__order=order)
        rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
        FC = _dev(prev_fc if order == "fc" else np.zeros(n))
        r.react(rY, _dev(fY), T, rE, _dev(fH), FC, _dev(mask, torch.int32), d["dt"], 0.0)
        torch.cuda.synchronize()
        st = r.stats()
        r.close()
        assert st["nfail"] == 0
        out[order] = [x.cpu().numpy() for x in (rY, T, rE, FC)]
        print(f"{schedule} graphs {graphs} order {order}: {st['ms_kernel']:.2f} ms, {st['ntrips']} trips")
        if order == "none":
            prev_fc = out[order][3]
    for order in ("rate", "fc"):
        for a, b in zip(out["none"], out[order]):
            np.testing.assert_array_equal(a, b)
    gY, gT, _, gFC = out["none"]
    _parity(f"ordered {schedule}", gY, gT, gFC, ref)
