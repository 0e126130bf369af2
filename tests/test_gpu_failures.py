"""Failure paths of the HIP reactor, each forced from the host so every kernel still
exits normally: too few steps (cvode.max_num_steps), the trip cap (mi.max_trips, which
ends in the abandon kernel), refinement that cannot converge (mi.ir_tol = 0 forces the
FP32 -> FP64 factor fallback), and the return-code policy (mi.on_fail).  Also the
pass-through ReactorNull.  Every call goes through the C ABI.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from pelelm_amd import cells, reactor  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _mk(mech, name="ReactorCvode", **keys):
    r = reactor.ReactorBase.create(name)
    if name != "ReactorNull":
        r.set("mi.mechanism", mech)
    for k, v in keys.items():
        r.set(k.replace("__", "."), v)
    r.init(2, 1)
    return r


def _dev(a, dtype=torch.float64):
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def _run(r, d, n):
    rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
    FC = torch.zeros(n, dtype=torch.float64, device=DEV)
    err = None
    try:
        r.react(rY, None, T, rE, None, FC, None, d["dt"], 0.0)
    except reactor.CellFailure as e:
        err = e
    torch.cuda.synchronize()
    return rY.cpu().numpy(), T.cpu().numpy(), FC.cpu().numpy(), r.stats(), err


@pytest.mark.parametrize("on_fail", ["continue", "abort"])
def test_too_few_steps_fail_the_oracles_cells(on_fail):
    """cvode.max_num_steps = 20 (these cells need 7 to 39 steps, median 22): the cells that
    need more steps fail -- the same cells as
    in the oracle with the same limit -- keep their input state and get FC < 0; the call
    returns MI_ECELLFAIL under mi.on_fail = abort and MI_OK under continue."""
    n = 1024
    d = cells.generate_workload("drm19_128_mix", ncell=n)
    ref = O.Mech("drm19").react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10, max_steps=20)
    assert 0 < ref["nfail"] < n
    r = _mk("drm19", ode__rtol=1e-6, ode__atol=1e-10, cvode__max_num_steps=20, mi__on_fail=on_fail)
    gY, gT, gFC, st, err = _run(r, d, n)
    r.close()
    failed = gFC < 0
    print(f"max_num_steps=20, on_fail={on_fail}: {failed.sum()} of {n} cells failed "
          f"(oracle {ref['nfail']}), error {err}")
    assert (err is not None) == (on_fail == "abort")
    assert st["nfail"] == failed.sum() == ref["nfail"]
    np.testing.assert_array_equal(failed, ref["FC"] < 0)
    np.testing.assert_array_equal(gY[:, failed], d["rY"][:, failed])
    np.testing.assert_array_equal(gT[failed], d["T"][failed])
    ok = ~failed
    np.testing.assert_array_equal(gFC[ok], ref["FC"][ok])
    tol = 1e-6 * np.abs(ref["rY"][:, ok]) + 1e-10
    assert (np.abs(gY[:, ok] - ref["rY"][:, ok]) / tol).max() <= 1e-2


def test_trip_cap_abandons_the_unfinished_cells():
    """mi.max_trips = 6: the cells still integrating after six trips are marked failed by
    the abandon kernel (FC = -(RHS count), state untouched); no cell can finish a stiff
    advance that fast, and the cells that do finish are unaffected."""
    n = 1024
    d = cells.generate_workload("drm19_128_mix", ncell=n)
    r = _mk("drm19", ode__rtol=1e-6, ode__atol=1e-10, mi__max_trips=6, mi__on_fail="continue")
    gY, gT, gFC, st, err = _run(r, d, n)
    r.close()
    failed = gFC < 0
    print(f"max_trips=6: {failed.sum()} of {n} cells abandoned, trips {st['ntrips']}")
    assert err is None and st["ntrips"] == 6
    assert st["nfail"] == failed.sum() > 0
    assert st["ncells"] == n
    assert np.all(gFC[failed] <= -1.0) and np.all(np.abs(gFC) <= 6)
    np.testing.assert_array_equal(gY[:, failed], d["rY"][:, failed])


@pytest.mark.parametrize("wl,mech", [("drm19_64", "drm19"), ("dodecane53_64_mix", "dodecane53")])
def test_refinement_failure_falls_back_to_fp64_factors(wl, mech):
    """mi.lu = fp32_ir with mi.ir_tol = 0: no refinement can meet the tolerance, so every
    cell's first Newton solve falls back to FP64 factors (nlu_fallback counts them) and
    the cell restarts that iteration from its prediction.  The states are then the FP64
    form's; only the RHS counts differ (the restarted iteration evaluates once more)."""
    n = 1024
    d = cells.generate_workload(wl, ncell=n)
    ref = O.Mech(mech).react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10)
    r = _mk(mech, ode__rtol=1e-6, ode__atol=1e-10, mi__lu="fp32_ir", mi__ir_tol=0)
    gY, gT, gFC, st, err = _run(r, d, n)
    r.close()
    print(f"{wl}: FP64 fallbacks {st['nlu_fallback']} of {n} cells, refinement sweeps {st['nlu_refine']}")
    assert err is None and st["nfail"] == 0
    assert st["nlu_fallback"] == n
    assert np.all(gFC >= ref["FC"])
    tol = 1e-6 * np.abs(ref["rY"]) + 1e-10
    assert (np.abs(gY - ref["rY"]) / tol).max() <= 1e-2
    assert np.abs(gT - ref["T"]).max() <= 1e-7


def test_reactor_null_passes_the_forcing_through():
    """ReactorNull (the pass-through most decks use): rhoY += dt F_Y, rhoh += dt F_h, T
    untouched, FC = 0, masked cells untouched."""
    n = 2048
    d = cells.generate_workload("drm19_64", ncell=n)
    rng = np.random.default_rng(3)
    fY = rng.normal(size=d["rY"].shape) * d["rY"]
    fH = rng.normal(size=n) * 1e8
    mask = np.ones(n, dtype=np.int32)
    mask[::5] = -1
    r = _mk("drm19", name="ReactorNull")
    rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
    FC = torch.full((n,), 7.0, dtype=torch.float64, device=DEV)
    dt, tm = r.react(rY, _dev(fY), T, rE, _dev(fH), FC, _dev(mask, torch.int32), d["dt"], 1.5)
    torch.cuda.synchronize()
    r.close()
    assert tm == pytest.approx(1.5 + d["dt"])
    act = mask >= 0
    gY, gT, gE, gFC = rY.cpu().numpy(), T.cpu().numpy(), rE.cpu().numpy(), FC.cpu().numpy()
    # x + dt f may be one fused multiply-add on the device: equal to within its rounding
    np.testing.assert_allclose(gY[:, act], d["rY"][:, act] + d["dt"] * fY[:, act], rtol=1e-15, atol=0)
    np.testing.assert_allclose(gE[act], d["rE"][act] + d["dt"] * fH[act], rtol=1e-15, atol=0)
    np.testing.assert_array_equal(gY[:, ~act], d["rY"][:, ~act])
    np.testing.assert_array_equal(gE[~act], d["rE"][~act])
    np.testing.assert_array_equal(gT, d["T"])
    np.testing.assert_array_equal(gFC, 0.0)
