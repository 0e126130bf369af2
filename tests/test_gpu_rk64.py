"""ReactorRK64 on the GPU against the oracle's restatement of the same explicit method.

Cells that are not stiff over their dt (LiDryer and drm19 at dt = 1e-6 s) take the
oracle's sub-steps: the same RHS count in at least 99 % of the cells and states within
1e-2 of the tolerance.  Stiff flame-front cells over dt = 1e-5 s run at the explicit
stability limit, where the accept/reject sequence follows rounding: there the same cells
fail at the budget of 500 sub-steps (FC < 0, state unchanged, MI_ECELLFAIL under
mi.on_fail = abort) and the others agree within the method's accuracy.  Every call goes
through the C ABI, flat and Array4 forms, with forcing and masks.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from pelelm_amd import cells, reactor  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _mk(mech, **keys):
    r = reactor.ReactorBase.create("ReactorRK64")
    r.set("mi.mechanism", mech)
    for k, v in keys.items():
        r.set(k.replace("__", "."), v)
    r.init(2, 1)
    return r


def _dev(a, dtype=torch.float64):
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


@pytest.mark.parametrize("wl,mech,n", [("lidryer_16", "LiDryer", 4096), ("drm19_64", "drm19", 2048)])
def test_rk64_matches_oracle(wl, mech, n):
    d = cells.generate_workload(wl, ncell=n)
    rng = np.random.default_rng(12)
    fY = rng.normal(size=d["rY"].shape) * 1e-3 * d["rY"] / d["dt"]
    fH = rng.normal(size=n) * 1e-3 * np.abs(d["rE"]) / d["dt"]
    mask = np.where(rng.random(n) < 0.1, -1, 1).astype(np.int32)
    ref = O.Mech(mech).rk64_react(d["rY"], d["T"], d["rE"], d["dt"], rYsrc=fY, rEsrc=fH, mask=mask,
                                  nsub=(10, 5, 500))
    assert ref["nfail"] == 0
    r = _mk(mech, ode__rtol=1e-6, ode__atol=1e-10)
    rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
    FC = torch.zeros(n, dtype=torch.float64, device=DEV)
    dt, tm = r.react(rY, _dev(fY), T, rE, _dev(fH), FC, _dev(mask, torch.int32), d["dt"], 0.0)
    torch.cuda.synchronize()
    st = r.stats()
    r.close()
    assert tm == pytest.approx(d["dt"])
    gY, gT, gE, gFC = rY.cpu().numpy(), T.cpu().numpy(), rE.cpu().numpy(), FC.cpu().numpy()
    tol = 1e-6 * np.abs(ref["rY"]) + 1e-10
    ey = float((np.abs(gY - ref["rY"]) / tol).max())
    et = float(np.abs(gT - ref["T"]).max())
    same = float(np.mean(gFC == ref["FC"]))
    print(f"{wl} RK64: max |dY|/tol {ey:.3g}, max |dT| {et:.3g} K, identical FC {same:.4f}, "
          f"{st['nfe'] / max(1, st['ncells']):.1f} RHS per cell, {st['ms_kernel']:.2f} ms")
    assert st["nfail"] == 0 and st["ncells"] == int((mask >= 0).sum())
    assert ey <= 1e-2 and et <= 1e-7 and same >= 0.99
    np.testing.assert_array_equal(gFC[mask < 0], 0.0)
    np.testing.assert_allclose(gE[mask >= 0], (d["rE"] + d["dt"] * fH)[mask >= 0], rtol=1e-15)
    assert st["nfe"] == int(gFC.sum())


def test_rk64_fails_stiff_cells_as_the_oracle_does():
    n = 1024
    d = cells.generate_workload("drm19_128_mix", ncell=n)
    ref = O.Mech("drm19").rk64_react(d["rY"], d["T"], d["rE"], d["dt"], nsub=(10, 5, 500))
    assert 0 < ref["nfail"] < n
    for on_fail in ("continue", "abort"):
        r = _mk("drm19", ode__rtol=1e-6, ode__atol=1e-10, mi__on_fail=on_fail)
        rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
        FC = torch.zeros(n, dtype=torch.float64, device=DEV)
        err = None
        try:
            r.react(rY, None, T, rE, None, FC, None, d["dt"], 0.0)
        except reactor.CellFailure as e:
            err = e
        torch.cuda.synchronize()
        st = r.stats()
        r.close()
        assert (err is not None) == (on_fail == "abort")
        gY, gFC = rY.cpu().numpy(), FC.cpu().numpy()
        failed = gFC < 0
        print(f"RK64 stiff cells, on_fail={on_fail}: {failed.sum()} failed (oracle {ref['nfail']})")
        np.testing.assert_array_equal(failed, ref["FC"] < 0)
        assert st["nfail"] == failed.sum()
        np.testing.assert_array_equal(gY[:, failed], d["rY"][:, failed])
        ok = ~failed
        tol = 1e-6 * np.abs(ref["rY"]) + 1e-10
        assert np.max(np.abs(gY - ref["rY"])[:, ok] / tol[:, ok]) < 10.0
        assert np.max(np.abs(gFC[ok] - ref["FC"][ok]) / ref["FC"][ok]) < 0.05


def test_rk64_on_array4_views():
    """The box form on views inside a fab with ghost cells: the interior matches the flat
    call, nothing outside the box is written."""
    nv, g = 8, 2
    nf = nv + 2 * g
    d = cells.generate_workload("drm19_64", ncell=nv ** 3)
    N = d["rY"].shape[0]
    fab = np.full((N + 2, nf, nf, nf), 5.0)
    inner = (slice(None), slice(g, g + nv), slice(g, g + nv), slice(g, g + nv))
    fab[inner] = np.concatenate([d["rY"], d["rE"][None], d["T"][None]]).reshape(N + 2, nv, nv, nv)
    F = _dev(fab)
    FC = torch.zeros((1, nf, nf, nf), dtype=torch.float64, device=DEV)
    lo_fab = (-g, -g, -g)
    lo, hi = (0, 0, 0), (nv - 1, nv - 1, nv - 1)
    r = _mk("drm19", ode__rtol=1e-6, ode__atol=1e-10)
    v = lambda t, c0, nc: reactor.array4(t[c0:c0 + nc], lo_fab, None, nc)  # noqa: E731
    r.react_box(lo, hi, v(F, 0, N), reactor.null_real4(), v(F, N + 1, 1), v(F, N, 1), reactor.null_real4(),
                reactor.array4(FC, lo_fab, None, 1), reactor.int4(None, lo_fab), d["dt"], 0.0)
    torch.cuda.synchronize()
    r.close()
    out = F.cpu().numpy()
    ref = O.Mech("drm19").rk64_react(d["rY"], d["T"], d["rE"], d["dt"], nsub=(10, 5, 500))
    got = out[inner].reshape(N + 2, -1)
    tol = 1e-6 * np.abs(ref["rY"]) + 1e-10
    assert np.max(np.abs(got[:N] - ref["rY"]) / tol) <= 1e-2
    assert np.max(np.abs(got[N + 1] - ref["T"])) <= 1e-7
    ghost = np.ones(out.shape, dtype=bool)
    ghost[inner] = False
    assert np.all(out[ghost] == 5.0)
