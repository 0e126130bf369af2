"""The persistent schedule (mi.schedule = persistent) against the CPU oracle.

One kernel per react call: waves keep cells in their lanes and claim new ones from a
device queue, so a cell runs all its trips without a host round trip.  Few waves are
used here so that every lane slot integrates many cells in turn (refills after
uneven finishes are what the schedule adds).  The FP64 form runs the oracle's algorithm
step for step: the same bounds as the other schedules (1e-2 of the tolerance, 1e-7 K,
the oracle's RHS count in >= 99 % of the cells).  Every call goes through the C ABI.
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
    r.set("mi.schedule", "persistent")
    for k, v in keys.items():
        r.set(k.replace("__", "."), v)
    r.init(2, 1)
    return r


def _dev(a, dtype=torch.float64):
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def _parity(label, gY, gT, gFC, ref, y_frac=1e-2, dT=1e-7, fc=0.99):
    tol = 1e-6 * np.abs(ref["rY"]) + 1e-10
    ey = float((np.abs(gY - ref["rY"]) / tol).max())
    et = float(np.abs(gT - ref["T"]).max())
    same = float(np.mean(gFC == ref["FC"])) if gFC is not None else 1.0
    print(f"{label}: max |dY|/tol {ey:.3g}, max |dT| {et:.3g} K, identical FC {same:.4f}")
    assert ey <= y_frac and et <= dT and same >= fc, (label, ey, et, same)


@pytest.mark.parametrize("wl,mech,n,waves,chunk", [("drm19_128_mix", "drm19", 8192, 8, 1),
                                                   ("drm19_128_mix", "drm19", 8192, 16, 16),
                                                   ("lidryer_16", "LiDryer", 4096, 4, 64),
                                                   ("dodecane53_64_mix", "dodecane53", 2048, 8, 4)])
@pytest.mark.parametrize("lu", ["fp64", "fp32_ir"])
def test_persistent_matches_oracle(wl, mech, n, waves, chunk, lu):
    d = cells.generate_workload(wl, ncell=n)
    ref = O.Mech(mech).react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10)
    r = _mk(mech, ode__rtol=1e-6, ode__atol=1e-10, mi__persist_waves=waves, mi__persist_chunk=chunk, mi__lu=lu)
    rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
    FC = torch.zeros(n, dtype=torch.float64, device=DEV)
    dt, tm = r.react(rY, None, T, rE, None, FC, None, d["dt"], 0.0)
    torch.cuda.synchronize()
    st = r.stats()
    r.close()
    print(f"{wl} waves {waves} chunk {chunk} {lu}: {st['ntrips']} iterations of the longest wave, "
          f"{n * 1.0 / (waves * 64):.1f} cells per slot, {st['ms_kernel']:.2f} ms")
    assert st["nfail"] == 0 and st["ncells"] == n and st["schedule_used"] == 3
    assert tm == pytest.approx(d["dt"])
    gY, gT, gE = rY.cpu().numpy(), T.cpu().numpy(), rE.cpu().numpy()
    if lu == "fp64":
        _parity(f"{wl} persistent", gY, gT, FC.cpu().numpy(), ref)
    else:
        assert st["nlu_refine"] > 0
        _parity(f"{wl} persistent fp32_ir", gY, gT, FC.cpu().numpy(), ref, y_frac=1.0, dT=1e-3, fc=0.95)
    np.testing.assert_array_equal(gE, ref["rE"])
    np.testing.assert_allclose(gY.sum(0), d["rY"].sum(0), rtol=1e-12)
    # the counters agree with the per-cell counts
    assert st["nfe"] == int(FC.sum().item())


def test_persistent_forcing_mask_and_default_grid():
    """Forcing, masked cells and the default grid (as many waves as the device holds
    resident, capped by the cell count)."""
    n = 4096
    d = cells.generate_workload("drm19_64", ncell=n)
    rng = np.random.default_rng(5)
    fY = rng.normal(size=d["rY"].shape) * 1e-3 * d["rY"]
    fH = rng.normal(size=n) * 1e6
    mask = np.ones(n, dtype=np.int32)
    mask[::7] = -1
    ref = O.Mech("drm19").react(d["rY"], d["T"], d["rE"], d["dt"], rYsrc=fY, rEsrc=fH, mask=mask, rtol=1e-6,
                                atol=1e-10)
    r = _mk("drm19", ode__rtol=1e-6, ode__atol=1e-10)
    rY, T, rE = _dev(d["rY"]), _dev(d["T"]), _dev(d["rE"])
    FC = torch.full((n,), 5.0, dtype=torch.float64, device=DEV)
    r.react(rY, _dev(fY), T, rE, _dev(fH), FC, _dev(mask, torch.int32), d["dt"], 0.0)
    torch.cuda.synchronize()
    st = r.stats()
    r.close()
    act = mask >= 0
    assert st["nfail"] == 0 and st["ncells"] == act.sum()
    gY, gT, gE, gFC = rY.cpu().numpy(), T.cpu().numpy(), rE.cpu().numpy(), FC.cpu().numpy()
    np.testing.assert_array_equal(gFC[~act], 0.0)
    np.testing.assert_array_equal(gY[:, ~act], d["rY"][:, ~act])
    sub = {"rY": ref["rY"][:, act], "T": ref["T"][act], "FC": ref["FC"][act]}
    _parity("drm19_64 persistent forced+masked", gY[:, act], gT[act], gFC[act], sub)
    np.testing.assert_allclose(gE[act], d["rE"][act] + d["dt"] * fH[act], rtol=1e-15)


def test_persistent_react_box_on_array4_views():
    nv, ng = 12, 2
    nf = nv + 2 * ng
    d = cells.generate_workload("drm19_128_mix", ncell=nf ** 3)
    N = d["rY"].shape[0]
    fab = np.concatenate([d["rY"], d["rE"][None], d["T"][None]], axis=0).reshape(N + 2, nf, nf, nf)
    S = _dev(fab)
    FCf = torch.zeros((1, nf, nf, nf), dtype=torch.float64, device=DEV)
    lo_fab = (-ng, -ng, -ng)
    r = _mk("drm19", ode__rtol=1e-6, ode__atol=1e-10, mi__persist_waves=4)
    v = lambda t, nc: reactor.array4(t, lo_fab, None, nc)  # noqa: E731
    z = reactor.null_real4()
    r.react_box((0, 0, 0), (nv - 1, nv - 1, nv - 1), v(S[:N], N), z, v(S[N + 1:N + 2], 1), v(S[N:N + 1], 1), z,
                v(FCf, 1), reactor.int4(None, lo_fab), d["dt"], 0.0)
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
    _parity("drm19 persistent Array4 box", got[:N], got[N + 1], FCf.cpu().numpy()[0][ng:ng + nv, ng:ng + nv,
                                                                                   ng:ng + nv].ravel(), ref)


def test_persistent_react_captured_in_a_graph():
    """With mi.on_fail = continue a persistent react call enqueues kernels only (no host
    synchronisation, no allocation after the first call of its size), so a caller can
    capture it into a HIP graph.  Replayed three times, with eager work and
    synchronisations between the replays, every replay must give the eager result."""
    n = 4096
    d = cells.generate_workload("drm19_128_mix", ncell=n)
    r = _mk("drm19", ode__rtol=1e-6, ode__atol=1e-10, mi__on_fail="continue", mi__persist_waves=16)
    src = {k: _dev(d[k]) for k in ("rY", "T", "rE")}
    rY, T, rE = (torch.empty_like(src[k]) for k in ("rY", "T", "rE"))
    FC = torch.zeros(n, dtype=torch.float64, device=DEV)

    def step():
        rY.copy_(src["rY"])
        T.copy_(src["T"])
        rE.copy_(src["rE"])
        r.react(rY, None, T, rE, None, FC, None, d["dt"], 0.0)

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step()                                  # eager: sizes the workspace
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    eager = (rY.clone(), T.clone(), FC.clone())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step()
    for k in range(3):
        rY.zero_()
        T.zero_()
        FC.fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(rY, eager[0]) and torch.equal(T, eager[1]) and torch.equal(FC, eager[2]), k
        _ = (src["T"] * 2.0).sum().item()      # eager work between replays
    st = r.stats()
    r.close()
    assert st["nfail"] == 0 and st["ncells"] == n
