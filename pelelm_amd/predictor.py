"""The learned Newton predictor over the C ABI ``include/pelelm_mlp.h``.

``Mlp`` wraps one network (n_in -> 128 -> 128 -> 128 -> n_out, GELU, BF16 matrix-core
products with FP32 accumulation) living on the GPU; ``train`` runs the GPU trainer
(Adam on mean-squared error) over device tensors.  Feature and target definitions
for the reactor are in ``features`` / ``targets``: per cell the inputs are
(T, log p, log10 Y_k, log dt) and the outputs (dT, dY_k) over dt.  The reference has
no learned component; nothing here replaces a reference interface.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from .reactor import lib

_vp = C.c_void_p


def _bind():
    L = lib()
    if getattr(L, "_mlp_bound", False):
        return L
    L.mi_mlp_create.argtypes = [C.c_int, C.c_int, C.c_ulonglong, C.POINTER(_vp)]
    L.mi_mlp_nparams.argtypes = [_vp]
    L.mi_mlp_get_params.argtypes = [_vp, _vp, C.c_int]
    L.mi_mlp_set_params.argtypes = [_vp, _vp, C.c_int]
    L.mi_mlp_set_normalisation.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.mi_mlp_forward.argtypes = [_vp, _vp, _vp, C.c_long, _vp]
    L.mi_mlp_grad.argtypes = [_vp, _vp, _vp, C.c_long, _vp, C.POINTER(C.c_float), _vp]
    L.mi_mlp_train_step.argtypes = [_vp, _vp, _vp, C.c_long, C.c_float, C.POINTER(C.c_float), _vp]
    L.mi_mlp_last_error.argtypes = [_vp]
    L.mi_mlp_last_error.restype = C.c_char_p
    L.mi_mlp_close.argtypes = [_vp]
    L.mi_chem_attach_predictor.argtypes = [_vp, _vp]
    for f in ("mi_mlp_create", "mi_mlp_nparams", "mi_mlp_get_params", "mi_mlp_set_params",
              "mi_mlp_set_normalisation", "mi_mlp_forward", "mi_mlp_grad", "mi_mlp_train_step", "mi_mlp_close",
              "mi_chem_attach_predictor"):
        getattr(L, f).restype = C.c_int
    L._mlp_bound = True
    return L


MLP_EXPORTS = ["mi_chem_attach_predictor", "mi_mlp_close", "mi_mlp_create", "mi_mlp_forward", "mi_mlp_get_params",
               "mi_mlp_grad", "mi_mlp_last_error", "mi_mlp_nparams", "mi_mlp_set_normalisation",
               "mi_mlp_set_params", "mi_mlp_train_step"]


def attach(reactor, net) -> None:
    """Make ``net`` the Newton predictor of ``reactor`` (a ``ReactorBase``); None detaches."""
    rc = _bind().mi_chem_attach_predictor(reactor._h, net.handle if net is not None else None)
    if rc != 0:
        raise MlpError(f"{rc}: {lib().mi_chem_last_error(reactor._h).decode()}")


class MlpError(RuntimeError):
    pass


def _stream():
    import torch
    return _vp(torch.cuda.current_stream().cuda_stream)


def _dptr(t):
    import torch
    if not (t.is_cuda and t.is_contiguous() and t.dtype == torch.float32):
        raise ValueError("network arrays must be contiguous float32 device tensors")
    return _vp(t.data_ptr())


class Mlp:
    def __init__(self, n_in: int, n_out: int, seed: int = 1):
        L = _bind()
        h = _vp()
        rc = L.mi_mlp_create(n_in, n_out, seed, C.byref(h))
        if rc != 0:
            raise MlpError(f"mi_mlp_create failed ({rc})")
        self._h, self.n_in, self.n_out = h, n_in, n_out

    @property
    def handle(self):
        return self._h

    def _check(self, rc):
        if rc != 0:
            raise MlpError(f"{rc}: {_bind().mi_mlp_last_error(self._h).decode()}")

    @property
    def nparams(self) -> int:
        return _bind().mi_mlp_nparams(self._h)

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.nparams, dtype=np.float32)
        self._check(_bind().mi_mlp_get_params(self._h, p.ctypes.data_as(_vp), p.size))
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(_bind().mi_mlp_set_params(self._h, p.ctypes.data_as(_vp), p.size))

    def set_normalisation(self, in_mean, in_std, out_mean, out_std) -> None:
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (in_mean, in_std, out_mean, out_std)]
        self.norm = arrs
        self._check(_bind().mi_mlp_set_normalisation(self._h, *[a.ctypes.data_as(_vp) for a in arrs]))

    def forward(self, X):
        import torch
        Y = torch.empty((X.shape[0], self.n_out), dtype=torch.float32, device=X.device)
        self._check(_bind().mi_mlp_forward(self._h, _dptr(X), _dptr(Y), X.shape[0], _stream()))
        return Y

    def grad(self, X, T):
        import torch
        g = torch.empty(self.nparams, dtype=torch.float32, device=X.device)
        loss = C.c_float()
        self._check(_bind().mi_mlp_grad(self._h, _dptr(X), _dptr(T), X.shape[0], _dptr(g), C.byref(loss),
                                        _stream()))
        return g, loss.value

    def train_step(self, X, T, lr: float, want_loss: bool = False):
        loss = C.c_float()
        self._check(_bind().mi_mlp_train_step(self._h, _dptr(X), _dptr(T), X.shape[0], lr,
                                              C.byref(loss) if want_loss else None, _stream()))
        return loss.value if want_loss else None

    def close(self):
        if self._h:
            _bind().mi_mlp_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def train(net: Mlp, X, T, epochs: int = 10, batch: int = 8192, lr: float = 1e-3, lr_final: float = 1e-4,
          seed: int = 0, log=None):
    """Adam over shuffled mini-batches of the device tensors X [n][n_in], T [n][n_out]
    (already normalised), learning rate decaying geometrically from lr to lr_final.
    Returns the mean loss of each epoch."""
    import torch
    n = X.shape[0]
    gen = torch.Generator(device=X.device)
    gen.manual_seed(seed)
    steps = epochs * max(1, n // batch)
    decay = (lr_final / lr) ** (1.0 / max(1, steps - 1))
    hist, k = [], 0
    for ep in range(epochs):
        perm = torch.randperm(n, device=X.device, generator=gen)
        tot, cnt = 0.0, 0
        for s in range(0, n - batch + 1, batch):
            idx = perm[s:s + batch]
            xb, tb = X[idx].contiguous(), T[idx].contiguous()
            want = (s // batch) % 16 == 0
            loss = net.train_step(xb, tb, lr * decay ** k, want_loss=want)
            k += 1
            if want:
                tot += loss
                cnt += 1
        hist.append(tot / max(1, cnt))
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {hist[-1]:.4g}")
    return hist


# ---------------------------------------------------------------------------
# reactor features: (T, log p, log10 Y_k, log h) -> (dT, dY_k); the same definitions as
# the reactor's predictor kernel (p = rho R T sum_k Y_k / W_k, Y floored at 1e-20)
# ---------------------------------------------------------------------------
Y_FLOOR = 1e-20


def _mech_W(mech: str) -> np.ndarray:
    from .mechanisms import mechanism
    return np.array(mechanism.load(mech).W)


def features(mech: str, rY: np.ndarray, T: np.ndarray, h) -> np.ndarray:
    """Raw (unnormalised) inputs [ncell][N + 3] from SoA rY [N][ncell], T, step h."""
    from .mechanisms import mechanism
    W = _mech_W(mech)
    rho = rY.sum(axis=0)
    Y = rY / rho
    n = T.shape[0]
    h = np.broadcast_to(np.asarray(h, dtype=np.float64), (n,))
    p = rho * mechanism.R_CGS * T * (Y / W[:, None]).sum(axis=0)
    return np.concatenate([T[:, None], np.log(p)[:, None], np.log10(np.maximum(Y, Y_FLOOR)).T,
                           np.log(h)[:, None]], axis=1)


def targets(rY0, T0, rY1, T1) -> np.ndarray:
    """Raw outputs [ncell][N + 1]: (T1 - T0, Y1 - Y0) with Y = rY / rho (rho is constant)."""
    rho = rY0.sum(axis=0)
    return np.concatenate([(T1 - T0)[:, None], ((rY1 - rY0) / rho).T], axis=1)


def generate_pairs(workload: str = "drm19_128_mix", n: int = 1_000_000, seed: int = 9104, groups: int = 32,
                   h_range=(1e-10, 1e-5), tau_range=(1e-9, 1e-5), rtol: float = 1e-6, atol: float = 1e-10,
                   log=None):
    """Training pairs from the plain FP64 integrator on the GPU: seeded cells of the
    workload's distribution (a seed of their own) are first advanced by a random time
    tau (so intermediate integrator states are represented), then by a step h; the pair
    is (features of the state after tau, h) -> increment over h.  Cells form `groups`
    groups, each with its own log-uniform tau and h.  Returns (X, T) raw numpy arrays."""
    import torch

    from . import cells
    from .reactor import ReactorBase
    w = cells.WORKLOADS[workload]
    d = cells.generate(w.mech, w.fuel, w.o2_per_fuel, w.phi, w.mode, seed, n)
    rng = np.random.default_rng(seed)
    taus = np.exp(rng.uniform(math.log(tau_range[0]), math.log(tau_range[1]), groups))
    hs = np.exp(rng.uniform(math.log(h_range[0]), math.log(h_range[1]), groups))
    r = ReactorBase.create("ReactorCvode")
    r.set("mi.mechanism", w.mech)
    r.set("ode.rtol", rtol)
    r.set("ode.atol", atol)
    r.init(2, 1)
    m = n // groups
    X, Tt = [], []
    for g in range(groups):
        sl = slice(g * m, (g + 1) * m)
        y = torch.tensor(np.ascontiguousarray(d["rY"][:, sl]), device="cuda")
        t = torch.tensor(d["T"][sl], device="cuda")
        e = torch.tensor(d["rE"][sl], device="cuda")
        fc = torch.zeros(m, dtype=torch.float64, device="cuda")
        r.react(y, None, t, e, None, fc, None, float(taus[g]), 0.0)
        y0, t0 = y.clone(), t.clone()
        r.react(y, None, t, e, None, fc, None, float(hs[g]), 0.0)
        torch.cuda.synchronize()
        y0n, t0n, y1n, t1n = y0.cpu().numpy(), t0.cpu().numpy(), y.cpu().numpy(), t.cpu().numpy()
        X.append(features(w.mech, y0n, t0n, hs[g]))
        Tt.append(targets(y0n, t0n, y1n, t1n))
        if log and (g + 1) % 8 == 0:
            log(f"pairs: group {g + 1}/{groups} (tau {taus[g]:.2e} s, h {hs[g]:.2e} s)")
    r.close()
    return np.concatenate(X), np.concatenate(Tt)


def fit(mech: str, X: np.ndarray, T: np.ndarray, epochs: int = 10, batch: int = 8192, lr: float = 2e-3,
        lr_final: float = 1e-4, seed: int = 1, log=None) -> Mlp:
    """Normalise the pairs, train a network on the GPU and set its normalisation."""
    import torch
    in_mean, in_std = X.mean(0), np.maximum(X.std(0), 1e-3)
    out_mean, out_std = T.mean(0), np.maximum(T.std(0), 1e-30)
    Xn = torch.tensor(((X - in_mean) / in_std).astype(np.float32), device="cuda")
    Tn = torch.tensor(((T - out_mean) / out_std).astype(np.float32), device="cuda")
    net = Mlp(X.shape[1], T.shape[1], seed=seed)
    net.losses = train(net, Xn, Tn, epochs=epochs, batch=batch, lr=lr, lr_final=lr_final, seed=seed, log=log)
    net.set_normalisation(in_mean, in_std, out_mean, out_std)
    return net
