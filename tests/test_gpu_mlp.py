"""The network kernels (pelelm_amd/csrc/mlp.hip) on the GPU against the CPU network
oracle (oracle/mlp_ref.py), which rounds to bfloat16 at the same points the kernels
do: the forward pass, the parameter gradient of the mean-squared error (the kernels'
own backward pass), and Adam training that reduces the loss.  Sample counts are not
multiples of the 64-row tiles, so partial tiles are covered."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import mlp_ref as R  # noqa: E402
from pelelm_amd import predictor as P  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _f32(a):
    return torch.tensor(np.ascontiguousarray(a, dtype=np.float32), device=DEV)


@pytest.mark.parametrize("n_in,n_out,B", [(24, 22, 1000), (56, 54, 333), (5, 3, 64)])
def test_forward_matches_oracle(n_in, n_out, B):
    net = P.Mlp(n_in, n_out, seed=3)
    p = net.get_params().astype(np.float64)
    rng = np.random.default_rng(4)
    # non-zero biases exercise the bias path
    lay, _ = R.layout(n_in, n_out)
    for fi, fo, ow, ob in lay:
        p[ob:ob + fo] = rng.normal(size=fo) * 0.1
    net.set_params(p)
    p = net.get_params().astype(np.float64)
    x = rng.normal(size=(B, n_in))
    print(f"forward {n_in}->{n_out}, {B} rows", flush=True)
    y = net.forward(_f32(x))
    torch.cuda.synchronize()
    y = y.cpu().numpy().astype(np.float64)
    ref, _, _ = R.forward(p, x, n_out, bf16=True)
    scale = np.abs(ref).max(axis=1, keepdims=True) + 1e-6
    assert np.max(np.abs(y - ref) / scale) < 2e-2
    assert np.median(np.abs(y - ref) / scale) < 1e-3
    net.close()


@pytest.mark.parametrize("n_in,n_out,B", [(24, 22, 700), (56, 54, 200)])
def test_gradient_matches_oracle(n_in, n_out, B):
    net = P.Mlp(n_in, n_out, seed=5)
    p = net.get_params().astype(np.float64)
    rng = np.random.default_rng(6)
    x = rng.normal(size=(B, n_in))
    t = rng.normal(size=(B, n_out))
    print(f"gradient {n_in}->{n_out}, {B} rows", flush=True)
    g, loss = net.grad(_f32(x), _f32(t))
    torch.cuda.synchronize()
    g = g.cpu().numpy().astype(np.float64)
    lref, gref = R.loss_and_grad(p, x, t, bf16=True)
    _, gexact = R.loss_and_grad(p, x, t, bf16=False)
    assert loss == pytest.approx(lref, rel=1e-2)
    cos = g @ gref / (np.linalg.norm(g) * np.linalg.norm(gref))
    assert cos > 0.9999, cos
    assert np.linalg.norm(g - gref) / np.linalg.norm(gref) < 1e-2
    cos_exact = g @ gexact / (np.linalg.norm(g) * np.linalg.norm(gexact))
    assert cos_exact > 0.999, cos_exact
    # every layer's weights and biases, not just the total
    for fi, fo, ow, ob in R.layout(n_in, n_out)[0]:
        for a, b in ((ow, ob), (ob, ob + fo)):
            seg, sref = g[a:b], gref[a:b]
            assert np.linalg.norm(seg - sref) <= 2e-2 * np.linalg.norm(sref) + 1e-9
    net.close()


def test_adam_training_reduces_loss():
    rng = np.random.default_rng(7)
    n_in, n_out, n = 24, 22, 16384
    x = rng.normal(size=(n, n_in))
    A = rng.normal(size=(n_in, n_out)) / np.sqrt(n_in)
    t = np.tanh(x @ A) + 0.1 * (x[:, :n_out] ** 2 - 1.0)
    net = P.Mlp(n_in, n_out, seed=8)
    X, T = _f32(x), _f32(t)
    print("training 30 epochs on 16384 samples", flush=True)
    hist = P.train(net, X, T, epochs=30, batch=1024, lr=3e-3, lr_final=3e-4)
    torch.cuda.synchronize()
    assert hist[-1] < 0.1 * hist[0], hist
    y = net.forward(X).cpu().numpy()
    assert np.mean((y - t) ** 2) < 0.2 * np.mean(t ** 2)
    net.close()


def test_params_round_trip():
    net = P.Mlp(24, 22, seed=9)
    p = np.random.default_rng(10).normal(size=net.nparams).astype(np.float32)
    net.set_params(p)
    np.testing.assert_array_equal(net.get_params(), p)
    net.close()
