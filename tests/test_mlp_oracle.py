"""Pins the CPU FP64 oracle of the warm-start network (oracle/mlp_oracle.py)
without a GPU: bf16 rounding against torch's own float -> bfloat16
conversion, the unrounded backward pass against torch FP64 autograd (an
independent implementation) and against central finite differences, and a
single linear layer against the closed-form least-squares gradient."""
import sys
from pathlib import Path

import numpy as np
import pytest
import torch

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
import mlp_oracle as mo  # noqa: E402


def _net(dims, seed):
    g = np.random.default_rng(seed)
    return [(g.standard_normal((dims[l + 1], dims[l])) / np.sqrt(dims[l]), 0.1 * g.standard_normal(dims[l + 1]))
            for l in range(len(dims) - 1)]


def test_bf16_rounding_matches_torch():
    g = np.random.default_rng(0)
    v = np.concatenate([g.standard_normal(20000) * 10.0 ** g.integers(-30, 30, 20000),
                        [0.0, -0.0, 1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 65504.0, 3.0e38, 1e-40]])
    ours = mo.bf16(v)
    theirs = torch.from_numpy(v.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(ours, theirs)
    # ties to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7
    assert mo.bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0
    assert mo.bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6


@pytest.mark.parametrize("dims", [[8, 16, 16, 8], [5, 7, 3]])
def test_unrounded_backward_matches_torch_autograd(dims):
    params = _net(dims, 1)
    g = np.random.default_rng(2)
    x, t = g.standard_normal((11, dims[0])), g.standard_normal((11, dims[-1]))
    loss, dW, db = mo.loss_and_grads(params, x, t, rounding=False)
    tp = [(torch.tensor(W, requires_grad=True), torch.tensor(b, requires_grad=True)) for W, b in params]
    h = torch.tensor(x)
    for l, (W, b) in enumerate(tp):
        h = h @ W.T + b
        if l < len(tp) - 1:
            h = torch.relu(h)
    tl = torch.mean((h - torch.tensor(t)) ** 2)
    tl.backward()
    assert abs(loss - tl.item()) <= 1e-12 * (1 + abs(loss))
    for l, (W, b) in enumerate(tp):
        assert np.allclose(dW[l], W.grad.numpy(), rtol=1e-10, atol=1e-14)
        assert np.allclose(db[l], b.grad.numpy(), rtol=1e-10, atol=1e-14)


def test_unrounded_backward_matches_finite_differences():
    dims = [6, 12, 12, 4]
    params = _net(dims, 3)
    g = np.random.default_rng(4)
    x, t = g.standard_normal((9, dims[0])), g.standard_normal((9, dims[-1]))
    _, dW, db = mo.loss_and_grads(params, x, t, rounding=False)
    for layer in range(3):
        for idx in [(0, 0), (dims[layer + 1] - 1, dims[layer] - 1), (2, 1)]:
            fd = mo.finite_difference(params, x, t, layer, idx)
            assert abs(fd - dW[layer][idx]) <= 1e-6 * (1 + abs(fd))
        fd = mo.finite_difference(params, x, t, layer, 1)
        assert abs(fd - db[layer][1]) <= 1e-6 * (1 + abs(fd))


def test_single_linear_layer_closed_form():
    g = np.random.default_rng(5)
    n, din, dout = 13, 7, 5
    W, b = g.standard_normal((dout, din)), g.standard_normal(dout)
    x, t = g.standard_normal((n, din)), g.standard_normal((n, dout))
    loss, dW, db = mo.loss_and_grads([(W, b)], x, t, rounding=False)
    r = x @ W.T + b - t
    assert np.isclose(loss, np.sum(r * r) / (n * dout))
    assert np.allclose(dW[0], 2.0 / (n * dout) * r.T @ x)
    assert np.allclose(db[0], 2.0 / (n * dout) * r.sum(0))


def test_rounded_gradients_stay_near_the_exact_ones():
    dims = [64, 128, 128, 64]
    params = _net(dims, 6)
    g = np.random.default_rng(7)
    x, t = g.standard_normal((64, dims[0])), g.standard_normal((64, dims[-1]))
    _, dW, _ = mo.loss_and_grads(params, x, t, rounding=True)
    _, dW0, _ = mo.loss_and_grads(params, x, t, rounding=False)
    for a, b in zip(dW, dW0):
        assert np.linalg.norm(a - b) <= 0.1 * np.linalg.norm(b)   # bf16 through two backward GEMMs: ~5%
