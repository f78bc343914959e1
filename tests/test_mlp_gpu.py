"""The warm-start network's gfx950 kernels against the CPU FP64 oracle
(oracle/mlp_oracle.py, pinned in test_mlp_oracle.py).

  GEMM: every transpose layout against FP64 products of the same
        bf16-rounded operands, |C - C_ref| <= 2^-20 (|A||B|) (FP32 accumulation).
  forward: against the exact FP64 network within a propagated bf16 bound
        e_0 = u|x|, e_l = |W_l| e_{l-1} + 2u |W_l||H_{l-1}| (+ u|H_l| after rounding),
        u = 2^-8; and against the oracle with the same rounding points to 1e-2
        relative (rounding flips of hidden activations).
  gradients: against the oracle with the same rounding points (relative
        Frobenius error <= 2e-2 per tensor, cosine >= 0.999) and against the
        exact FP64 gradients (<= 0.1).
  training: Adam on a teacher-network regression lowers the loss, to within
        25% of an FP32 torch run of the same recipe from the same weights."""
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
import mlp_oracle as mo  # noqa: E402
from optimisation_solver_amd.warmstart.mlp import Mlp, MlpError  # noqa: E402

pytestmark = pytest.mark.gpu
U = 2.0 ** -8


@pytest.mark.parametrize("ta", [False, True])
@pytest.mark.parametrize("tb", [False, True])
@pytest.mark.parametrize("shape", [(64, 64, 32), (128, 192, 96), (64, 128, 224), (128, 64, 256), (192, 64, 480),
                                   (256, 512, 512)])
def test_gemm_layouts(shape, ta, tb):
    """Every layout, K within one LDS chunk (224) and across chunk boundaries."""
    M, N, K = shape
    g = np.random.default_rng(M + N + K)
    A = g.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = g.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    print(f"gemm {shape} ta={ta} tb={tb}", flush=True)
    with Mlp([64, 64], 64) as net:
        C, _ = net.gemm(A, B, ta, tb)
    Ar, Br = mo.bf16(A), mo.bf16(B)
    Am = Ar.T if ta else Ar
    Bm = Br.T if tb else Br
    ref = Am @ Bm
    bound = 2.0 ** -20 * (np.abs(Am) @ np.abs(Bm)) + 1e-30
    assert np.all(np.abs(C - ref) <= bound)


def _setup(dims, batch, seed):
    g = np.random.default_rng(seed)
    params = [((g.standard_normal((dims[l + 1], dims[l])) / np.sqrt(dims[l])).astype(np.float32),
               (0.1 * g.standard_normal(dims[l + 1])).astype(np.float32)) for l in range(len(dims) - 1)]
    x = g.standard_normal((batch, dims[0])).astype(np.float32)
    t = g.standard_normal((batch, dims[-1])).astype(np.float32)
    return params, x, t


DIMS = [256, 512, 512, 512]


def test_forward_within_bf16_bound():
    params, x, t = _setup(DIMS, 200, 11)   # 200 rows: forward pads to 256 internally
    with Mlp(DIMS, 256) as net:
        net.set_params(params)
        y = net.forward(x)
    y_exact, H = mo.forward(params, x, rounding=False)   # H[l] = exact input of layer l
    y_round, _ = mo.forward(params, x, rounding=True)
    # Propagated bound on |y - y_exact|: bf16 rounding of x, W and the hidden
    # activations (u each), FP32 accumulation over K terms (K 2^-24), ReLU
    # 1-Lipschitz.
    e = U * np.abs(x.astype(np.float64))
    L = len(params)
    for l, (W, b) in enumerate(params):
        Wa = np.abs(W.astype(np.float64)).T
        K = W.shape[1]
        h = np.abs(H[l])
        ez = (1 + U) * (e @ Wa) + U * (h @ Wa) + (K + 1) * 2.0 ** -24 * ((h + e) @ Wa + np.abs(b))
        e = ez + U * (np.abs(H[l + 1]) + ez) if l < L - 1 else ez
    assert np.all(np.abs(y - y_exact) <= e)
    assert np.linalg.norm(y - y_round) <= 1e-2 * np.linalg.norm(y_round)


def test_gradients_match_oracle():
    params, x, t = _setup(DIMS, 256, 12)
    with Mlp(DIMS, 256) as net:
        net.set_params(params)
        loss, dW, db = net.gradients(x, t)
    l_r, dW_r, db_r = mo.loss_and_grads(params, x, t, rounding=True)
    l_e, dW_e, db_e = mo.loss_and_grads(params, x, t, rounding=False)
    print(f"loss gpu {loss:.6g} oracle {l_r:.6g} exact {l_e:.6g}", flush=True)
    assert abs(loss - l_r) <= 1e-2 * l_r
    for l in range(3):
        for got, ref, ex in ((dW[l], dW_r[l], dW_e[l]), (db[l], db_r[l], db_e[l])):
            rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            cos = float(np.sum(got * ref) / (np.linalg.norm(got) * np.linalg.norm(ref)))
            rel_e = np.linalg.norm(got - ex) / np.linalg.norm(ex)
            print(f"  layer {l} {got.shape}: rel vs oracle {rel:.2e} cos {cos:.6f} rel vs exact {rel_e:.2e}",
                  flush=True)
            assert rel <= 2e-2 and cos >= 0.999 and rel_e <= 0.1


def _torch_reference_training(params, x, t, epochs, batch, lr, final_lr):
    """The same Adam recipe (cosine learning rate, bias correction) in FP32
    torch on the CPU from the same initial weights (its own batch order)."""
    import math

    import torch
    torch.manual_seed(0)
    Ws = [torch.tensor(W, requires_grad=True) for W, _ in params]
    bs = [torch.tensor(b, requires_grad=True) for _, b in params]
    opt = torch.optim.Adam(Ws + bs, lr=lr, betas=(0.9, 0.999), eps=1e-8)
    X, T = torch.tensor(x), torch.tensor(t)
    steps = x.shape[0] // batch
    last = None
    for e in range(epochs):
        perm = torch.randperm(x.shape[0])
        tot = 0.0
        for k in range(steps):
            frac = (e * steps + k) / (epochs * steps)
            for gp in opt.param_groups:
                gp["lr"] = final_lr + 0.5 * (lr - final_lr) * (1 + math.cos(math.pi * frac))
            idx = perm[k * batch:(k + 1) * batch]
            h = X[idx]
            for l in range(len(Ws)):
                h = h @ Ws[l].T + bs[l]
                if l < len(Ws) - 1:
                    h = torch.relu(h)
            loss = torch.mean((h - T[idx]) ** 2)
            opt.zero_grad()
            loss.backward()
            opt.step()
            tot += loss.item()
        last = tot / steps
    return last


def test_training_lowers_the_loss_like_a_reference_implementation():
    g = np.random.default_rng(13)
    dims = [256, 512, 512, 512]
    teacher, _, _ = _setup([256, 256, 512], 1, 14)
    x = g.standard_normal((2048, 256)).astype(np.float32)
    t, _ = mo.forward(teacher, x, rounding=False)
    t = (t / t.std()).astype(np.float32)
    init, _, _ = _setup(dims, 1, 15)
    with Mlp(dims, 256, seed=3) as net:
        net.set_params(init)
        net.set_data(x, t)
        out = net.train(40, 256, lr=1e-3, final_lr=1e-5, seed=9)
        y = net.forward(x[:256])
    first, last = out["epoch_loss"][0], out["epoch_loss"][-1]
    ref = _torch_reference_training(init, x, t, 40, 256, 1e-3, 1e-5)
    print(f"loss epoch 1 {first:.4g} -> {last:.4g} (torch FP32 reference {ref:.4g}) in {out['seconds']:.3f}s "
          f"({out['gemm_flops'] / out['seconds'] / 1e12:.2f} TFLOP/s on the matrix cores)", flush=True)
    assert last <= 0.4 * first
    assert abs(last - ref) <= 0.25 * ref
    assert np.mean((y - t[:256]) ** 2) <= 0.5 * first


def test_invalid_shapes_rejected():
    with pytest.raises(MlpError):
        Mlp([100, 64])
    with Mlp([64, 64], 64) as net:
        with pytest.raises(MlpError):
            net.gradients(np.zeros((32, 64)), np.zeros((32, 64)))   # batch not a multiple of 64
        with pytest.raises(MlpError):
            net.gemm(np.zeros((64, 48)), np.zeros((48, 64)))       # K not a multiple of 32
