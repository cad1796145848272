"""Corrector network kernels on the GPU against an FP32 torch reference (CPU)
with the same BF16 roundings (tests/nn_ref.py).

Floating-point kernels with BF16 storage: the bar is a tolerance, written in
each test (MFMA accumulation order differs from torch's, and the stored BF16
activations can round the other way after a last-bit FP32 difference).
"""
import numpy as np
import pytest

import nn_ref

pytestmark = pytest.mark.gpu


def _bf16_round(a):
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(3 * nn_ref.LAYER, np.float32)
    for l, (cin, cout) in enumerate(((12, 16), (16, 16), (16, 12))):
        w = np.zeros((27, 16, 16), np.float32)
        w[:, :cout, :cin] = rng.standard_normal((27, cout, cin)) * scale * np.sqrt(2.0 / (27 * cin))
        b = np.zeros(16, np.float32)
        b[:cout] = rng.standard_normal(cout) * 0.1
        p[l * nn_ref.LAYER:(l + 1) * nn_ref.LAYER] = np.concatenate([w.ravel(), b])
    return p


def _input(shape, seed):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    x = np.zeros((nz, ny, nx, 16), np.float32)
    x[..., :12] = rng.standard_normal((nz, ny, nx, 12))
    return x


def test_mfma_operand_layout(gd):
    rng = np.random.default_rng(0)
    a = _bf16_round(rng.standard_normal((16, 16)))
    b = _bf16_round(rng.standard_normal((16, 16)))
    d = gd.mfma_selftest(a, b)
    want = a.astype(np.float64) @ b.astype(np.float64)
    # products of BF16 values are exact in FP32; only the 16-term sum rounds
    assert np.max(np.abs(d - want)) <= 1e-5 * np.max(np.abs(want))
    assert not np.allclose(d, a @ b.T, atol=1e-3)  # a transposed layout would fail


@pytest.mark.parametrize("shape", [(37, 21, 13), (16, 16, 16), (5, 9, 7)])
def test_forward_matches_torch(gd, shape):
    p = _params(1)
    x = _input(shape, 2)
    with gd.Net(*shape) as net:
        net.set_params(p)
        net.set_input(x)
        y = net.forward()
    ref, _ = nn_ref.forward(p, x)
    ref = ref.detach().numpy()
    # tolerance: BF16 activations between layers (relative 2^-8 per rounding)
    err = np.max(np.abs(y[..., :12] - ref[..., :12])) / np.max(np.abs(ref[..., :12]))
    assert err < 2e-2, err


def test_gradients_match_torch(gd):
    shape = (29, 18, 11)
    nx, ny, nz = shape
    p = _params(3)
    x = _input(shape, 4)
    rng = np.random.default_rng(5)
    t = np.zeros((nz, ny, nx, 16), np.float32)
    t[..., :12] = rng.standard_normal((nz, ny, nx, 12))
    mask = (rng.random((nz, ny, nx)) < 0.8).astype(np.uint8)
    with gd.Net(*shape) as net:
        net.set_params(p)
        net.set_input(x)
        net.set_target(t, mask)
        loss, g = net.train_step(0.0, want_grads=True)
        assert np.array_equal(net.get_params(), p)  # lr 0: no update
    rloss, rg = nn_ref.loss_and_grads(p, x, t, mask)
    assert abs(loss - rloss) <= 1e-2 * rloss
    # tolerance: BF16 gradients between layers (relative 2^-8 per rounding)
    for l in range(3):
        s = slice(l * nn_ref.LAYER, (l + 1) * nn_ref.LAYER)
        rel = np.linalg.norm(g[s] - rg[s]) / np.linalg.norm(rg[s])
        assert rel < 5e-2, (l, rel)


def test_training_reduces_loss(gd):
    shape = (24, 20, 16)
    nx, ny, nz = shape
    x = _input(shape, 6)
    # a target the network can represent: a fixed random 3x3x3 linear map
    tp = _params(7, scale=0.5)
    ref, _ = nn_ref.forward(tp, x)
    t = ref.detach().numpy().copy()
    t[..., 12:] = 0
    mask = np.ones((nz, ny, nx), np.uint8)
    with gd.Net(*shape) as net:
        net.set_params(_params(8, scale=0.5))
        net.set_input(x)
        net.set_target(t, mask)
        first = net.train_step(3e-3)
        for _ in range(150):
            last = net.train_step(3e-3)
    assert np.isfinite(last) and last < 0.2 * first, (first, last)
