"""The surrogate network's GPU kernels (BF16 MFMA products, FP32 accumulation) against
the FP64 oracle network: forward values, every gradient block, and training."""
import numpy as np
import pytest

import oracle_lib as O
from powder6_amd import pt

pytestmark = pytest.mark.gpu
N_IN, H = 64, 256
OFFS = np.cumsum([0, N_IN * H, H, H * H, H, H * H, H, H, 1])
BLOCKS = ['W1', 'b1', 'W2', 'b2', 'W3', 'b3', 'w4', 'b4']


def bf16(a):
    """Round to BF16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7fff + ((u >> 16) & 1)) & 0xffff0000
    return u.astype(np.uint32).view(np.float32)


def as_used(p):
    """The parameters as the kernels use them: BF16 weight matrices, FP32 biases."""
    q = np.asarray(p, dtype=np.float64).copy()
    for k in (0, 2, 4, 6):
        q[OFFS[k]:OFFS[k + 1]] = bf16(p[OFFS[k]:OFFS[k + 1]])
    return q


def _pad(x):
    xp = np.zeros((x.shape[0], N_IN))
    xp[:, :x.shape[1]] = bf16(x)
    return xp


def test_forward_matches_oracle():
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, size=(300, 53)).astype(np.float32)
    with pt.Mlp(seed=5) as net:
        p = net.params()
        y = net.forward(x)
    yo = O.mlp_forward(as_used(p), N_IN, H, _pad(x))
    err = np.sqrt(np.mean((y - yo) ** 2)) / np.sqrt(np.mean(yo ** 2))
    print('forward relative rms error vs FP64', err)
    assert err < 2e-2


def test_gradients_match_oracle():
    rng = np.random.default_rng(2)
    n = 192
    x = rng.uniform(-1, 1, size=(n, 53)).astype(np.float32)
    t = rng.normal(size=n).astype(np.float32)
    mask = (rng.uniform(size=n) > 0.1).astype(np.float32)
    with pt.Mlp(seed=6) as net:
        p = net.params()
        loss, g = net.gradient(x, t, mask)
    lo, go = O.mlp_loss_grad(as_used(p), N_IN, H, _pad(x), t.astype(np.float64), mask.astype(np.float64))
    assert abs(loss - lo) <= 2e-2 * lo, (loss, lo)
    for k, name in enumerate(BLOCKS):
        a, b = g[OFFS[k]:OFFS[k + 1]].astype(np.float64), go[OFFS[k]:OFFS[k + 1]]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        print('%s gradient: relative error %.2e, cosine %.6f' % (name, rel, cos))
        assert rel < 3e-2 and cos > 0.999, name


def test_training_fits_a_smooth_function():
    rng = np.random.default_rng(3)

    def f(x):
        return 0.6 * np.sin(2.0 * x[:, 0]) + x[:, 1] ** 2 - 0.4 * x[:, 2] * x[:, 3] + 0.3 * x.mean(axis=1)

    x = rng.uniform(-1, 1, size=(8192, 53)).astype(np.float32)
    t = f(x).astype(np.float32)
    xh = rng.uniform(-1, 1, size=(1024, 53)).astype(np.float32)
    with pt.Mlp(seed=7) as net:
        before = np.mean((net.forward(xh) - f(xh)) ** 2)
        net.train(x, t, iters=600, batch=256, lr=2e-3)
        after = np.mean((net.forward(xh) - f(xh)) ** 2)
    print('held-out mse before %.4f after %.4f (target variance %.4f)' % (before, after, np.var(f(xh))))
    assert after < 0.1 * np.var(f(xh)) and after < 0.1 * before
