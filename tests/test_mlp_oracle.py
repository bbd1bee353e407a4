"""The FP64 surrogate-network oracle (oracle/mlp_oracle.c) pinned by central finite
differences and by closed forms: the GPU training kernels are checked against it."""
import numpy as np

import oracle_lib as O

N_IN, H = 8, 16     # small widths: the oracle takes any; the GPU network is 64/256


def _setup(seed=0, n=6):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.4, O.mlp_param_count(N_IN, H))
    x = rng.normal(size=(n, N_IN))
    t = rng.normal(size=n)
    return rng, p, x, t


def test_param_count_matches_layout():
    assert O.mlp_param_count(64, 256) == 64 * 256 + 256 + 2 * (256 * 256 + 256) + 256 + 1 == 148481


def test_gradient_matches_central_differences():
    rng, p, x, t = _setup()
    mask = np.array([1, 1, 0, 1, 1, 1], dtype=np.float64)
    loss, g = O.mlp_loss_grad(p, N_IN, H, x, t, mask)
    # a sample from every parameter block: W1, b1, W2, b2, W3, b3, w4, b4
    offs = np.cumsum([0, N_IN * H, H, H * H, H, H * H, H, H, 1])
    idx = [int(rng.integers(offs[k], offs[k + 1])) for k in range(8) for _ in range(3)]
    eps = 1e-6
    for i in idx:
        pp = p.copy()
        pp[i] += eps
        lp, _ = O.mlp_loss_grad(pp, N_IN, H, x, t, mask)
        pp[i] -= 2 * eps
        lm, _ = O.mlp_loss_grad(pp, N_IN, H, x, t, mask)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_closed_forms():
    rng, p, x, t = _setup(1)
    n = len(t)
    offs = np.cumsum([0, N_IN * H, H, H * H, H, H * H, H, H, 1])
    p[offs[6]:offs[7]] = 0.0            # w4 = 0: y = b4 for every row
    p[offs[7]] = 0.3
    y = O.mlp_forward(p, N_IN, H, x)
    assert np.allclose(y, 0.3, rtol=0, atol=1e-15)
    loss, g = O.mlp_loss_grad(p, N_IN, H, x, t)
    assert abs(loss - np.mean((0.3 - t) ** 2)) < 1e-14
    assert abs(g[offs[7]] - 2 * np.mean(0.3 - t)) < 1e-14
    assert np.all(g[:offs[6]] == 0.0)   # nothing upstream of a zero w4 moves the loss
    # masked rows contribute nothing
    m = np.ones(n)
    m[0] = 0
    l2, g2 = O.mlp_loss_grad(p, N_IN, H, x, t, m)
    assert abs(l2 - np.sum((0.3 - t[1:]) ** 2) / n) < 1e-14
