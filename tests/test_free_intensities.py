"""Free-intensity (Le Bail / WPPF) phases in the oracle: the non-negative least-squares
solve (Lawson-Hanson on the weighted normal equations) against brute force over every
passive set, its optimality conditions, and the pattern-level facts the method fixes:
noiseless data of the phase's own Rietveld model are fitted exactly, and a free phase
never fits worse than any intensity set (js/pt_gpu.js:259-297,
js/rietveld_core.js:1019-1090)."""
import itertools

import numpy as np

import oracle_lib as O
from powder6_amd import problems as PR


def _nnls(H, b):
    m = len(b)
    x = np.zeros(m)
    H = np.ascontiguousarray(H, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    rc = O.lib().or_nnls_normal(m, O.d(H)[1], O.d(b)[1], x.ctypes.data_as(O.P_D))
    assert rc >= 0
    return x


def _brute(A, w, y):
    m = A.shape[1]
    best, bx = np.inf, None
    for k in range(m + 1):
        for S in itertools.combinations(range(m), k):
            x = np.zeros(m)
            if S:
                As = A[:, S] * np.sqrt(w)[:, None]
                xs, *_ = np.linalg.lstsq(As, y * np.sqrt(w), rcond=None)
                if np.any(xs < 0):
                    continue
                x[list(S)] = xs
            f = np.sum(w * (y - A @ x) ** 2)
            if f < best:
                best, bx = f, x
    return best, bx


def test_nnls_against_brute_force():
    rng = np.random.default_rng(3)
    for trial in range(25):
        n, m = 40, int(rng.integers(2, 7))
        A = np.abs(rng.normal(size=(n, m))) * (rng.random((n, m)) < 0.6)
        A[:, 0] += 0.1
        y = A @ rng.normal(size=m) + 0.3 * rng.normal(size=n)
        w = rng.uniform(0.2, 2.0, size=n)
        H = A.T @ (w[:, None] * A)
        b = A.T @ (w * y)
        x = _nnls(H, b)
        f = np.sum(w * (y - A @ x) ** 2)
        fb, xb = _brute(A, w, y)
        assert np.all(x >= 0)
        assert abs(f - fb) <= 1e-9 * max(1.0, fb), (trial, f, fb)
        assert np.allclose(x, xb, atol=1e-7 * max(1.0, np.max(np.abs(xb))))


def test_nnls_optimality_conditions():
    """At the solution: x >= 0, the gradient b - Hx is zero where x > 0 and <= 0 where
    x = 0 (Karush-Kuhn-Tucker)."""
    rng = np.random.default_rng(11)
    n, m = 300, 40
    t = np.linspace(0, 1, n)
    centres = np.sort(rng.uniform(0, 1, m))
    A = np.exp(-((t[:, None] - centres[None, :]) / 0.02) ** 2)
    y = A @ np.abs(rng.normal(size=m)) - 0.4 * A[:, ::3].sum(axis=1) + 0.01 * rng.normal(size=n)
    w = np.ones(n)
    H, b = A.T @ A, A.T @ y
    x = _nnls(H, b)
    g = b - H @ x
    scale = np.max(np.abs(b))
    assert np.all(x >= 0)
    assert np.all(np.abs(g[x > 0]) <= 1e-8 * scale)
    assert np.all(g[x == 0] <= 1e-8 * scale)
    assert np.sum(x > 0) < m          # some intensities are held at zero


def test_single_column_closed_form():
    for bv in (3.0, -2.0):
        x = _nnls(np.array([[4.0]]), np.array([bv]))
        assert x[0] == max(0.0, bv / 4.0)


def test_lebail_fits_its_own_rietveld_pattern_exactly():
    """Noiseless data from the phase's Rietveld model: the free intensities reproduce it
    (chi^2 at the true cell and widths ~ 0), and at any vector the free phase fits at
    least as well as the Rietveld intensities."""
    pb = PR.lebail_probe()
    x = pb.start_vector()
    pb.phases[0].method = 'rietveld'
    y_true = O.Model(pb).pattern(x)
    pb.y = y_true.copy()
    pb.w = 1.0 / np.maximum(y_true, 1.0)
    rv = O.Model(pb)
    pb.phases[0].method = 'lebail'
    lb = O.Model(pb)
    c0 = lb.cost(x)
    assert c0 <= 1e-12 * np.sum(pb.w * pb.y ** 2)
    rng = np.random.default_rng(5)
    for _ in range(4):
        xx = x + rng.uniform(-1, 1, size=pb.nv) * np.array(pb.scale)
        xx = np.minimum(np.array(pb.hi), np.maximum(np.array(pb.lo), xx))
        assert lb.cost(xx) <= rv.cost(xx) * (1 + 1e-12)
