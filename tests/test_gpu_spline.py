"""Spline background (monotone cubic Hermite through refined anchors,
js/rietveld_core.js:549-612 and 932-936) on the HIP path against the FP64 oracle:
chi^2 of the plain, batched, PT and shared-profile forms at <= 1e-5 relative, and the
pattern."""
import numpy as np
import pytest

import oracle_lib as O
from powder6_amd import problems as PR
from powder6_amd import pt

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope='module')
def spline():
    return PR.nacl_spline(simulate=O.oracle_simulate)


def _probes(pb, n, seed, spread=2.0):
    rng = np.random.default_rng(seed)
    x0 = pb.start_vector()
    X = x0 + rng.uniform(-spread, spread, size=(n, pb.nv)) * np.array(pb.scale)
    X = np.minimum(np.array(pb.hi), np.maximum(np.array(pb.lo), X))
    X[0] = x0
    return X


def _rel(g, c):
    return np.abs(np.asarray(g) - np.asarray(c)) / np.maximum(1.0, np.abs(np.asarray(c)))


def test_spline_batched_and_plain_chi2_match_oracle(spline):
    m = O.Model(spline)
    X = _probes(spline, 24, 21)
    # anchors pushed far apart so that some intervals are monotone and some are not,
    # and the tangent limiter acts
    X[5, [spline.var_index('bga%d' % j) for j in range(8)]] = [400, 50, 50, 300, 20, 25, 200, 10]
    with pt.Plan(spline) as pl:
        b = pl.eval(X, pt.FORM_BATCHED)
        p = pl.eval(X[:6], pt.FORM_PLAIN)
    cpu = np.array([m.cost(x) for x in X])
    print('spline batched max rel err', _rel(b, cpu).max(), 'plain', _rel(p, cpu[:6]).max())
    assert _rel(b, cpu).max() <= TOL
    assert _rel(p, cpu[:6]).max() <= TOL


def test_spline_pattern_matches_oracle(spline):
    m = O.Model(spline)
    x = _probes(spline, 3, 4)[2]
    with pt.Plan(spline) as pl:
        y = pl.pattern(x)
    yo = m.pattern(x)
    err = np.max(np.abs(y - yo) / np.maximum(1.0, np.abs(yo)))
    print('spline pattern max rel err', err)
    assert err < 2e-6


def test_spline_pt_batch_states_rescore(spline):
    """PT over the spline problem: current and best costs of the chains re-scored by the
    oracle at the vectors the chains hold."""
    m = O.Model(spline)
    with pt.Plan(spline) as pl:
        pl.initial([spline.start_vector()])
        pl.batch(1, bootstrap=True)
        st = pl.batch(3)
        nv, stride = pl.nv, pl.stride
    S = st.reshape(-1, stride)
    for c in range(0, S.shape[0], 3):
        assert _rel(S[c, 3 * nv + pt.S_COST], m.cost(S[c, :nv])) <= TOL
        assert _rel(S[c, 3 * nv + pt.S_BEST], m.cost(S[c, 2 * nv:3 * nv])) <= TOL


def test_spline_matrix_form_matches_oracle():
    pb = PR.nacl_spline(simulate=O.oracle_simulate, atoms_only=True)
    m = O.Model(pb)
    X = _probes(pb, 16, 8)
    with pt.Plan(pb) as pl:
        g = pl.eval(X, pt.FORM_MATRIX)
    cpu = np.array([m.cost(x) for x in X])
    print('spline matrix form max rel err', _rel(g, cpu).max())
    assert _rel(g, cpu).max() <= TOL
