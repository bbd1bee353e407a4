"""Multi-dataset plans (pt_plan_create_multi): an X-ray and a neutron pattern of one phase
sharing cell, coordinates and B, each with its own zero, scale, widths and background.
The cost of a trial is the sum of the datasets' chi^2 (js/pt_gpu.js:259-297); checked
against the sum of the FP64 oracle's costs at <= 1e-5, through batched evaluation and the
PT states, and each dataset's pattern against the oracle."""
import numpy as np
import pytest

import oracle_lib as O
from powder6_amd import problems as PR
from powder6_amd import pt

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope='module')
def pair():
    return PR.xray_neutron_pair(simulate=O.oracle_simulate)


def _probes(pb, n, seed, spread=1.5):
    rng = np.random.default_rng(seed)
    x0 = pb.start_vector()
    X = x0 + rng.uniform(-spread, spread, size=(n, pb.nv)) * np.array(pb.scale)
    X = np.minimum(np.array(pb.hi), np.maximum(np.array(pb.lo), X))
    X[0] = x0
    return X


def _rel(g, c):
    return np.abs(np.asarray(g) - np.asarray(c)) / np.maximum(1.0, np.abs(np.asarray(c)))


def test_multi_batched_cost_is_the_sum_of_the_datasets(pair):
    ms = [O.Model(pb) for pb in pair]
    X = _probes(pair[0], 24, 2)
    with pt.Plan(pair) as pl:
        g = pl.eval(X, pt.FORM_BATCHED)
        y0 = pl.pattern(X[3])
        y1 = pl.dataset_pattern(1, X[3])
    cpu = np.array([sum(m.cost(x) for m in ms) for x in X])
    print('multi-dataset batched max rel err', _rel(g, cpu).max())
    assert _rel(g, cpu).max() <= TOL
    for y, m in ((y0, ms[0]), (y1, ms[1])):
        yo = m.pattern(X[3])
        assert np.max(np.abs(y - yo) / np.maximum(1.0, np.abs(yo))) < 2e-6


def test_multi_pt_states_rescore(pair):
    ms = [O.Model(pb) for pb in pair]
    with pt.Plan(pair) as pl:
        pl.initial([pair[0].start_vector()])
        s0 = pl.batch(1, bootstrap=True)
        st = pl.batch(3)
        nv, stride = pl.nv, pl.stride
    init = s0[3 * nv + pt.S_COST]
    assert _rel(init, sum(m.cost(pair[0].start_vector()) for m in ms)) <= TOL
    S = st.reshape(-1, stride)
    assert S[:, 3 * nv + pt.S_PROPOSALS].sum() == pl.chains * 3 * nv
    assert S[:, 3 * nv + pt.S_SWAPS].sum() > 0
    assert S[:, 3 * nv + pt.S_BEST].min() <= init
    for c in range(0, S.shape[0], 3):
        assert _rel(S[c, 3 * nv + pt.S_COST], sum(m.cost(S[c, :nv]) for m in ms)) <= TOL
        assert _rel(S[c, 3 * nv + pt.S_BEST], sum(m.cost(S[c, 2 * nv:3 * nv]) for m in ms)) <= TOL


def test_one_dataset_multi_plan_is_the_plain_plan(pair):
    """A multi-dataset plan of one dataset runs exactly the single-dataset pipeline."""
    pb = pair[0]
    with pt.Plan([pb]) as a, pt.Plan(pb) as b:
        for pl in (a, b):
            pl.initial([pb.start_vector()])
            pl.batch(1, bootstrap=True)
        sa, sb = a.batch(2), b.batch(2)
    assert np.array_equal(sa, sb)
