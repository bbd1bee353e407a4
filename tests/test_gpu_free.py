"""Free-intensity (Le Bail / WPPF) phases on the HIP path: every trial's chi^2 is the least
one over non-negative reflection intensities (js/pt_gpu.js:259-297), solved per trial on
the GPU in FP64 (k_stage_nnls), against the FP64 oracle's Lawson-Hanson solve at
<= 1e-5 relative; a Le Bail phase alone and next to a Rietveld phase; PT states."""
import numpy as np
import pytest

import oracle_lib as O
from powder6_amd import problems as PR
from powder6_amd import pt

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _probes(pb, n, seed, spread=1.5):
    rng = np.random.default_rng(seed)
    x0 = pb.start_vector()
    X = x0 + rng.uniform(-spread, spread, size=(n, pb.nv)) * np.array(pb.scale)
    X = np.minimum(np.array(pb.hi), np.maximum(np.array(pb.lo), X))
    X[0] = x0
    return X


def _rel(g, c):
    return np.abs(np.asarray(g) - np.asarray(c)) / np.maximum(1.0, np.abs(np.asarray(c)))


@pytest.mark.parametrize('mixed', [False, True])
def test_free_intensity_chi2_matches_oracle(mixed):
    pb = PR.lebail_probe(mixed=mixed, simulate=O.oracle_simulate)
    m = O.Model(pb)
    X = _probes(pb, 16, 4)
    with pt.Plan(pb) as pl:
        g = pl.eval(X, pt.FORM_BATCHED)
    cpu = np.array([m.cost(x) for x in X])
    print('free intensities (mixed=%s) max rel err' % mixed, _rel(g, cpu).max())
    assert _rel(g, cpu).max() <= TOL


def test_free_intensity_pt_states_rescore():
    pb = PR.lebail_probe(simulate=O.oracle_simulate)
    m = O.Model(pb)
    with pt.Plan(pb) as pl:
        pl.initial([pb.start_vector()])
        s0 = pl.batch(1, bootstrap=True)
        st = pl.batch(2)
        nv, stride = pl.nv, pl.stride
        with pytest.raises(pt.PTError, match='free intensities'):
            pl.eval(pb.start_vector()[None, :], pt.FORM_PLAIN)
    assert _rel(s0[3 * nv + pt.S_COST], m.cost(pb.start_vector())) <= TOL
    S = st.reshape(-1, stride)
    assert S[:, 3 * nv + pt.S_PROPOSALS].sum() == pl.chains * 2 * nv
    for c in range(S.shape[0]):
        assert _rel(S[c, 3 * nv + pt.S_COST], m.cost(S[c, :nv])) <= TOL
        assert _rel(S[c, 3 * nv + pt.S_BEST], m.cost(S[c, 2 * nv:3 * nv])) <= TOL


@pytest.mark.parametrize('method', ['lebail', 'wppf'])
def test_more_live_intensities_than_lds_holds(method):
    """The P-1 probe over 8-120 deg: about 200 reflections have a window in every trial, more
    than the 120 columns the LDS solve holds, so each trial solves in its slot's global
    workspace. chi^2 of probes and of PT states against the oracle at 1e-5 (a trial over the
    LDS limit without that workspace would cost 3e38)."""
    pb = PR.triclinic_probe('tch_aniso', simulate=O.oracle_simulate, method=method, n_points=480, span=112.0)
    m = O.Model(pb)
    with pt.Plan(pb) as pl:
        X = _probes(pb, pl.chains, 11, spread=0.5)   # pt_costs scores one vector per chain
        g = pl.costs(X)
        pl.initial([pb.start_vector()])
        pl.batch(1, bootstrap=True)
        st = pl.batch(1)
        nv, stride = pl.nv, pl.stride
    c = np.array([m.cost(x) for x in X])
    assert np.all(g < 1e30)
    assert _rel(g, c).max() <= TOL
    S = st.reshape(-1, stride)
    for k in range(S.shape[0]):
        assert _rel(S[k, 3 * nv + pt.S_COST], m.cost(S[k, :nv])) <= TOL
