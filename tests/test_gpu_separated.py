"""Separated (GSAS) profiles on the HIP path against the FP64 oracle: the dataset's
instrument function (TCH, simple or split pseudo-Voigt) plus isotropic, uniaxial or
ellipsoidal size and isotropic, uniaxial or Stephens strain broadening
(js/profile.js:384-422, js/rietveld_core.js:694-708).  chi^2 of the batched and plain
forms at <= 1e-5 relative, patterns, the PT states and the shared-profile form."""
import numpy as np
import pytest

import oracle_lib as O
from powder6_amd import problems as PR
from powder6_amd import pt

pytestmark = pytest.mark.gpu
TOL = 1e-5
CASES = [('isotropic', 'isotropic', 'gsas_cw'), ('uniaxial', 'uniaxial', 'gsas_cw'), ('ellipsoid', 'stephens', 'gsas_cw'),
         ('isotropic', 'uniaxial', 'simple_pvoigt'), ('uniaxial', 'stephens', 'split_pvoigt')]


def _probes(pb, n, seed, spread=1.5):
    rng = np.random.default_rng(seed)
    x0 = pb.start_vector()
    X = x0 + rng.uniform(-spread, spread, size=(n, pb.nv)) * np.array(pb.scale)
    X = np.minimum(np.array(pb.hi), np.maximum(np.array(pb.lo), X))
    X[0] = x0
    return X


def _rel(g, c):
    return np.abs(np.asarray(g) - np.asarray(c)) / np.maximum(1.0, np.abs(np.asarray(c)))


@pytest.mark.parametrize('size,strain,family', CASES)
def test_separated_chi2_and_pattern_match_oracle(size, strain, family):
    pb = PR.separated_probe(size, strain, family, simulate=O.oracle_simulate)
    m = O.Model(pb)
    X = _probes(pb, 20, 3)
    with pt.Plan(pb) as pl:
        b = pl.eval(X, pt.FORM_BATCHED)
        p = pl.eval(X[:4], pt.FORM_PLAIN)
        y = pl.pattern(X[1])
    cpu = np.array([m.cost(x) for x in X])
    yo = m.pattern(X[1])
    perr = np.max(np.abs(y - yo) / np.maximum(1.0, np.abs(yo)))
    print(size, strain, family, 'batched max rel err', _rel(b, cpu).max(), 'plain', _rel(p, cpu[:4]).max(), 'pattern', perr)
    assert _rel(b, cpu).max() <= TOL
    assert _rel(p, cpu[:4]).max() <= TOL
    assert perr < 2e-6


def test_separated_pt_states_rescore():
    pb = PR.separated_probe('ellipsoid', 'stephens', 'gsas_cw', simulate=O.oracle_simulate)
    m = O.Model(pb)
    with pt.Plan(pb) as pl:
        pl.initial([pb.start_vector()])
        pl.batch(1, bootstrap=True)
        st = pl.batch(4)
        nv, stride = pl.nv, pl.stride
    S = st.reshape(-1, stride)
    assert S[:, 3 * nv + pt.S_PROPOSALS].sum() > 0
    for c in range(S.shape[0]):
        assert _rel(S[c, 3 * nv + pt.S_COST], m.cost(S[c, :nv])) <= TOL
        assert _rel(S[c, 3 * nv + pt.S_BEST], m.cost(S[c, 2 * nv:3 * nv])) <= TOL
        assert m.violation(S[c, :nv]) == 0.0


def test_separated_stephens_guard_rejects_like_the_oracle():
    """A trial whose strain quartic goes negative is infeasible on both sides: its cost
    is evaluated, but PT never accepts it (the guard)."""
    pb = PR.separated_probe('isotropic', 'stephens', 'gsas_cw', simulate=O.oracle_simulate)
    m = O.Model(pb)
    X = _probes(pb, 6, 9)
    j = pb.var_index('p1.S220')
    X[3, j] = -2000.0
    assert m.violation(X[3]) > 0 and m.violation(X[0]) == 0
    with pt.Plan(pb) as pl:
        b = pl.eval(X, pt.FORM_BATCHED)
    cpu = np.array([m.cost(x) for x in X])
    assert _rel(b, cpu).max() <= TOL


def test_separated_matrix_form_matches_oracle():
    """Structure-only moves over a separated profile: the shared-profile MFMA form."""
    pb = PR.separated_probe('uniaxial', 'stephens', 'gsas_cw', simulate=O.oracle_simulate)
    names = [n for n in pb.vars]
    keep = [i for i, n in enumerate(names) if PR.classify(n)[0] == 'atom' or n == 'p1.scale']
    # freeze every other variable at its start value
    x0 = pb.start_vector()
    base = pb.base.copy()
    for (k, j, c) in pb.terms:
        if j not in keep:
            base[k] += c * x0[j]
    pb.base = base
    remap = {j: i for i, j in enumerate(keep)}
    pb.terms = [(k, remap[j], c) for (k, j, c) in pb.terms if j in keep]
    pb.vars = [names[j] for j in keep]
    for attr in ('x0', 'lo', 'hi', 'scale'):
        setattr(pb, attr, [getattr(pb, attr)[j] for j in keep])
    m = O.Model(pb)
    X = _probes(pb, 12, 5)
    with pt.Plan(pb) as pl:
        g = pl.eval(X, pt.FORM_MATRIX)
    cpu = np.array([m.cost(x) for x in X])
    print('separated matrix form max rel err', _rel(g, cpu).max())
    assert _rel(g, cpu).max() <= TOL
