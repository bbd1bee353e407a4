"""The oracle's normal equations of one LM step (oracle/p6_oracle.c or_normal: the columns of
Model.derivatives, js/rietveld_core.js:1245-1298, assembled as normal(), :1953-1983),
pinned independently of the way the oracle builds them:

* every variable's column against a central difference of the oracle's whole pattern at a
  step four times the reference's (analytic columns -- scale, atoms, background -- and
  finite-difference ones alike), so a wrong sign, factor, step or phase list fails;
* A and g against J'WJ and J'W(y - ycalc) formed here from those columns and the pattern,
  and chi^2 against cost();
* linked variables (one B for several atoms) as the sum of their parameters' columns.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from powder6_amd import problems as PR

CASES = {
    'pnma_tch': lambda: PR.pbso4_pnma(simulate=O.oracle_simulate, chains=16),
    'p1_tch_po_shifts': lambda: PR.triclinic_probe('tch_aniso', simulate=O.oracle_simulate),
    'p1_tch_asymmetric': lambda: PR.triclinic_probe('tch_aniso', simulate=O.oracle_simulate, sl=.02, hl=.01),
    'p1_simple': lambda: PR.triclinic_probe('simple_pvoigt', simulate=O.oracle_simulate),
    'p1_split_neutron': lambda: PR.triclinic_probe('split_pvoigt', simulate=O.oracle_simulate, neutron=True),
    'nacl_spline': lambda: PR.nacl_spline(simulate=O.oracle_simulate),
    'two_phase': lambda: PR.two_phase(simulate=O.oracle_simulate),
    'p21c_shared_b': lambda: PR.molecular_p21c(simulate=O.oracle_simulate, chains=8, shared_b=True),
}


def _fd_step(name, v):
    """Four times js/rietveld_core.js:526-537 fdStep (a different step than the oracle's)."""
    kind, comp = PR.classify(name)
    a = abs(v)
    if kind == 'cell':
        h = max(1e-6 * a, 1e-6) if comp in ('a', 'b', 'c') else 1e-4
    else:
        h = {'zero': 1e-4, 'disp': 1e-2, 'transp': 1e-6, 'po': 1e-4}.get(kind, max(1e-4 * a, 1e-5))
        if kind == 'bga':
            h = max(1e-4 * a, 1e-3)
    return 4 * h


def _pattern(m, pb, x):
    y = np.zeros(len(pb.tth))
    assert O.lib().or_pattern(m.h, np.ascontiguousarray(x).ctypes.data_as(O.P_D), y.ctypes.data_as(O.P_D)) == 0
    return y


@pytest.fixture(scope='module', params=sorted(CASES))
def case(request):
    pb = CASES[request.param]()
    m = O.Model(pb)
    x = pb.start_vector()
    A, g, f, J = m.normal(x)
    yield request.param, pb, m, x, A, g, f, J
    m.close()


def test_columns_match_central_differences_of_the_pattern(case):
    name, pb, m, x, A, g, f, J = case
    worst = 0.0
    for j, var in enumerate(pb.vars):
        h = _fd_step(var.split('@')[0], x[j])
        xp, xm = x.copy(), x.copy()
        xp[j] += h
        xm[j] -= h
        fd = (_pattern(m, pb, xp) - _pattern(m, pb, xm)) / (2 * h)
        scale = max(np.abs(J[j]).max(), np.abs(fd).max(), 1e-300)
        err = np.abs(J[j] - fd).max() / scale
        worst = max(worst, err)
        # the reference's own steps leave a truncation error of up to a few 1e-4 (disp: 1e-2)
        assert err <= 1e-3, (name, var, err)
    assert worst <= 1e-3


def test_normal_matrix_is_the_weighted_product_of_the_columns(case):
    name, pb, m, x, A, g, f, J = case
    w = pb.w
    r = np.where(w > 0, pb.y - _pattern(m, pb, x), 0.0)
    An = (J * w) @ J.T
    gn = (J * w) @ r
    assert np.allclose(A, An, rtol=1e-12, atol=1e-12 * np.abs(An).max())
    assert np.allclose(g, gn, rtol=1e-12, atol=1e-12 * np.abs(gn).max())
    assert f == pytest.approx(m.cost(x), rel=1e-14)
    assert np.allclose(A, A.T)


def test_linked_variable_is_the_sum_of_its_parameters():
    """A variable linked to several model parameters (one B for all 17 atoms) has the sum
    of their columns: checked against a central difference of the pattern in it."""
    pb = PR.molecular_p21c(simulate=O.oracle_simulate, chains=8, shared_b=True)
    jb = pb.vars.index('p1.Biso')
    m = O.Model(pb)
    x = pb.start_vector()
    J = m.normal(x)[3]
    # the same column by hand: the central difference of the pattern in the shared variable
    h = 4e-4
    xp, xm = x.copy(), x.copy()
    xp[jb] += h
    xm[jb] -= h
    fd = (_pattern(m, pb, xp) - _pattern(m, pb, xm)) / (2 * h)
    assert np.abs(J[jb] - fd).max() <= 1e-5 * np.abs(fd).max()
    assert sum(1 for (k, j, c) in pb.terms if j == jb) > 1
    m.close()
