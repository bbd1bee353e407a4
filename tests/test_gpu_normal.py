"""The GPU normal equations of one LM step (pt_normal_equations: FP64 derivative columns of
js/rietveld_core.js:1245-1298 and normal() :1953-1983) against the oracle's, which
tests/test_normal_oracle.py pins by central differences of the whole pattern.  Both sides
run in FP64, so the bar is 1e-6 of each column's scale: the finite-difference columns divide
ulp-level differences of the transcendentals by steps as small as 1e-6 of a cell edge, or
1e-5 one-sided for a width term at its bound (the P-1 probe's Y = 0 column agrees to 1.2e-7)."""
import numpy as np
import pytest

import oracle_lib as O
from powder6_amd import problems as PR
from powder6_amd import pt

pytestmark = pytest.mark.gpu
TOL = 1e-6

CASES = {
    'pnma_tch': lambda: PR.pbso4_pnma(simulate=O.oracle_simulate, chains=16),
    'p1_tch_po_shifts': lambda: PR.triclinic_probe('tch_aniso', simulate=O.oracle_simulate),
    'p1_tch_asymmetric': lambda: PR.triclinic_probe('tch_aniso', simulate=O.oracle_simulate, sl=.02, hl=.01),
    'p1_simple': lambda: PR.triclinic_probe('simple_pvoigt', simulate=O.oracle_simulate),
    'p1_split_neutron': lambda: PR.triclinic_probe('split_pvoigt', simulate=O.oracle_simulate, neutron=True),
    'nacl_spline': lambda: PR.nacl_spline(simulate=O.oracle_simulate),
    'two_phase': lambda: PR.two_phase(simulate=O.oracle_simulate),
    'p21c_shared_b': lambda: PR.molecular_p21c(simulate=O.oracle_simulate, chains=8, shared_b=True),
    'zns_anomalous': lambda: PR.zincblende_f43m(simulate=O.oracle_simulate, anomalous=True),
}


def _compare(A, g, f, J, A0, g0, f0, J0, label):
    for j in range(J0.shape[0]):
        scale = np.abs(J0[j]).max()
        if scale == 0:
            assert np.abs(J[j]).max() == 0, (label, j)
            continue
        err = np.abs(J[j] - J0[j]).max() / scale
        assert err <= TOL, (label, j, err)
    d = np.sqrt(np.maximum(np.diag(A0), 1e-300))
    assert np.all(np.abs(A - A0) <= TOL * np.outer(d, d) + 1e-300), label
    assert np.all(np.abs(g - g0) <= TOL * d * np.sqrt(f0)), label
    assert abs(f - f0) <= 1e-12 * f0, label


@pytest.mark.parametrize('name', sorted(CASES))
def test_normal_equations_match_oracle(name):
    pb = CASES[name]()
    m = O.Model(pb)
    rng = np.random.default_rng(7)
    x0 = pb.start_vector()
    xs = [x0, np.clip(x0 + rng.uniform(-.3, .3, len(x0)) * np.array(pb.scale), pb.lo, pb.hi)]
    with pt.Plan(pb) as pl:
        for x in xs:
            A0, g0, f0, J0 = m.normal(x)
            A, g, f, J = pl.normal(x, columns=True)
            _compare(A, g, f, J, A0, g0, f0, J0, name)
    m.close()


def test_two_datasets_sum_their_normal_equations():
    """An X-ray and a neutron dataset (js/rietveld_core.js:1955-1982 loops the datasets):
    A, g and chi^2 are the sums of the per-dataset oracle results, J the stacked columns."""
    pbs = PR.xray_neutron_pair(simulate=O.oracle_simulate)
    ms = [O.Model(pb) for pb in pbs]
    x = pbs[0].start_vector()
    parts = [m.normal(x) for m in ms]
    A0 = sum(p[0] for p in parts)
    g0 = sum(p[1] for p in parts)
    f0 = sum(p[2] for p in parts)
    J0 = np.concatenate([p[3] for p in parts], axis=1)
    with pt.Plan(pbs) as pl:
        A, g, f, J = pl.normal(x, columns=True)
    _compare(A, g, f, J, A0, g0, f0, J0, 'xray_neutron')


def test_gauss_newton_step_lowers_chi2():
    """The use of the normal equations: from a perturbed start, the damped step
    (A + lambda diag A) dx = g lowers the FP64 chi^2."""
    pb = PR.pbso4_pnma(simulate=O.oracle_simulate, chains=16)
    m = O.Model(pb)
    x = pb.start_vector()
    with pt.Plan(pb) as pl:
        A, g, f = pl.normal(x)
    d = np.diag(A)
    dx = np.linalg.solve(A + 1e-3 * np.diag(d), g)
    x1 = np.clip(x + dx, pb.lo, pb.hi)
    assert m.cost(x1) < f
    m.close()


def test_separated_profiles_are_refused():
    pb = PR.triclinic_probe('gsas_cw', simulate=O.oracle_simulate)
    with pt.Plan(pb) as pl:
        with pytest.raises(pt.PTError, match='normal equations'):
            pl.normal(pb.start_vector())
