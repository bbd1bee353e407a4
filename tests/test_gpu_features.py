"""GPU chi^2 against the FP64 oracle over the model's feature matrix: groups without an
inversion centre (the full sine and cosine structure factor), anomalous scattering,
neutron lengths, SL/HL asymmetry, several phases in one pattern, and the shared-profile
matrix form on a non-centrosymmetric phase."""
import numpy as np
import pytest

import oracle_lib as O
from powder6_amd import problems as PR
from powder6_amd import pt

pytestmark = pytest.mark.gpu
TOL = 1e-5

CASES = {
    'F-43m': lambda: PR.zincblende_f43m(simulate=O.oracle_simulate),
    'F-43m_anomalous': lambda: PR.zincblende_f43m(simulate=O.oracle_simulate, anomalous=True),
    'F-43m_neutron': lambda: PR.zincblende_f43m(simulate=O.oracle_simulate, neutron=True),
    'P2_1': lambda: PR.general_noncentro(4, simulate=O.oracle_simulate),
    'P2_12_12_1': lambda: PR.general_noncentro(19, simulate=O.oracle_simulate),
    'two_phases': lambda: PR.two_phase(simulate=O.oracle_simulate),
    'P-1_neutron': lambda: PR.triclinic_probe('tch_aniso', simulate=O.oracle_simulate, neutron=True),
    'P-1_SL_HL': lambda: PR.triclinic_probe('tch_aniso', simulate=O.oracle_simulate, sl=0.02, hl=0.015),
}


def _probes(pb, n, seed):
    rng = np.random.default_rng(seed)
    x0 = np.array(pb.start_vector())
    X = x0 + rng.uniform(-1.5, 1.5, size=(n, pb.nv)) * np.array(pb.scale)
    X = np.minimum(np.array(pb.hi), np.maximum(np.array(pb.lo), X))
    X[0] = x0
    return X


@pytest.mark.parametrize('name', sorted(CASES))
def test_feature_chi2_and_pattern_match_oracle(name):
    pb = CASES[name]()
    m = O.Model(pb)
    X = _probes(pb, 12, 3)
    with pt.Plan(pb) as pl:
        g = pl.eval(X, pt.FORM_BATCHED)
        y = pl.pattern(X[0])
    cpu = np.array([m.cost(x) for x in X])
    err = np.abs(g - cpu) / np.maximum(1.0, np.abs(cpu))
    yo = m.pattern(X[0])
    yerr = np.max(np.abs(y - yo)) / max(1.0, np.max(np.abs(yo)))
    print('%s: chi2 max rel err %.2e, pattern max rel err %.2e' % (name, err.max(), yerr))
    assert err.max() <= TOL, list(zip(g, cpu))
    assert yerr <= 1e-5


def test_matrix_form_on_a_noncentrosymmetric_phase():
    """F-43m with anomalous scattering: only B and scale move, so the shared-profile form
    applies; it must agree with the oracle and the VALU form."""
    pb = PR.zincblende_f43m(simulate=O.oracle_simulate, anomalous=True)
    m = O.Model(pb)
    x0 = np.array(pb.start_vector())
    rng = np.random.default_rng(8)
    X = np.repeat(x0[None, :], 40, axis=0)
    for j, n in enumerate(pb.vars):
        if n.endswith('.B') or n.endswith('.scale'):
            X[:, j] = np.clip(x0[j] + rng.uniform(-1, 1, size=40) * pb.scale[j] * 2, pb.lo[j], pb.hi[j])
    with pt.Plan(pb) as pl:
        gm = pl.eval(X, pt.FORM_MATRIX)
        gv = pl.eval(X, pt.FORM_BATCHED)
    cpu = np.array([m.cost(x) for x in X])
    em = np.max(np.abs(gm - cpu) / np.maximum(1.0, cpu))
    ev = np.max(np.abs(gm - gv) / np.maximum(1.0, gv))
    print('F-43m matrix form: max rel err vs FP64 %.2e, vs VALU form %.2e' % (em, ev))
    assert em <= TOL and ev <= 2e-6
