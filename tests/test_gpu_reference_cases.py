"""The cases of the reference's own WebGPU PT test page (tests/webgpu-pt.js:47-124) on the
HIP path, each against the FP64 oracle at 1e-5 instead of the page's 1e-3: every method
and profile family, the separated broadening models, a distant second wavelength, the
negative-variance continuation, the variance guard under PT and a 12,000-point pattern.
Free-intensity (Le Bail / WPPF) phases run on the page's P-1 probe (544 envelope
reflections, about 60 with a window in a trial) and on the NaCl pattern."""
import numpy as np
import pytest

import oracle_lib as O
from powder6_amd import problems as PR
from powder6_amd import pt

pytestmark = pytest.mark.gpu
TOL = 1e-5
FAMILIES = ['simple_pvoigt', 'split_pvoigt', 'tch_aniso', 'gsas_cw']


def _page_probes(pb):
    """x and x moved by -0.12, 0.12 and 0.3 PT scales, clipped (webgpu-pt.js:49)."""
    x0 = pb.start_vector()
    return np.array([x0] + [np.clip(x0 + s * np.array(pb.scale), pb.lo, pb.hi) for s in (-.12, .12, .3)])


def _check(pb):
    m = O.Model(pb)
    X = _page_probes(pb)
    with pt.Plan(pb) as pl:
        g = pl.costs(X)[:len(X)]
    cpu = np.array([m.cost(x) for x in X])
    err = np.abs(g - cpu) / np.maximum(1.0, cpu)
    return err.max()


@pytest.mark.parametrize('family', FAMILIES)
def test_rietveld_each_family(family):
    assert _check(PR.triclinic_probe(family, simulate=O.oracle_simulate)) <= TOL


@pytest.mark.parametrize('method', ['wppf', 'lebail'])
@pytest.mark.parametrize('family', FAMILIES)
def test_free_intensities_each_family(method, family):
    assert _check(PR.triclinic_probe(family, simulate=O.oracle_simulate, method=method)) <= TOL
    assert _check(PR.lebail_probe(simulate=O.oracle_simulate, profile=family, method=method)) <= TOL


@pytest.mark.parametrize('broadening', [
    {'size': 'uniaxial', 'strain': 'uniaxial', 'sizeAxis': [1, 0, 1], 'strainAxis': [1, 1, 0]},
    {'size': 'ellipsoid', 'strain': 'stephens'}])
def test_separated_broadening(broadening):
    """webgpu-pt.js:57-60 (the page's simple pseudo-Voigt instrument)."""
    pb = PR.triclinic_probe('simple_pvoigt', simulate=O.oracle_simulate, broadening=broadening)
    assert _check(pb) <= TOL


def test_second_line_size_width_follows_its_wavelength():
    """webgpu-pt.js:63-68: lambda2 = 1.9 and size 0.02 make the Ka2 line's size term large."""
    pb = PR.triclinic_probe('gsas_cw', lambda2=1.9)
    pb.set('p1.size', .02)
    pb.x0[pb.var_index('p1.size')] = .02
    pb.y = np.asarray(O.oracle_simulate(pb, pb.start_vector())) + 1.0
    assert _check(pb) <= TOL


@pytest.mark.parametrize('U', [3.96361, -1e5])
def test_negative_variance_continuation(U):
    """webgpu-pt.js:71-77: a negative Gaussian variance takes the same soft continuation."""
    inst = {'U': U, 'V': -43.01288, 'W': 18.71003, 'X': 1.40615, 'Y': 4.800405, 'Z': .01277393}
    pb = PR.triclinic_probe('gsas_cw', inst=inst)
    for n in ('W', 'X'):
        pb.x0[pb.var_index('inst.' + n)] = inst[n]
    pb.y = np.asarray(O.oracle_simulate(pb, pb.start_vector())) + 1.0
    assert _check(pb) <= TOL


def test_variance_guard_keeps_replicas_physical():
    """webgpu-pt.js:106-112: inst.W free in [-100, 100] with U = V = 0; after PT every
    replica and every best vector keeps W >= 0."""
    pb = PR.triclinic_probe('gsas_cw', inst={'U': 0.0, 'V': 0.0, 'W': 1.0}, simulate=O.oracle_simulate)
    pb.base = pb.values()
    PR.link_datasets([pb], ['inst.W'], [[]], limits={'inst.W': (-100.0, 100.0)})
    pb.scale[0] = 10.0
    pb.chains = 8
    m = O.Model(pb)
    with pt.Plan(pb) as pl:
        pl.initial([pb.start_vector()])
        pl.batch(1, bootstrap=True)
        st = pl.batch(10)
        nv, stride = pl.nv, pl.stride
    S = st.reshape(-1, stride)
    assert S[:, 3 * nv + pt.S_PROPOSALS].sum() > 0
    for c in range(S.shape[0]):
        assert S[c, 0] >= 0 and S[c, 2 * nv] >= 0
        assert m.violation(S[c, 2 * nv:3 * nv]) == 0.0


def test_twelve_thousand_points():
    """webgpu-pt.js:120-124: 12,000 points over 8-72 deg."""
    pb = PR.triclinic_probe('gsas_cw', n_points=12000, span=64.0, simulate=O.oracle_simulate)
    pb.w = np.full(len(pb.tth), .01)
    assert _check(pb) <= TOL


def _multi_probes(pbs):
    """x and x moved by 0.1 PT scales, clipped (webgpu-pt.js:97, 105), and three more
    seeded moves across the boxes."""
    pb = pbs[0]
    x0 = pb.start_vector()
    sc, lo, hi = np.array(pb.scale), np.array(pb.lo), np.array(pb.hi)
    rng = np.random.default_rng(83)
    X = [x0, np.clip(x0 + .1 * sc, lo, hi)] + [np.clip(x0 + rng.uniform(-.5, .5, len(x0)) * sc, lo, hi) for _ in range(3)]
    return np.array(X)


def _check_multi(pbs):
    ms = [O.Model(pb) for pb in pbs]
    X = _multi_probes(pbs)
    with pt.Plan(pbs) as pl:
        g = pl.costs(X)
        # a few PT sweeps: every chain's state re-scored by the oracle
        pl.initial([X[0]])
        pl.batch(1, bootstrap=True)
        S = pl.batch(3).reshape(pl.chains, pl.stride)
        nv = pl.nv
    worst = 0.0
    for c, gc in enumerate(g):
        cpu = sum(m.cost(X[c % len(X)]) for m in ms)
        worst = max(worst, abs(gc - cpu) / max(1.0, cpu))
    for c in range(S.shape[0]):
        cpu = sum(m.cost(S[c, :nv]) for m in ms)
        worst = max(worst, abs(S[c, 3 * nv] - cpu) / max(1.0, cpu))
    assert S[:, 3 * nv + pt.S_PROPOSALS].sum() == S.shape[0] * 3 * nv
    return worst


def test_joint_constraints_mixed_phases_absent_phase():
    """webgpu-pt.js:83-99: held B, equivalence C2.x = 0.5 C1.x, C1.occ + C2.occ = 1, a WPPF
    phase beside the Rietveld phase in the X-ray dataset and absent from the neutron
    dataset, per-dataset bkg0 and zero."""
    assert _check_multi(PR.page_joint_case(simulate=O.oracle_simulate)) <= TOL


def test_scoped_resolution_functions_and_cross_dataset_link():
    """webgpu-pt.js:100-106: resolutionType simple_pvoigt in one dataset and split_pvoigt
    in the other, and C1.B of the second dataset = 1.2 C1.B of the first + 0.1."""
    assert _check_multi(PR.page_scoped_case(simulate=O.oracle_simulate)) <= TOL
