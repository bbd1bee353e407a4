"""The Node host (powder6_amd/node/pt_hip.js over the N-API addon) driven with a PT context
shaped like the reference's (tests/js_ctx.py): compile() without a device matches the C
plan; on the GPU, costs() of the probe vectors match the FP64 oracle and run() completes
a PT search with the reference's invariants."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import js_ctx
import oracle_lib as O
from powder6_amd import problems as PR
from powder6_amd import pt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, 'tests', 'js', 'pt_hip_probe.js')
ADDON = os.path.join(ROOT, 'powder6_amd', '_lib', 'pt_addon.node')
NODE = shutil.which('node')
needs_node = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON), reason='node or the addon is not available')


def _node(ctx_path, mode):
    r = subprocess.run([NODE, PROBE, ctx_path, mode], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    line = r.stdout.strip().splitlines()[-1] if r.stdout.strip() else '{}'
    return json.loads(line), r


@needs_node
@pytest.mark.parametrize('profile', ['tch_aniso', 'simple_pvoigt', 'split_pvoigt'])
def test_compile_from_reference_context(profile, tmp_path):
    pb = PR.triclinic_probe(profile)
    pb.y = np.ones_like(pb.tth)
    path = js_ctx.write_ctx(pb, str(tmp_path / 'ctx.json'))
    out, r = _node(path, 'compile')
    assert 'compile' in out, r.stderr
    with pt.Plan(pb, device=False) as pl:
        info = pl.info
    c = out['compile']
    assert c['nv'] == info['nv'] and c['chains'] == info['chains'] and c['stride'] == info['stride']
    assert c['reflectionCount'] == info['reflection_count'] and c['pointCount'] == info['point_count']
    assert c['paramCount'] == info['n_params']


@needs_node
def test_compile_spline_background_context(tmp_path):
    """A dataset with a spline background (js/rietveld_core.js:639-673 bkgX) compiles:
    the anchors become the plan's background terms."""
    pb = PR.nacl_spline()
    pb.y = np.ones_like(pb.tth)
    path = js_ctx.write_ctx(pb, str(tmp_path / 'ctx.json'))
    out, r = _node(path, 'compile')
    assert 'compile' in out, r.stderr
    with pt.Plan(pb, device=False) as pl:
        info = pl.info
    c = out['compile']
    assert c['nv'] == info['nv'] == 16 and c['pointCount'] == info['point_count'] and c['paramCount'] == info['n_params']


@needs_node
@pytest.mark.parametrize('family', ['gsas_cw', 'split_pvoigt'])
def test_compile_separated_profile_context(family, tmp_path):
    """A phase with a separated (GSAS) profile and the dataset's instrument function
    (js/rietveld_core.js:460, 694-708) compiles with every sample and inst.* name mapped."""
    pb = PR.separated_probe('ellipsoid', 'stephens', family)
    pb.y = np.ones_like(pb.tth)
    path = js_ctx.write_ctx(pb, str(tmp_path / 'ctx.json'))
    out, r = _node(path, 'compile')
    assert 'compile' in out, r.stderr
    with pt.Plan(pb, device=False) as pl:
        info = pl.info
    c = out['compile']
    assert c['nv'] == info['nv'] and c['paramCount'] == info['n_params'] and c['reflectionCount'] == info['reflection_count']


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@needs_node
@pytest.mark.skipif(_has_gpu(), reason='checks the behaviour without a device')
def test_create_rejects_without_device(tmp_path):
    pb = PR.triclinic_probe('tch_aniso')
    pb.y = np.ones_like(pb.tth)
    path = js_ctx.write_ctx(pb, str(tmp_path / 'ctx.json'), probes=[list(pb.x0)])
    out, r = _node(path, 'costs')
    assert 'unavailable' in out.get('error', ''), (out, r.stderr)


def _probes(pb):
    x0 = pb.start_vector()
    return [x0] + [np.clip(x0 + s * np.array(pb.scale), pb.lo, pb.hi) for s in (-.12, .12, .3)]


@needs_node
@pytest.mark.gpu
@pytest.mark.parametrize('profile', ['tch_aniso', 'split_pvoigt'])
def test_node_costs_match_oracle(profile, tmp_path):
    pb = PR.triclinic_probe(profile, simulate=O.oracle_simulate)
    m = O.Model(pb)
    probes = _probes(pb)
    path = js_ctx.write_ctx(pb, str(tmp_path / 'ctx.json'), probes=[list(map(float, p)) for p in probes])
    out, r = _node(path, 'costs')
    assert 'costs' in out, (out, r.stderr)
    for c, g in enumerate(out['costs']):
        cpu = m.cost(probes[c % len(probes)])
        assert abs(g - cpu) / max(1.0, cpu) <= 1e-5


@needs_node
@pytest.mark.gpu
def test_node_costs_spline_background_match_oracle(tmp_path):
    pb = PR.nacl_spline(simulate=O.oracle_simulate)
    m = O.Model(pb)
    probes = _probes(pb)
    path = js_ctx.write_ctx(pb, str(tmp_path / 'ctx.json'), probes=[list(map(float, p)) for p in probes])
    out, r = _node(path, 'costs')
    assert 'costs' in out, (out, r.stderr)
    for c, g in enumerate(out['costs']):
        cpu = m.cost(probes[c % len(probes)])
        assert abs(g - cpu) / max(1.0, cpu) <= 1e-5


@needs_node
@pytest.mark.gpu
def test_node_costs_separated_profile_match_oracle(tmp_path):
    pb = PR.separated_probe('uniaxial', 'stephens', 'gsas_cw', simulate=O.oracle_simulate)
    m = O.Model(pb)
    probes = _probes(pb)
    path = js_ctx.write_ctx(pb, str(tmp_path / 'ctx.json'), probes=[list(map(float, p)) for p in probes])
    out, r = _node(path, 'costs')
    assert 'costs' in out, (out, r.stderr)
    for c, g in enumerate(out['costs']):
        cpu = m.cost(probes[c % len(probes)])
        assert abs(g - cpu) / max(1.0, cpu) <= 1e-5


@needs_node
@pytest.mark.gpu
def test_node_costs_two_datasets_match_oracle(tmp_path):
    """A context with an X-ray and a neutron dataset (variables acting in one or both,
    ctx.info[j].ds): the costs are the sums of the oracle's per-dataset costs."""
    pbs = PR.xray_neutron_pair(simulate=O.oracle_simulate)
    ms = [O.Model(pb) for pb in pbs]
    probes = _probes(pbs[0])
    path = js_ctx.write_ctx(pbs, str(tmp_path / 'ctx.json'), probes=[list(map(float, p)) for p in probes])
    out, r = _node(path, 'costs')
    assert 'costs' in out, (out, r.stderr)
    for c, g in enumerate(out['costs']):
        cpu = sum(m.cost(probes[c % len(probes)]) for m in ms)
        assert abs(g - cpu) / max(1.0, cpu) <= 1e-5


@needs_node
def test_compile_two_dataset_context(tmp_path):
    pbs = PR.xray_neutron_pair()
    for pb in pbs:
        pb.y = np.ones_like(pb.tth)
    path = js_ctx.write_ctx(pbs, str(tmp_path / 'ctx.json'))
    out, r = _node(path, 'compile')
    assert 'compile' in out, r.stderr
    c = out['compile']
    infos = []
    for pb in pbs:
        with pt.Plan(pb, device=False) as pl:
            infos.append(pl.info)
    assert c['datasets'] == 2 and c['nv'] == infos[0]['nv']
    assert c['pointCount'] == sum(i['point_count'] for i in infos)
    assert c['reflectionCount'] == sum(i['reflection_count'] for i in infos)


@needs_node
@pytest.mark.gpu
def test_node_costs_lebail_phase_match_oracle(tmp_path):
    """A Le Bail phase next to a Rietveld phase: the free intensities are solved per trial."""
    pb = PR.lebail_probe(mixed=True, simulate=O.oracle_simulate)
    m = O.Model(pb)
    probes = _probes(pb)
    path = js_ctx.write_ctx(pb, str(tmp_path / 'ctx.json'), probes=[list(map(float, p)) for p in probes])
    out, r = _node(path, 'costs')
    assert 'costs' in out, (out, r.stderr)
    for c, g in enumerate(out['costs']):
        cpu = m.cost(probes[c % len(probes)])
        assert abs(g - cpu) / max(1.0, cpu) <= 1e-5


@needs_node
def test_compile_lebail_context(tmp_path):
    pb = PR.lebail_probe(mixed=True)
    pb.y = np.ones_like(pb.tth)
    path = js_ctx.write_ctx(pb, str(tmp_path / 'ctx.json'))
    out, r = _node(path, 'compile')
    assert 'compile' in out, r.stderr
    with pt.Plan(pb, device=False) as pl:
        info = pl.info
    assert out['compile']['reflectionCount'] == info['reflection_count']


@needs_node
def test_cost_pipe_answers_with_the_oracle(tmp_path):
    """The FIFO cost server the run() tests use: Node's synchronous ctx.cost gets the
    oracle's FP64 cost of any vector, bit for bit."""
    pb = PR.triclinic_probe('tch_aniso', simulate=O.oracle_simulate)
    m = O.Model(pb)
    probes = _probes(pb)
    with js_ctx.OracleCostPipe(m, tmp_path) as pipe:
        path = js_ctx.write_ctx(pb, str(tmp_path / 'ctx.json'), costPipe=pipe.paths,
                                probes=[list(map(float, p)) for p in probes])
        out, r = _node(path, 'cost')
    assert 'cost' in out, (out, r.stderr)
    want = [m.cost(pb.start_vector())] + [m.cost(p) for p in probes]
    assert out['cost'] == want


def _python_terms(pb):
    return sorted((int(k), int(j), float(c)) for (k, j, c) in pb.terms)


@needs_node
@pytest.mark.parametrize('case', ['page_joint_case', 'page_scoped_case'])
def test_context_constraints_flatten_like_the_problem(case, tmp_path):
    """The page's joint and scoped contexts (tests/webgpu-pt.js:83-106) carry their
    atomic constraints and cross-dataset link as ctx.constraints.links; the Node host
    flattens them into exactly the model-parameter terms and bases of the Python problem."""
    pbs = getattr(PR, case)()
    for pb in pbs:
        pb.y = np.ones_like(pb.tth)
    path = js_ctx.write_ctx(pbs, str(tmp_path / 'ctx.json'))
    out, r = _node(path, 'flat')
    assert 'flat' in out, r.stderr
    assert len(out['flat']) == len(pbs)
    for pb, f in zip(pbs, out['flat']):
        js_terms = sorted(zip(f['termParam'], f['termVar'], f['termCoeff']))
        assert js_terms == _python_terms(pb)
        assert np.array_equal(np.array(f['paramBase']), pb.base)
        assert f['resolutionType'] == (PR.PROFILE[pb.resolution_type] if pb.resolution_type else 0)
        assert [p['free'] for p in f['phases']] == [0 if ph.method == 'rietveld' else 1 for ph in pb.phases]


@needs_node
@pytest.mark.gpu
@pytest.mark.parametrize('case', ['page_joint_case', 'page_scoped_case'])
def test_node_costs_page_joint_and_scoped_cases(case, tmp_path):
    """The page's joint and scoped cases through the Node host: costs() of the probe
    vectors equal the oracle's sums over the datasets to 1e-5."""
    pbs = getattr(PR, case)(simulate=O.oracle_simulate)
    ms = [O.Model(pb) for pb in pbs]
    x0 = pbs[0].start_vector()
    probes = [x0, np.clip(x0 + .1 * np.array(pbs[0].scale), pbs[0].lo, pbs[0].hi)]
    path = js_ctx.write_ctx(pbs, str(tmp_path / 'ctx.json'), probes=[list(map(float, p)) for p in probes])
    out, r = _node(path, 'costs')
    assert 'costs' in out, (out, r.stderr)
    for c, g in enumerate(out['costs']):
        cpu = sum(m.cost(probes[c % len(probes)]) for m in ms)
        assert abs(g - cpu) / max(1.0, cpu) <= 1e-5


@needs_node
@pytest.mark.gpu
def test_node_normal_equations_match_oracle(tmp_path):
    """session.normal(x) through the Node host: A, g and chi^2 of one LM step against the
    oracle's normal() (the columns the reference's Model.derivatives builds)."""
    pb = PR.pbso4_pnma(simulate=O.oracle_simulate, chains=16)
    m = O.Model(pb)
    A0, g0, f0, _ = m.normal(pb.start_vector())
    path = js_ctx.write_ctx(pb, str(tmp_path / 'ctx.json'))
    out, r = _node(path, 'normal')
    assert 'A' in out, (out, r.stderr)
    A, g = np.array(out['A']).reshape(pb.nv, pb.nv), np.array(out['g'])
    d = np.sqrt(np.diag(A0))
    assert np.all(np.abs(A - A0) <= 1e-6 * np.outer(d, d))
    assert np.all(np.abs(g - g0) <= 1e-6 * d * np.sqrt(f0))
    assert abs(out['f'] - f0) <= 1e-12 * f0
