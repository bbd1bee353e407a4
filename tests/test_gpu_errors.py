"""Error semantics at the boundary, as the reference's callers see them:

* an intensity solve that does not converge sets the chain's unsolved flag (state +9), its
  trial costs 3e38, and costs() / batch() / run() fail with "did not converge"
  (js/pt_gpu.js:363, 402, 418) -- forced from the host with a one-iteration solve limit,
  and by a trial with more live free intensities than one solve takes;
* a GPU cost that disagrees with the caller's FP64 model makes run() fail with
  /disagrees/ (js/pt_gpu.js:415; the page's forced case, tests/webgpu-pt.js:143-148);
* shouldStop() ends run() between batches with stopped = true and the best kept
  (js/pt_gpu.js:419);
* run() through the Node host re-scores every candidate with the caller's FP64 cost.
"""
import math

import numpy as np
import pytest

import js_ctx
import oracle_lib as O
from powder6_amd import problems as PR
from powder6_amd import pt
from test_node_addon import _node, needs_node

pytestmark = pytest.mark.gpu


def test_costs_report_a_capped_solve_and_recover():
    pb = PR.lebail_probe(simulate=O.oracle_simulate)
    m = O.Model(pb)
    x0 = pb.start_vector()
    with pt.Plan(pb) as pl:
        pl.solve_limit(1)
        with pytest.raises(pt.PTError, match='did not converge'):
            pl.costs([x0])
        pl.solve_limit(0)
        c = pl.costs([x0])
    ref = m.cost(x0)
    assert np.all(np.abs(c - ref) / max(1.0, ref) <= 1e-5)


def test_batch_sets_the_sticky_unsolved_flag():
    pb = PR.lebail_probe(simulate=O.oracle_simulate)
    x0 = pb.start_vector()
    with pt.Plan(pb) as pl:
        nv, stride = pl.nv, pl.stride
        pl.initial([x0])
        pl.solve_limit(1)
        with pytest.raises(pt.PTError, match='did not converge') as ei:
            pl.batch(1, bootstrap=True)
        S = ei.value.state.reshape(pl.chains, stride)
        assert np.all(S[:, 3 * nv + pt.S_UNSOLVED] == 1.0)
        assert np.all(S[:, 3 * nv + 10] >= 3e38)
        # the flag stays until the states are set again
        pl.solve_limit(0)
        with pytest.raises(pt.PTError, match='did not converge'):
            pl.batch(1)
        pl.initial([x0])
        pl.batch(1, bootstrap=True)
        S = pl.batch(2).reshape(pl.chains, stride)
        assert np.all(S[:, 3 * nv + pt.S_UNSOLVED] == 0.0)
        assert np.all(np.isfinite(S[:, 3 * nv]) & (S[:, 3 * nv] < 3e38))


def _wide_lebail_probe():
    """The reference page's P-1 probe as a Le Bail phase over 8-168 deg: about 290
    reflections have a window in the pattern, more than one trial's solve takes (256)."""
    pb = PR.triclinic_probe('tch_aniso', method='lebail', n_points=2300, span=160.0)
    pb.y = 100.0 + 50.0 * np.cos(pb.tth * 0.37) ** 2
    pb.w = 1.0 / pb.y
    return pb


def test_costs_report_a_trial_beyond_the_live_ceiling():
    pb = _wide_lebail_probe()
    with pt.Plan(pb) as pl:
        with pytest.raises(pt.PTError, match='did not converge'):
            pl.costs([pb.start_vector()])


def _run_ctx(pb, m, **extra):
    ctx = {'problem': pb, 'x': pb.start_vector(), 'cost': m.cost, 'initial_cost': m.cost(pb.start_vector())}
    ctx.update(extra)
    return ctx


def test_run_rejects_a_disagreeing_model():
    """tests/webgpu-pt.js:143-148: the caller's cost + 1000 must be rejected."""
    pb = PR.triclinic_probe('tch_aniso', simulate=O.oracle_simulate)
    m = O.Model(pb)
    ctx = _run_ctx(pb, m, cost=lambda x: m.cost(x) + 1000.0)
    with pytest.raises(pt.PTError, match='disagrees'):
        pt.run(ctx)


def test_run_stops_when_asked():
    pb = PR.triclinic_probe('tch_aniso', simulate=O.oracle_simulate)
    m = O.Model(pb)
    asked = []

    def should_stop():
        asked.append(1)
        return len(asked) >= 3

    r = pt.run(_run_ctx(pb, m, should_stop=should_stop))
    assert r['stopped'] is True
    assert 0 < r['pt']['sweeps'] < pb.sweeps
    assert r['pt']['proposals'] == pb.chains * r['pt']['sweeps'] * pb.nv
    assert r['pt']['bestChi2'] <= r['pt']['initialChi2']
    assert m.cost(r['x']) == r['pt']['bestChi2']


def test_run_reports_unsolved_trials():
    pb = _wide_lebail_probe()
    m = O.Model(pb)
    with pytest.raises(pt.PTError, match='did not converge'):
        pt.run(_run_ctx(pb, m))


@needs_node
def test_node_run_pt_rescored_with_the_fp64_model(tmp_path):
    """tests/webgpu-pt.js:131-136 through the Node host, with ctx.cost the FP64 oracle for
    every vector: proposals = chains x sweeps x nv, swaps > 0, the best strictly below the
    initial chi^2 and equal to the oracle's cost of the returned x."""
    pb = PR.triclinic_probe('tch_aniso', simulate=O.oracle_simulate)
    m = O.Model(pb)
    init = m.cost(pb.start_vector())
    with js_ctx.OracleCostPipe(m, tmp_path) as pipe:
        path = js_ctx.write_ctx(pb, str(tmp_path / 'ctx.json'), initial_cost=init, costPipe=pipe.paths)
        out, r = _node(path, 'run')
    assert 'pt' in out, (out, r.stderr)
    p = out['pt']
    assert p['proposals'] == pb.chains * pb.sweeps * pb.nv
    assert p['swaps'] > 0 and p['swapAttempts'] > 0
    assert p['validationError'] <= 1e-5
    assert p['bestChi2'] < init
    assert m.cost(np.array(out['x'])) == p['bestChi2']
    assert out['costCalls'] > pb.chains and out['iters'] >= 1


@needs_node
def test_node_run_rejects_a_disagreeing_model(tmp_path):
    pb = PR.triclinic_probe('tch_aniso', simulate=O.oracle_simulate)
    m = O.Model(pb)
    with js_ctx.OracleCostPipe(m, tmp_path) as pipe:
        path = js_ctx.write_ctx(pb, str(tmp_path / 'ctx.json'), initial_cost=m.cost(pb.start_vector()),
                                costPipe=pipe.paths, costOffset=1000.0)
        out, r = _node(path, 'run')
    assert 'disagrees' in out.get('error', ''), (out, r.stderr)


@needs_node
def test_node_run_stops_when_asked(tmp_path):
    pb = PR.triclinic_probe('tch_aniso', simulate=O.oracle_simulate)
    m = O.Model(pb)
    init = m.cost(pb.start_vector())
    with js_ctx.OracleCostPipe(m, tmp_path) as pipe:
        path = js_ctx.write_ctx(pb, str(tmp_path / 'ctx.json'), initial_cost=init, costPipe=pipe.paths, stopAfter=3)
        out, r = _node(path, 'run')
    assert 'pt' in out, (out, r.stderr)
    assert out['stopped'] is True
    assert 0 < out['pt']['sweeps'] < pb.sweeps
    assert out['pt']['bestChi2'] <= init


@needs_node
def test_node_reports_unsolved_trials(tmp_path):
    pb = _wide_lebail_probe()
    x0 = pb.start_vector()
    path = js_ctx.write_ctx(pb, str(tmp_path / 'ctx.json'), probes=[list(map(float, x0))])
    out, r = _node(path, 'costs')
    assert 'did not converge' in out.get('error', ''), (out, r.stderr)
