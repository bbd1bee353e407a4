"""HIP path vs the FP64 oracle: chi^2 of the plain, batched and PT forms at <= 1e-5
relative (the bar the PT trial evaluation must meet), patterns, and PT invariants."""
import math

import numpy as np
import pytest

import oracle_lib as O
from powder6_amd import problems as PR
from powder6_amd import pt

pytestmark = pytest.mark.gpu
TOL = 1e-5          # chi^2 relative error against the FP64 oracle


def _probes(pb, n, seed, spread=2.0):
    """n trial vectors around the start: every variable moved by up to `spread` PT
    scales, clipped to its box."""
    rng = np.random.default_rng(seed)
    x0 = pb.start_vector()
    lo, hi, sc = np.array(pb.lo), np.array(pb.hi), np.array(pb.scale)
    X = x0 + rng.uniform(-spread, spread, size=(n, pb.nv)) * sc
    X = np.minimum(hi, np.maximum(lo, X))
    X[0] = x0
    return X


def _rel(a, b):
    return abs(a - b) / max(1.0, abs(b))


@pytest.fixture(scope='module')
def nacl():
    return PR.nacl_fm3m(simulate=O.oracle_simulate)


@pytest.fixture(scope='module')
def pbso4():
    return PR.pbso4_pnma(simulate=O.oracle_simulate, chains=16, steps_per_sweep=8)


@pytest.mark.parametrize('name', ['nacl', 'pbso4'])
def test_batched_chi2_matches_oracle(name, request):
    pb = request.getfixturevalue(name)
    m = O.Model(pb)
    X = _probes(pb, 24, 5)
    with pt.Plan(pb) as pl:
        g = pl.eval(X, pt.FORM_BATCHED)
    cpu = np.array([m.cost(x) for x in X])
    errs = [_rel(a, b) for a, b in zip(g, cpu)]
    print(name, 'max rel err', max(errs))
    assert max(errs) <= TOL, list(zip(g, cpu, errs))


@pytest.mark.parametrize('name', ['nacl', 'pbso4'])
def test_plain_chi2_matches_oracle(name, request):
    pb = request.getfixturevalue(name)
    m = O.Model(pb)
    X = _probes(pb, 6, 9)
    with pt.Plan(pb) as pl:
        g = pl.eval(X, pt.FORM_PLAIN)
        b = pl.eval(X, pt.FORM_BATCHED)
    cpu = np.array([m.cost(x) for x in X])
    errs = [_rel(a, c) for a, c in zip(g, cpu)]
    print(name, 'plain max rel err', max(errs))
    assert max(errs) <= TOL
    assert np.allclose(g, b, rtol=2e-6)


@pytest.mark.parametrize('profile', ['tch_aniso', 'simple_pvoigt', 'split_pvoigt'])
def test_triclinic_probe_families(profile):
    """tests/webgpu-pt.js:47-55 family, at 1e-5 instead of 1e-3."""
    pb = PR.triclinic_probe(profile, simulate=O.oracle_simulate)
    m = O.Model(pb)
    x0 = pb.start_vector()
    X = [x0] + [np.clip(x0 + s * np.array(pb.scale), pb.lo, pb.hi) for s in (-.12, .12, .3)]
    with pt.Plan(pb) as pl:
        g = pl.costs(np.array(X))
    for j, x in enumerate(X):
        cpu = m.cost(x)
        assert _rel(g[j], cpu) <= TOL, (profile, j, g[j], cpu)


def test_pattern_matches_oracle(pbso4):
    m = O.Model(pbso4)
    x = pbso4.start_vector()
    with pt.Plan(pbso4) as pl:
        y = pl.pattern(x)
    yo = m.pattern(x)
    err = np.max(np.abs(y - yo) / np.maximum(1.0, np.abs(yo)))
    print('pattern max rel err', err)
    assert err < 2e-6


def test_pt_batch_invariants(pbso4):
    """tests/webgpu-pt.js:131-136: proposals = chains*sweeps*steps, swaps > 0, best <= initial;
    the best vectors re-scored by the oracle."""
    m = O.Model(pbso4)
    x0 = pbso4.start_vector()
    with pt.Plan(pbso4) as pl:
        pl.initial([x0])
        s0 = pl.batch(1, bootstrap=True)
        nv, stride = pl.nv, pl.stride
        init = s0[3 * nv + pt.S_COST]
        assert _rel(init, m.cost(x0)) <= TOL
        sweeps = 4
        st = pl.batch(sweeps)
    S = st.reshape(pl.chains, stride)
    prop = S[:, 3 * nv + pt.S_PROPOSALS].sum()
    assert prop == pl.chains * sweeps * pl.info['steps_per_sweep']
    assert S[:, 3 * nv + pt.S_SWAPS].sum() > 0
    assert S[:, 3 * nv + pt.S_ATTEMPTS].sum() > 0
    best = S[:, 3 * nv + pt.S_BEST]
    assert best.min() <= init
    for c in range(0, pl.chains, 5):
        xb = S[c, 2 * nv:3 * nv]
        assert _rel(best[c], m.cost(xb)) <= TOL
        xc = S[c, :nv]
        assert _rel(S[c, 3 * nv + pt.S_COST], m.cost(xc)) <= TOL


def test_sweep_graph_replays_match_direct_launches(pbso4):
    """The captured sweep graph, replayed several times with host work in between, gives
    the same chain states as launching every kernel directly (deterministic pipeline)."""
    x0 = pbso4.start_vector()
    with pt.Plan(pbso4) as pl:
        pl.initial([x0])
        pl.batch(1, bootstrap=True)
        graphed = [pl.batch(1) for _ in range(3)]
        pl.initial([x0])
        pl.batch(1, bootstrap=True)
        pl.set_kernel_timing(True)
        direct = [pl.batch(1) for _ in range(3)]
    for g, d in zip(graphed, direct):
        assert np.array_equal(g, d)


@pytest.mark.parametrize('chains', [200, 256])
def test_sweep_parts_on_streams_match_one_stream(chains):
    """With 128 chains or more the sweep graph runs the chains as parts on their own streams
    (200 chains: parts of 72, 72 and 56; 256: four of 64); the timed direct form runs them as
    one part.  Chains are independent between exchanges, so the states are bit-identical."""
    pb = PR.pbso4_pnma(simulate=O.oracle_simulate, chains=chains, steps_per_sweep=8)
    x0 = pb.start_vector()
    with pt.Plan(pb) as pl:
        pl.initial([x0])
        pl.batch(1, bootstrap=True)
        graphed = [pl.batch(1) for _ in range(3)]
        pl.initial([x0])
        pl.batch(1, bootstrap=True)
        pl.set_kernel_timing(True)
        direct = [pl.batch(1) for _ in range(3)]
    for g, d in zip(graphed, direct):
        assert np.array_equal(g, d)
    S = graphed[-1].reshape(chains, -1)
    nv = pb.nv
    assert np.all(S[:, 3 * nv + pt.S_PROPOSALS] == 3 * 8)


def test_tile_shapes_give_identical_states(pbso4):
    """The chi^2 partials are per 512-point group whatever the tile shape, so the PT
    trajectory does not depend on the shape the tuner picks: bit-identical states."""
    x0 = pbso4.start_vector()
    runs = []
    for tiles in (1, 2, 4):
        with pt.Plan(pbso4) as pl:
            pl.tiles(tiles)
            pl.initial([x0])
            pl.batch(1, bootstrap=True)
            runs.append(pl.batch(2))
            assert pl.tiles()[0] == tiles
    assert np.array_equal(runs[0], runs[1]) and np.array_equal(runs[1], runs[2])


def test_batched_eval_chunks_and_repeats(pbso4):
    """Evaluating 17,000 vectors (two passes of the 16,384-slot buffer) equals evaluating
    them in other splits, and a repeat is bit-identical."""
    X = _probes(pbso4, 17000, 21, spread=1.0)
    with pt.Plan(pbso4) as pl:
        a = pl.eval(X, pt.FORM_BATCHED)
        b = pl.eval(X, pt.FORM_BATCHED)
        c = np.concatenate([pl.eval(X[:1000], pt.FORM_BATCHED), pl.eval(X[1000:], pt.FORM_BATCHED)])
    assert np.array_equal(a, b)
    assert np.array_equal(a, c)
