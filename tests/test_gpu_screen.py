"""Surrogate-screened PT on the P2_1/c molecular problem: with the network's picks held
back it is plain PT bit for bit; with them on, the exact chi^2 still decides every step."""
import numpy as np
import pytest

import oracle_lib as O
from powder6_amd import problems as PR
from powder6_amd import pt

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def mol():
    return PR.molecular_p21c(simulate=O.oracle_simulate, chains=64, shared_b=True)


def _run(pb, sweeps, **screen):
    with pt.Plan(pb) as pl:
        if screen:
            pl.screen(**screen)
        pl.initial([pb.start_vector()])
        pl.batch(1, bootstrap=True)
        st = pl.batch(sweeps)
        stats = pl.screen_stats() if screen else None
    return st, stats


def test_screen_held_back_is_plain_pt(mol):
    plain, _ = _run(mol, 2)
    held, stats = _run(mol, 2, warmup_iters=10 ** 9, batch=128, capacity=8192)
    assert stats['iterations'] > 0 and stats['picked'] == 0
    assert np.array_equal(plain, held)


def test_screened_pt_invariants(mol):
    st, stats = _run(mol, 3, warmup_iters=20, batch=128, capacity=8192)
    nv, C = mol.nv, mol.chains
    S = st.reshape(C, 3 * nv + 12)
    print('screen stats', stats)
    assert stats['iterations'] > 20 and stats['picked'] > 0 and stats['ring'] == min(8192, 3 * nv * C)
    assert np.all(S[:, 3 * nv + 4] == 3 * nv)                     # one exact evaluation per step
    x = S[:, :nv]
    assert np.all(x >= np.array(mol.lo)) and np.all(x <= np.array(mol.hi))
    m = O.Model(mol)
    for c in range(0, C, 16):                                     # the recorded best is exact
        best = S[c, 3 * nv + 1]
        if best < 1e38:
            ref = m.cost(S[c, 2 * nv:3 * nv])
            assert abs(best - ref) <= 1e-5 * max(1.0, ref), (c, best, ref)


def _far_start(pb, seed, frac=0.6):
    """Every atom coordinate moved by up to `frac` of its half box: a start well above the
    noise floor, so that PT has a descent to make."""
    rng = np.random.default_rng(seed)
    x = np.array(pb.start_vector())
    for j, n in enumerate(pb.vars):
        if n.endswith(('.x', '.y', '.z')):
            x[j] = np.clip(x[j] + rng.uniform(-1, 1) * frac * (pb.hi[j] - pb.lo[j]) / 2, pb.lo[j], pb.hi[j])
    return x


def _spearman(a, b):
    ra, rb = np.argsort(np.argsort(a)), np.argsort(np.argsort(b))
    return float(np.corrcoef(ra, rb)[0, 1])


def test_trained_network_ranks_trials_and_pt_keeps_its_quality(mol):
    """What the trained network gains: on held-out trial vectors around the chains' states
    its predicted log Rwp ranks the exact values (Spearman) far better than the untrained
    network does.  At the PT level the screened run, with the same number of exact
    evaluations, reaches a best Rwp within 2% of plain PT (both printed)."""
    x0 = _far_start(mol, 3)
    wss = float(np.sum(mol.w * mol.y ** 2))
    cfg = dict(warmup_iters=50, batch=256, capacity=16384, margin=0.002, cold_fraction=0.5)
    sweeps = 4
    nv, C = mol.nv, mol.chains
    rng = np.random.default_rng(11)
    with pt.Plan(mol) as pl:
        pl.screen(**cfg)
        pl.initial([x0])
        pl.batch(1, bootstrap=True)
        # untrained network on probes around the start
        base = np.repeat(np.atleast_2d(x0), 256, axis=0)
        scale = np.array(mol.scale)
        P0 = np.clip(base + rng.normal(size=base.shape) * scale * rng.uniform(0, 3, size=(256, 1)), mol.lo, mol.hi)
        untrained = _spearman(pl.screen_predict(P0), pl.eval(P0, pt.FORM_BATCHED))
        st = pl.batch(sweeps).reshape(C, -1)
        screened_best = float(np.min(st[:, 3 * nv + 1]))
        stats = pl.screen_stats()
        # held-out probes around the cold half of the chains' final states
        picks = rng.integers(0, C // 2, size=256)
        B = st[picks, :nv]
        P1 = np.clip(B + rng.normal(size=B.shape) * scale * rng.uniform(0, 3, size=(256, 1)), mol.lo, mol.hi)
        chi = pl.eval(P1, pt.FORM_BATCHED)
        trained = _spearman(pl.screen_predict(P1), chi)
    with pt.Plan(mol) as pl:
        pl.initial([x0])
        pl.batch(1, bootstrap=True)
        plain_best = float(np.min(pl.batch(sweeps).reshape(C, -1)[:, 3 * nv + 1]))
    rwp = lambda c: 100 * np.sqrt(c / wss)
    print('rank correlation predicted vs exact log Rwp: untrained %.3f, trained %.3f; stats %s' %
          (untrained, trained, stats))
    print('best Rwp after %d sweeps at equal exact evaluations: plain %.3f%%, screened %.3f%%' %
          (sweeps, rwp(plain_best), rwp(screened_best)))
    assert trained > 0.8 and trained > untrained + 0.3
    assert rwp(screened_best) <= 1.02 * rwp(plain_best)


def test_skip_screening_saves_exact_evaluations_at_equal_quality(mol):
    """Skip mode: a chain's proposal whose predicted Metropolis acceptance (after a margin
    in log Rwp) is below skip_eps is rejected without an exact evaluation.  From a far start
    the skipping run reaches a best Rwp within 2% of plain PT's with fewer exact
    evaluations, and every recorded best is still the exact chi^2 of its vector."""
    x0 = _far_start(mol, 5)
    wss = float(np.sum(mol.w * mol.y ** 2))
    sweeps = 4
    nv, C = mol.nv, mol.chains
    cfg = dict(oversample=1, warmup_iters=50, batch=256, capacity=16384, cold_fraction=1.0,
               skip_eps=1e-2, skip_margin=0.02)
    with pt.Plan(mol) as pl:
        pl.screen(**cfg)
        pl.initial([x0])
        pl.batch(1, bootstrap=True)
        st = pl.batch(sweeps).reshape(C, -1)
        stats = pl.screen_stats()
    with pt.Plan(mol) as pl:
        pl.initial([x0])
        pl.batch(1, bootstrap=True)
        plain = pl.batch(sweeps).reshape(C, -1)
    proposals = float(np.sum(st[:, 3 * nv + 4]))
    rwp = lambda c: 100 * np.sqrt(c / wss)
    skip_best, plain_best = float(np.min(st[:, 3 * nv + 1])), float(np.min(plain[:, 3 * nv + 1]))
    print('skip stats %s; exact evaluations %.1f%% of proposals; best Rwp plain %.3f%%, skipping %.3f%%' %
          (stats, 100 * (1 - stats['skipped'] / proposals), rwp(plain_best), rwp(skip_best)))
    assert stats['picked'] == 0 and 0 < stats['skipped'] < proposals
    assert np.all(st[:, 3 * nv + 4] == sweeps * nv)               # a skip still counts as a step
    assert rwp(skip_best) <= 1.02 * rwp(plain_best)
    m = O.Model(mol)
    for c in range(0, C, 16):
        best = st[c, 3 * nv + 1]
        if best < 1e38:
            ref = m.cost(st[c, 2 * nv:3 * nv])
            assert abs(best - ref) <= 1e-5 * max(1.0, ref), (c, best, ref)


def test_in_step_skip_test_without_skips_is_plain_pt(mol):
    """One proposal per chain and training once per sweep: the skip test runs inside the
    PT setup and the sweep keeps its parts.  With skipping off the network still trains
    (train_iters Adam steps before every sweep once the ring holds a minibatch: from the
    second sweep on) but takes no part in the PT steps, so the states are plain PT's bit for
    bit."""
    plain, _ = _run(mol, 2)
    fused, stats = _run(mol, 2, oversample=1, train_every=mol.nv, train_iters=3, warmup_iters=0, batch=128, capacity=8192,
                        skip_eps=0.0)
    assert stats['iterations'] == 3 and stats['skipped'] == 0
    assert stats['ring'] == min(8192, 2 * mol.nv * mol.chains)   # every evaluated trial recorded
    assert np.array_equal(plain, fused)


def test_in_step_skip_test_skips_and_keeps_exact_costs(mol):
    """The in-step skip test from a far start: after the warm-up some proposals are
    rejected unevaluated (counted by the plan's trial counts), every step still counts, the
    best Rwp after four sweeps stays within 2% of plain PT's and every recorded best is the
    exact chi^2 of its vector."""
    x0 = _far_start(mol, 5)
    wss = float(np.sum(mol.w * mol.y ** 2))
    sweeps, nv, C = 4, mol.nv, mol.chains
    cfg = dict(oversample=1, train_every=nv, train_iters=12, warmup_iters=24, batch=256, capacity=16384, cold_fraction=1.0,
               skip_eps=0.1, skip_margin=0.005)
    with pt.Plan(mol) as pl:
        pl.screen(**cfg)
        pl.initial([x0])
        pl.batch(1, bootstrap=True)
        st = pl.batch(sweeps).reshape(C, -1)
        stats = pl.screen_stats()
        evaluated, skipped = pl.trial_counts()
    with pt.Plan(mol) as pl:
        pl.initial([x0])
        pl.batch(1, bootstrap=True)
        plain = pl.batch(sweeps).reshape(C, -1)
    rwp = lambda c: 100 * np.sqrt(c / wss)
    skip_best, plain_best = float(np.min(st[:, 3 * nv + 1])), float(np.min(plain[:, 3 * nv + 1]))
    print('in-step skip stats %s; evaluated %d, not evaluated %d; best Rwp plain %.3f%%, skipping %.3f%%' %
          (stats, evaluated, skipped, rwp(plain_best), rwp(skip_best)))
    assert stats['iterations'] == 12 * (sweeps - 1) and stats['picked'] == 0 and stats['skipped'] > 0
    assert evaluated + skipped == sweeps * nv * C and skipped >= stats['skipped']
    assert np.all(st[:, 3 * nv + 4] == sweeps * nv)
    assert rwp(skip_best) <= 1.02 * rwp(plain_best)
    m = O.Model(mol)
    for c in range(0, C, 8):
        best = st[c, 3 * nv + 1]
        if best < 1e38:
            ref = m.cost(st[c, 2 * nv:3 * nv])
            assert abs(best - ref) <= 1e-5 * max(1.0, ref), (c, best, ref)
