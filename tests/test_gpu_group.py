"""Groups of independent problems in shared launches (pt_group_*): chi^2 of every member
against the FP64 oracle, the group's PT trajectories against the same problems run one
plan at a time, the tile-shape tuner, and PT invariants over uneven members."""
import numpy as np
import pytest

import oracle_lib as O
from powder6_amd import problems as PR
from powder6_amd import pt

pytestmark = pytest.mark.gpu
TOL = 1e-5
# members of the 48-pattern uneven batch with different groups, cells, point and
# reflection counts (cubic / tetragonal / monoclinic, 1-12 sites)
MEMBERS = [0, 4, 7, 12, 22, 33, 41, 47]


def _rel(a, b):
    return abs(a - b) / max(1.0, abs(b))


@pytest.fixture(scope='module')
def members():
    return [PR.uneven_pattern(i, simulate=O.oracle_simulate, chains=6) for i in MEMBERS]


def _probe(pb, seed):
    rng = np.random.default_rng(seed)
    x = np.array(pb.start_vector()) + rng.uniform(-1, 1, size=pb.nv) * np.array(pb.scale)
    return np.minimum(np.array(pb.hi), np.maximum(np.array(pb.lo), x))


def test_group_costs_match_oracle(members):
    with pt.Group(members) as g:
        probes = []
        for m, pb in enumerate(members):
            X = np.stack([pb.start_vector(), _probe(pb, 100 + m)])
            probes.append(X)
            g.initial(m, X)
        costs = g.costs()
    worst = 0.0
    for m, pb in enumerate(members):
        model = O.Model(pb)
        ref = [model.cost(x) for x in probes[m]]
        model.close()
        for c, v in enumerate(costs[m]):
            e = _rel(v, ref[c % 2])
            worst = max(worst, e)
            assert e <= TOL, (MEMBERS[m], c, v, ref[c % 2])
    print('group chi2 max rel err vs FP64', worst)


def test_group_matches_single_plans(members):
    """Same seeds, same sweep length, same tile shape: the group's states equal those of
    the members run as separate plans, bit for bit."""
    steps = 5
    with pt.Group(members, steps_per_sweep=steps) as g:
        g.tiles(2)
        for m, pb in enumerate(members):
            g.initial(m, [pb.start_vector()])
        g.batch(1, bootstrap=True)
        got = g.batch(3)
    for m, pb in enumerate(members):
        keep = pb.steps_per_sweep
        pb.steps_per_sweep = steps
        try:
            with pt.Plan(pb) as pl:
                pl.tiles(2)
                pl.initial([pb.start_vector()])
                pl.batch(1, bootstrap=True)
                want = pl.batch(3)
        finally:
            pb.steps_per_sweep = keep
        assert np.array_equal(got[m], want), (MEMBERS[m], np.max(np.abs(got[m] - want)))


def test_group_parts_match_single_plans():
    """With 256 chains (eight members of 32) the group sweep runs the members as four parts
    on their own streams; the states still equal the members run as separate plans."""
    pbs = [PR.uneven_pattern(i, simulate=O.oracle_simulate, chains=32) for i in MEMBERS]
    steps = 4
    with pt.Group(pbs, steps_per_sweep=steps) as g:
        g.tiles(1)
        for m, pb in enumerate(pbs):
            g.initial(m, [pb.start_vector()])
        g.batch(1, bootstrap=True)
        got = g.batch(2)
    for m, pb in enumerate(pbs):
        pb.steps_per_sweep = steps
        with pt.Plan(pb) as pl:
            pl.tiles(1)
            pl.initial([pb.start_vector()])
            pl.batch(1, bootstrap=True)
            want = pl.batch(2)
        assert np.array_equal(got[m], want), (MEMBERS[m], np.max(np.abs(got[m] - want)))


def test_group_tuner_and_invariants(members):
    with pt.Group(members) as g:
        for m, pb in enumerate(members):
            g.initial(m, [pb.start_vector()])
        g.batch(1, bootstrap=True)
        init = [s.copy() for s in g.batch(0)]
        tiles, ms = g.tiles(-1)
        assert tiles in (1, 2, 4) and np.all(ms > 0), (tiles, ms)
        states = g.batch(2)
        print('group tiles', tiles, 'tuner ms', ms, 'sweep ms', g.timing() / 2)
        for m, pb in enumerate(members):
            info = g.members[m]
            nv, stride, C = info['nv'], info['stride'], info['chains']
            S = states[m].reshape(C, stride)
            I = init[m].reshape(C, stride)
            steps = info['steps_per_sweep']
            assert np.all(S[:, 3 * nv + 4] == 2 * steps)                # proposals
            assert np.all(S[:, 3 * nv + 1] <= I[:, 3 * nv + 1])          # best never rises
            assert np.all(S[:-1, 3 * nv + 6] >= 1)                      # every lower chain tried a swap
            assert np.all(S[:, 3 * nv + 5] <= S[:, 3 * nv + 6])
            x = S[:, :nv]
            assert np.all(x >= np.array(pb.lo) - 1e-12) and np.all(x <= np.array(pb.hi) + 1e-12)
