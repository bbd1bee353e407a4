"""The library's host plan compiler (pt_plan_compile, no device needed) against the
oracle: the reflection envelope of js/pt_gpu.js:100-131 must equal the oracle's
generateReflections list bit for bit (h, k, l, multiplicity and order), in each of the
three branches (fixed cell; monotone envelope at the upper cell edges; full Miller box)."""
import ctypes
import math

import numpy as np
import pytest

import oracle_lib as O
from powder6_amd import problems as PR
from powder6_amd import pt


def _dmin_limits(pb):
    """dmin at the limit magnitudes of zero/disp/transp (rietveld_core.js:238-245)."""
    def mag(name):
        k = pb.index[name]
        los = his = pb.base[k]
        for (kk, j, c) in pb.terms:
            if kk == k:
                los += c * (pb.lo[j] if c > 0 else pb.hi[j])
                his += c * (pb.hi[j] if c > 0 else pb.lo[j])
        return max(abs(los), abs(his))
    cp = O.CProblem(pb) if hasattr(O, 'CProblem') else pt.CProblem(pb)
    return O.lib().or_reflection_dmin(ctypes.addressof(cp.struct), mag('zero'), mag('disp'), mag('transp'))


def _upper(pb, name):
    k = pb.index[name]
    for (kk, j, c) in pb.terms:
        if kk == k:
            return pb.hi[j]
    return pb.base[k]


CASES = {
    'p21c_fixed_cell': lambda: PR.molecular_p21c(),
    'pnma_monotone': lambda: PR.pbso4_pnma(),
    'triclinic_box': lambda: PR.triclinic_probe('tch_aniso'),
    'nacl_cubic': lambda: PR.nacl_fm3m(),
}


@pytest.mark.parametrize('case', sorted(CASES))
def test_envelope_matches_oracle(case):
    pb = CASES[case]()
    pb.y = np.ones_like(pb.tth)
    dmin = _dmin_limits(pb)
    ph = pb.phases[0]
    p = pb.values()
    cell = [p[pb.index['p1.' + n]] for n in ('a', 'b', 'c', 'alpha', 'beta', 'gamma')]
    if ph.system == 'cubic':
        cell[1] = cell[2] = cell[0]
    moves = any(n in pb.vars for n in ('p1.a', 'p1.b', 'p1.c', 'p1.alpha', 'p1.beta', 'p1.gamma'))
    if not moves:
        ref, _ = O.generate_reflections(cell, ph.op_r, ph.op_t, dmin)
    elif ph.system in ('orthorhombic', 'cubic', 'tetragonal'):
        env = list(cell)
        for i, n in enumerate('abc'):
            env[i] = _upper(pb, 'p1.' + n) if 'p1.' + n in pb.vars else cell[i]
        if ph.system == 'cubic':
            env[1] = env[2] = env[0]
        ref, _ = O.generate_reflections(env, ph.op_r, ph.op_t, dmin)
    else:
        box = [int(math.floor(_upper(pb, 'p1.' + n) / dmin)) for n in 'abc']
        ref, _ = O.generate_reflections(cell, ph.op_r, ph.op_t, dmin, box=box)
    with pt.Plan(pb, device=False) as pl:
        got = pl.reflections(0)
        assert pl.info['reflection_count'] == len(got)
    assert len(got) == len(ref) > 0
    assert np.array_equal(got, ref)


def test_envelope_contains_every_trial_list():
    """Any cell inside the limits enumerates a subset of the envelope (pt_gpu.js:110-131)."""
    pb = PR.pbso4_pnma()
    pb.y = np.ones_like(pb.tth)
    m = O.Model(pb)
    with pt.Plan(pb, device=False) as pl:
        keys = {tuple(r[:3]) for r in pl.reflections(0)}
    rng = np.random.default_rng(3)
    for _ in range(5):
        x = pb.start_vector()
        for n in ('p1.a', 'p1.b', 'p1.c', 'zero'):
            j = pb.var_index(n)
            x[j] = rng.uniform(pb.lo[j], pb.hi[j])
        hk, _, _, _ = m.reflections(x)
        assert all(tuple(r[:3]) in keys for r in hk)


def test_budget_and_argument_errors():
    pb = PR.triclinic_probe('tch_aniso')
    pb.y = np.ones_like(pb.tth)
    for n in ('p1.a', 'p1.b', 'p1.c'):
        pb.hi[pb.var_index(n)] = 400.0    # Miller box beyond 1e6 triples
    with pytest.raises(pt.PTError, match='budget'):
        pt.Plan(pb, device=False)
    pb = PR.triclinic_probe('tch_aniso')
    pb.y = np.ones_like(pb.tth)
    pb.hi[pb.var_index('p1.b')] = math.inf
    with pytest.raises(pt.PTError, match='finite positive upper bounds'):
        pt.Plan(pb, device=False)
    pb = PR.triclinic_probe('tch_aniso')
    pb.y = np.ones_like(pb.tth)
    pb.tth = pb.tth[::-1].copy()
    with pytest.raises(pt.PTError, match='ascending'):
        pt.Plan(pb, device=False)
