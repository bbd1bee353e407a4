"""The shared-profile MFMA form (PT_FORM_MATRIX): only structure variables move, the
pattern is P I with FP16 hi/lo operands on MFMA.  chi^2 against the FP64 oracle at 1e-5,
against the VALU form, patterns, chunking, and the refusal of non-structure moves."""
import numpy as np
import pytest

import oracle_lib as O
from powder6_amd import problems as PR
from powder6_amd import pt

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope='module')
def atoms_only():
    return PR.pbso4_pnma(simulate=O.oracle_simulate, chains=16, steps_per_sweep=8, atoms_only=True)


def _struct_probes(pb, n, seed, spread=2.0):
    rng = np.random.default_rng(seed)
    x0 = pb.start_vector()
    X = x0 + rng.uniform(-spread, spread, size=(n, pb.nv)) * np.array(pb.scale)
    X = np.minimum(np.array(pb.hi), np.maximum(np.array(pb.lo), X))
    X[0] = x0
    return X


def test_matrix_chi2_matches_oracle(atoms_only):
    pb = atoms_only
    m = O.Model(pb)
    X = _struct_probes(pb, 48, 3)
    with pt.Plan(pb) as pl:
        g = pl.eval(X, pt.FORM_MATRIX)
        b = pl.eval(X, pt.FORM_BATCHED)
    cpu = np.array([m.cost(x) for x in X])
    err = np.abs(g - cpu) / np.maximum(1.0, cpu)
    print('matrix form max rel err vs FP64', err.max(), 'vs VALU form', np.max(np.abs(g - b) / b))
    assert err.max() <= TOL
    assert np.max(np.abs(g - b) / b) <= 2e-6


def test_matrix_pattern_matches_oracle(atoms_only):
    pb = atoms_only
    m = O.Model(pb)
    x = _struct_probes(pb, 4, 7)[2]
    with pt.Plan(pb) as pl:
        y = pl.matrix_pattern(x)
    yo = m.pattern(x)
    err = np.max(np.abs(y - yo) / np.maximum(1.0, np.abs(yo)))
    print('matrix pattern max rel err', err)
    assert err < 2e-6


def test_matrix_chunks_and_repeats(atoms_only):
    """Repeats are identical; 33,000 vectors run as two passes (32,768 + 232, the second
    one's pair buffers the other half of the double buffer) and agree with the same
    vectors evaluated in other splits."""
    X = _struct_probes(atoms_only, 33000, 11, spread=1.0)
    with pt.Plan(atoms_only) as pl:
        a = pl.eval(X, pt.FORM_MATRIX)
        b = pl.eval(X, pt.FORM_MATRIX)
        c = np.concatenate([pl.eval(X[:100], pt.FORM_MATRIX), pl.eval(X[100:], pt.FORM_MATRIX)])
    assert np.array_equal(a, b)
    assert np.allclose(a, c, rtol=1e-6)     # per-trial scaling is per trial; tiles differ only by padding


def test_matrix_refuses_profile_moves():
    pb = PR.pbso4_pnma(simulate=O.oracle_simulate, chains=16, steps_per_sweep=8)
    X = pb.start_vector()[None, :].repeat(2, axis=0)
    X[1, pb.var_index('p1.a')] += 0.01
    with pt.Plan(pb) as pl:
        pl.eval(X[:1], pt.FORM_MATRIX)
        with pytest.raises(pt.PTError, match='only structure variables'):
            pl.eval(X, pt.FORM_MATRIX)


def test_bf16_pairs_against_fp16_pairs(atoms_only):
    """The precision choice of the matrix form, measured: the same product with BF16
    hi/lo pairs (16 significant bits) against FP16 pairs (22 bits), both against the FP64
    oracle.  FP16 meets the 1e-5 bar; BF16's error is larger, and both are printed."""
    pb = atoms_only
    m = O.Model(pb)
    X = _struct_probes(pb, 48, 5)
    with pt.Plan(pb) as pl:
        g16 = pl.eval(X, pt.FORM_MATRIX)
        gb = pl.eval(X, pt.FORM_MATRIX_BF16)
        g16b = pl.eval(X, pt.FORM_MATRIX)          # switching back refills P in FP16
    cpu = np.array([m.cost(x) for x in X])
    e16 = np.max(np.abs(g16 - cpu) / np.maximum(1.0, cpu))
    eb = np.max(np.abs(gb - cpu) / np.maximum(1.0, cpu))
    print('matrix form chi2 max rel err vs FP64: FP16 pairs %.2e, BF16 pairs %.2e' % (e16, eb))
    assert e16 <= TOL
    assert eb > e16
    assert np.array_equal(g16, g16b)
