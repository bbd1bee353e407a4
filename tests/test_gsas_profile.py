"""The separated (GSAS) profile of the oracle (js/profile.js:320-422 gsasStrainKeys,
gsasStephensBasis, gsasReciprocalVector, gsasCwComponents; js/rietveld_core.js:413-460,
694-708, 1413-1478), pinned by what the mathematics and the reference's own closed forms
fix: the generated Stephens basis against the hand-written quartic of each crystal system
(profile.js:371-374), size and strain models that reduce to the isotropic one, the
ellipsoid's rotation, the instrument part against the established profile path, a
separated pattern without sample broadening against the plain TCH pattern, and the
Stephens guard."""
import numpy as np
import pytest

import oracle_lib as O
from powder6_amd import problems as PR
from powder6_amd import symmetry as S

QUARTIC = PR.QUARTIC
WEIGHT = [1 if '4' in k else 2 if '3' in k else 3 if '0' in k else 4 for k in QUARTIC]


def _system_quartic(system, axis, h, S_):
    """profile.js:371-374: the hand-written quartic of each crystal system."""
    x, y, z = h
    x2, y2, z2 = x * x, y * y, z * z
    v = lambda k: S_.get(k, 0.0)
    if system == 'cubic':
        return v('S400') * (x2 * x2 + y2 * y2 + z2 * z2) + 3 * v('S220') * (x2 * y2 + x2 * z2 + y2 * z2)
    if system == 'tetragonal':
        return v('S400') * (x2 * x2 + y2 * y2) + v('S004') * z2 * z2 + 3 * v('S220') * x2 * y2 + 3 * v('S202') * (x2 + y2) * z2
    if system in ('hexagonal', 'trigonal'):
        xy = x2 + x * y + y2
        return v('S400') * xy * xy + v('S004') * z2 * z2 + 3 * v('S202') * xy * z2
    keys = ['S400', 'S040', 'S004', 'S220', 'S202', 'S022']
    if system == 'monoclinic':
        keys += {'a': ['S031', 'S013', 'S211'], 'c': ['S310', 'S130', 'S112']}.get(axis, ['S301', 'S103', 'S121'])
    elif system == 'triclinic':
        keys += ['S310', 'S301', 'S130', 'S103', 'S031', 'S013', 'S211', 'S121', 'S112']
    total = 0.0
    for key in keys:
        e = [int(c) for c in key[1:]]
        w = 1 if 4 in e else 2 if 3 in e else 3 if 0 in e else 4
        total += w * v(key) * x ** e[0] * y ** e[1] * z ** e[2]
    return total


def _closure(gens):
    """The rotation group generated by integer 3x3 matrices."""
    out = [np.eye(3, dtype=int)]
    grew = True
    while grew:
        grew = False
        for a in list(out):
            for g in gens:
                m = a @ np.array(g, dtype=int)
                if not any(np.array_equal(m, b) for b in out):
                    out.append(m)
                    grew = True
    return np.array(out).reshape(-1, 9)


# hexagonal axes: 6-fold about c, 2-fold about a, inversion (6/mmm); 3-fold, 2-fold,
# inversion (-3m1)
HEX_6MMM = _closure([[[1, -1, 0], [1, 0, 0], [0, 0, 1]], [[1, -1, 0], [0, -1, 0], [0, 0, -1]], -np.eye(3, dtype=int)])
HEX_3M = _closure([[[0, -1, 0], [1, -1, 0], [0, 0, 1]], [[1, -1, 0], [0, -1, 0], [0, 0, -1]], -np.eye(3, dtype=int)])


@pytest.mark.parametrize('group,axis', [(225, 'b'), (221, 'b'), (123, 'b'), (139, 'b'), ('6/mmm', 'b'), ('-3m1', 'b'),
                                        (62, 'b'), (19, 'b'), (14, 'b'), (2, 'b'), (1, 'b')])
def test_stephens_basis_reproduces_the_system_quartic(group, axis):
    if group == '6/mmm' or group == '-3m1':
        r, system = (HEX_6MMM, 'hexagonal') if group == '6/mmm' else (HEX_3M, 'trigonal')
    else:
        r, _ = S.group_ops(group)
        system = S.group_system(group)
    basis = O.stephens_basis(r)
    names = ['S' + QUARTIC[k] for k, _ in basis]
    rng = np.random.default_rng(abs(hash(str(group))) % 1000)
    S_ = {n: float(rng.uniform(-3, 3)) for n in names}
    if group == '-3m1':
        # -3m has a fourth invariant (S211) beyond the hexagonal quartic; with it at zero
        # the basis is the hand-written hexagonal form
        assert names == ['S400', 'S004', 'S202', 'S211']
        S_['S211'] = 0.0
    svec = np.zeros(15)
    for n, val in S_.items():
        svec[QUARTIC.index(n[1:])] = val
    keys = np.array([k for k, _ in basis], dtype=np.int32)
    coeff = np.concatenate([c for _, c in basis])
    for _ in range(12):
        h = rng.integers(-7, 8, size=3).astype(np.float64)
        got = O.lib().or_stephens_sum(O.d(h)[1], O.d(svec)[1], len(basis), keys.ctypes.data_as(O.P_I), O.d(coeff)[1])
        want = _system_quartic(system, axis, h, S_)
        assert abs(got - want) <= 1e-9 * max(1.0, abs(want)), (group, h, got, want)


@pytest.mark.parametrize('group', [225, 123, 139, 62, 14, 2, '6/mmm', '-3m1'])
def test_stephens_basis_is_invariant(group):
    """Every basis quartic takes the same value at h and at R^T h for every rotation R
    of the setting (the defining property, profile.js:334-363)."""
    r = HEX_6MMM if group == '6/mmm' else HEX_3M if group == '-3m1' else S.group_ops(group)[0]
    basis = O.stephens_basis(r)
    rots = np.asarray(r).reshape(-1, 3, 3)
    rng = np.random.default_rng(7)
    for k, c in basis:
        one = np.zeros(15)
        one[k] = 1.0
        keys = np.array([k], dtype=np.int32)
        for _ in range(4):
            h = rng.integers(-6, 7, size=3).astype(np.float64)
            v0 = O.lib().or_stephens_sum(O.d(h)[1], O.d(one)[1], 1, keys.ctypes.data_as(O.P_I), O.d(c)[1])
            for R in rots:
                hr = (R.T @ h).astype(np.float64)
                v1 = O.lib().or_stephens_sum(O.d(hr)[1], O.d(one)[1], 1, keys.ctypes.data_as(O.P_I), O.d(c)[1])
                assert abs(v1 - v0) <= 1e-9 * max(1.0, abs(v0)), (group, QUARTIC[k], h, R)


def test_stephens_basis_sizes():
    """Independent terms per Laue class: 2 cubic, 4 tetragonal 4/mmm, 3 hexagonal,
    6 orthorhombic, 9 monoclinic, 15 triclinic."""
    for group, n in ((225, 2), (123, 4), (62, 6), (14, 9), (2, 15)):
        assert len(O.stephens_basis(S.group_ops(group)[0])) == n, group


ORTHO = [8.48, 5.40, 6.96, 90, 90, 90]
SAMPLE = np.full(26, np.nan)


def _sample(**kw):
    s = SAMPLE.copy()
    for k, v in kw.items():
        s[PR.SAMPLE_SLOTS.index(k)] = v
    return s


def _d(cell, h):
    form = np.zeros(6)
    O.lib().or_inv_dsq(O.d(cell)[1], form.ctypes.data_as(O.P_D))
    q = h[0] ** 2 * form[0] + h[1] ** 2 * form[1] + h[2] ** 2 * form[2] + h[0] * h[1] * form[3] + h[1] * h[2] * form[4] + h[0] * h[2] * form[5]
    return 1 / np.sqrt(q)


MONO = [7.1, 9.3, 6.2, 90, 104.0, 90]


@pytest.mark.parametrize('h', [(1, 0, 0), (2, 1, 3), (0, 3, 1), (4, 0, 2)])
def test_uniaxial_and_ellipsoid_reduce_to_isotropic_size(h):
    d = _d(MONO, h)
    iso = O.gsas_widths(3, MONO, h, d, _sample(size=0.1, strain=600, sizeMix=0.7, strainMix=0.4), tth=55.0)
    uni = O.gsas_widths(3, MONO, h, d, _sample(size=0.1, sizeAxial=0.1, strain=600, sizeMix=0.7, strainMix=0.4),
                        size_model=1, size_axis=(1, 1, 0), tth=55.0)
    ell = O.gsas_widths(3, MONO, h, d, _sample(size=0.1, sizeB=0.1, sizeC=0.1, rotA=17, rotB=-33, rotC=71, strain=600,
                                               sizeMix=0.7, strainMix=0.4), size_model=2, tth=55.0)
    assert np.allclose(uni, iso, rtol=1e-12) and np.allclose(ell, iso, rtol=1e-12)


def test_uniaxial_size_along_and_across_its_axis():
    """Along the axis 1/size is 1/sizeAxial, across it 1/size (profile.js:393-396)."""
    along = O.gsas_widths(3, ORTHO, (0, 0, 2), _d(ORTHO, (0, 0, 2)), _sample(size=0.2, sizeAxial=0.05), size_model=1, tth=30.0)
    ref_a = O.gsas_widths(3, ORTHO, (0, 0, 2), _d(ORTHO, (0, 0, 2)), _sample(size=0.05), tth=30.0)
    across = O.gsas_widths(3, ORTHO, (2, 0, 0), _d(ORTHO, (2, 0, 0)), _sample(size=0.2, sizeAxial=0.05), size_model=1, tth=30.0)
    ref_c = O.gsas_widths(3, ORTHO, (2, 0, 0), _d(ORTHO, (2, 0, 0)), _sample(size=0.2), tth=30.0)
    assert np.allclose(along, ref_a, rtol=1e-12) and np.allclose(across, ref_c, rtol=1e-12)
    assert along[1] > across[1]          # the thin direction is the broad one


def test_uniaxial_strain_along_and_across_its_axis():
    along = O.gsas_widths(3, ORTHO, (3, 0, 0), _d(ORTHO, (3, 0, 0)), _sample(strain=500, strainAxial=2500), strain_model=1,
                          strain_axis=(1, 0, 0), tth=70.0)
    ref_a = O.gsas_widths(3, ORTHO, (3, 0, 0), _d(ORTHO, (3, 0, 0)), _sample(strain=2500), tth=70.0)
    across = O.gsas_widths(3, ORTHO, (0, 2, 0), _d(ORTHO, (0, 2, 0)), _sample(strain=500, strainAxial=2500), strain_model=1,
                           strain_axis=(1, 0, 0), tth=70.0)
    ref_c = O.gsas_widths(3, ORTHO, (0, 2, 0), _d(ORTHO, (0, 2, 0)), _sample(strain=500), tth=70.0)
    assert np.allclose(along, ref_a, rtol=1e-12) and np.allclose(across, ref_c, rtol=1e-12)


def test_ellipsoid_rotation_about_c_swaps_a_and_b():
    """Rotating the ellipsoid by 90 deg about c puts its b length along a* (orthogonal cell)."""
    h = (2, 0, 0)
    d = _d(ORTHO, h)
    rot = O.gsas_widths(3, ORTHO, h, d, _sample(size=0.2, sizeB=0.05, sizeC=0.1, rotC=90), size_model=2, tth=40.0)
    swp = O.gsas_widths(3, ORTHO, h, d, _sample(size=0.05, sizeB=0.2, sizeC=0.1), size_model=2, tth=40.0)
    assert np.allclose(rot, swp, rtol=1e-9)


def test_isotropic_stephens_tensor_is_isotropic_strain():
    """Cubic S220 = 2 S400 / 3 makes the quartic S400 (h^2 + k^2 + l^2)^2, so the
    strain d^2 sqrt(sum) is a^2 sqrt(S400) for every reflection (profile.js:410-412)."""
    a = 5.64
    cell = [a, a, a, 90, 90, 90]
    basis = O.stephens_basis(S.group_ops(225)[0])
    s400 = 40.0
    for h in ((1, 1, 1), (2, 0, 0), (3, 1, 1), (4, 2, 0)):
        st = O.gsas_widths(3, cell, h, _d(cell, h), _sample(S400=s400, S220=2 * s400 / 3), strain_model=2, basis=basis, tth=50.0)
        iso = O.gsas_widths(3, cell, h, _d(cell, h), _sample(strain=a * a * np.sqrt(s400)), tth=50.0)
        assert np.allclose(st, iso, rtol=1e-12), h


def test_instrument_part_matches_the_simple_profile():
    """Without sample broadening (size 1e9, strain 0) and with Y = Z = 0 the separated
    simple-pseudo-Voigt widths are the phase-owned simple profile's (GU = U, GV = V,
    GW = W, LX = X)."""
    inst = [9.0, -3.0, 4.0, 2.5, 0, 0, 0, 0, 0.5, 2, -2, 5, 1, 2, -2, 5, 1, 0.5]
    for tth in (12.0, 47.0, 96.0, 140.0):
        g, l = O.gsas_widths(1, ORTHO, (1, 1, 1), 3.0, _sample(size=1e9, strain=0.0), inst=inst, tth=tth, min_fwhm=0.01)
        v = O.prepare_voigt(1, [9.0, -3.0, 4.0, 0.0, 2.5, 0.5], 0.01, tth, tth)
        assert np.isclose(g, 1 / v.inv_g[0], rtol=1e-12) and np.isclose(l, 1 / v.inv_l[0], rtol=1e-12), tth


def test_separated_pattern_without_sample_broadening_is_the_tch_pattern():
    """A gsas_cw phase with no sample broadening renders exactly the pattern of a
    tch_aniso phase carrying the instrument's U, V, W, X, Y, SL, HL (Z = 0)."""
    ref = PR.separated_probe('isotropic', 'isotropic', 'gsas_cw')
    ph = ref.phases[0]
    sph = PR.Phase('p1', 62, ph.cell, ph.atoms, 'gsas_cw')
    tph = PR.Phase('p1', 62, ph.cell, ph.atoms, 'tch_aniso')
    sep, tch = [PR.Problem(ref.tth, [q], lambda1=ref.lambda1, lambda2=ref.lambda2, ratio=ref.ratio, pol_k=1.0, n_bkg=3)
                for q in (sph, tph)]
    for pb in (sep, tch):
        pb.default_values()
        pb.y = np.ones_like(pb.tth)
        pb.set('zero', 0.01)
    inst = dict(PR.SEPARATED_INST, Z=0.0)
    for n, v in inst.items():
        sep.set('inst.' + n, v)
    sep.set('p1.size', 1e12)
    sep.set('p1.strain', 0.0)
    for n in ('U', 'V', 'W', 'X', 'Y', 'SL', 'HL'):
        tch.set('p1.' + n, inst[n])
    sep.refine('p1.scale', 0, np.inf)
    tch.refine('p1.scale', 0, np.inf)
    ys = O.Model(sep).pattern(sep.start_vector())
    yt = O.Model(tch).pattern(tch.start_vector())
    assert np.max(np.abs(ys - yt)) <= 1e-9 * np.max(np.abs(yt))


def test_stephens_guard():
    """A strain quartic negative in some direction is a violation (rietveld_core.js:1469-1477);
    cubic S400 = 1, S220 = -1 is -0.25 at [110]."""
    pb = PR.separated_probe('isotropic', 'stephens', 'gsas_cw')
    pb.y = np.ones_like(pb.tth)
    pb.refine('p1.S220')
    m = O.Model(pb)
    x = pb.start_vector()
    assert m.violation(x) == 0.0
    x[pb.var_index('p1.S220')] = -400.0
    assert m.violation(x) > 0.0
