"""Pins the FP64 oracle against closed forms and the reference's own known-answer tests
(tests/numerics.test.cjs, tests/regressions.test.cjs), restated, and against brute force.

Each check is chosen so that a plausible mistake in the oracle (a dropped term, a wrong
index or sign, a transposed operand) fails it."""
import ctypes
import math

import numpy as np
import pytest

import oracle_lib as O
from powder6_amd import problems as PR
from powder6_amd import symmetry as S

L = O.lib()


def near(a, b, tol):
    assert abs(a - b) <= tol, (a, b)


def cell_arr(*c):
    return (ctypes.c_double * 6)(*c)


# ---- cell geometry: numerics.test.cjs:12-35 ------------------------------------------
def test_cell_volumes():
    near(L.or_cell_volume(cell_arr(5, 5, 5, 90, 90, 90)), 125, 1e-9)
    near(L.or_cell_volume(cell_arr(4, 4, 6, 90, 90, 120)), 16 * 6 * math.sin(math.pi * 2 / 3), 1e-9)
    near(L.or_cell_volume(cell_arr(5, 6, 7, 90, 100, 90)), 210 * math.sin(100 * math.pi / 180), 1e-9)
    near(L.or_cell_volume(cell_arr(5, 5, 5, 60, 60, 60)), 125 / math.sqrt(2), 1e-9)


def _dspacing(cell, h, k, l):
    f = (ctypes.c_double * 6)()
    assert L.or_inv_dsq(cell_arr(*cell), f)
    q = h * h * f[0] + k * k * f[1] + l * l * f[2] + h * k * f[3] + k * l * f[4] + h * l * f[5]
    return 1 / math.sqrt(q)


def test_d_spacings():
    near(_dspacing((5, 5, 5, 90, 90, 90), 1, 1, 1), 5 / math.sqrt(3), 1e-12)
    near(_dspacing((5, 5, 5, 90, 90, 90), 2, 0, 0), 2.5, 1e-12)
    near(_dspacing((5, 6, 7, 90, 90, 90), 1, 2, 3), 1 / math.sqrt(1 / 25 + 4 / 36 + 9 / 49), 1e-12)
    near(_dspacing((4, 4, 6, 90, 90, 120), 1, 0, 0), 4 * math.sqrt(3) / 2, 1e-12)
    near(_dspacing((4, 4, 6, 90, 90, 120), 0, 0, 1), 6, 1e-12)
    near(_dspacing((4, 4, 6, 90, 90, 120), 1, 1, 0), 2, 1e-12)
    sb, cb = math.sin(math.radians(100)), math.cos(math.radians(100))
    q = (1 / 25 + 4 / 49 - 2 * 2 * cb / 35) / (sb * sb) + 9 / 36
    near(_dspacing((5, 6, 7, 90, 100, 90), 1, 3, 2), 1 / math.sqrt(q), 1e-12)
    # general triclinic against the inverse of the real-space metric tensor
    cell = (3.7, 4.1, 4.3, 87, 94, 102)
    Ginv = np.linalg.inv(S.metric_tensor(cell))
    for hkl in [(1, 0, 0), (1, -1, 2), (2, 3, -1)]:
        h = np.array(hkl, dtype=float)
        near(_dspacing(cell, *hkl), 1 / math.sqrt(h @ Ginv @ h), 1e-12)


def test_degenerate_cell_rejected():
    f = (ctypes.c_double * 6)()
    assert not L.or_inv_dsq(cell_arr(5, 5, 5, 60, 60, 120), f)   # flat: volume factor 0
    assert math.isnan(L.or_cell_volume(cell_arr(5, -1, 5, 90, 90, 90)))


# ---- Lp: numerics.test.cjs:41-52 ---------------------------------------------------
def test_lorentz_polarisation():
    near(L.or_lorentz(90), 1 / (0.5 * math.sqrt(0.5)), 1e-12)
    assert math.isnan(L.or_lorentz(0)) and math.isnan(L.or_lorentz(180))
    near(L.or_polarization(90, 1), 0.5, 1e-12)
    near(L.or_polarization(0, 1), 1, 1e-12)
    near(L.or_polarization(60, 0), 1, 1e-12)
    near(L.or_polarization(60, 1), (1 + 0.25) / 2, 1e-12)
    near(L.or_lp(90, 1), 0.5 / (0.5 * math.sqrt(0.5)), 1e-12)
    assert L.or_lp(0, 1) == 1


# ---- erfc: numerics.test.cjs:91-96 ---------------------------------------------------
def test_erfc():
    near(L.or_erfc(0), 1, 1.2e-7)
    near(L.or_erfc(1), 0.15729920705028513, 1.2e-7)
    for x in (0.01, 0.5, 2, 5):
        near(L.or_erfc(-x), 2 - L.or_erfc(x), 1e-12)
        near(L.or_erfc(x), math.erfc(x), 1.2e-7 * max(1.0, math.erfc(x)))


# ---- softPositive: regressions.test.cjs:191-201 --------------------------------------
def test_soft_positive():
    assert abs(L.or_soft_positive(-1e10, 1e-3) / 1e-16 - 1) < 1e-14
    assert L.or_soft_positive(0, 0.25) == 0.25
    assert L.or_soft_positive(1e308, 0.25) == 1e308
    assert L.or_soft_positive(-1e308, 0.25) > 0
    for x in (-100, -1, 1, 100):
        assert abs(L.or_soft_positive(x, 0.5) - (x + math.sqrt(x * x + 1)) / 2) < 1e-12


def test_soft_floor_and_hypot():
    near(L.or_soft_floor(0.5, 1.0), (0.5 ** 4 + 1) ** 0.25, 1e-15)
    assert L.or_soft_floor(30.0, 1.0) == 30.0
    assert L.or_soft_floor(-1.0, 0.2) == 0.2
    near(L.or_hypot2(3.0, 4.0), 5.0, 0)
    near(L.or_hypot2(1e200, 1e200), math.sqrt(2) * 1e200, 1e186)


# ---- TCH FWHM and pseudo-Voigt shape -------------------------------------------------
def test_tch_fwhm_limits():
    # pure components and the published polynomial at G = L
    # (components are floored at 1e-9, so the pure limits carry a 1e-9 partner)
    g, l = 0.1, 1e-9
    exact = (g ** 5 + 2.69269 * g ** 4 * l + 2.42843 * g ** 3 * l ** 2 + 4.47163 * g ** 2 * l ** 3 + 0.07842 * g * l ** 4 + l ** 5) ** 0.2
    near(L.or_tch_fwhm(0.1, 1e-12), exact, 1e-15)
    near(L.or_tch_fwhm(0.1, 1e-12), 0.1, 1e-9)
    near(L.or_tch_fwhm(1e-12, 0.1), 0.1, 1e-9)
    s = 1 + 2.69269 + 2.42843 + 4.47163 + 0.07842 + 1
    near(L.or_tch_fwhm(0.1, 0.1), 0.1 * s ** 0.2, 1e-14)


@pytest.mark.parametrize('ptype,prof', [(3, [18.03, -9.02, 5.41, 5.0, 2.0, 0, 0]), (1, [2, -2, 5, 0, 1, 0.3]),
                                        (2, [2, -2, 5, 1, 8, -3, 6, 3, 0.6])])
def test_voigt_unit_height_and_pedestal(ptype, prof):
    v = O.prepare_voigt(ptype, prof, 0.04, 40.0, 40.0)
    # unit height at the centre before the pedestal, zero at both window edges
    near(O.eval_voigt(40.0, v) + v.pedestal, 1.0, 1e-12)
    assert O.eval_voigt(40.0 + v.max_d * (1 - 1e-12), v) < 1e-12
    assert O.eval_voigt(40.0 - v.max_d * (1 - 1e-12), v) < 1e-12
    assert O.eval_voigt(40.0 + v.max_d * 1.0001, v) == 0.0
    # analytic area of the untruncated shape vs a brute-force integral
    lo, hi, n = -40.0, 40.0, 400001
    xs = np.linspace(lo, hi, n)
    w = v.window_fwhm
    fv = O.OrVoigt()
    ctypes.pointer(fv)[0] = v
    fv.x0, fv.max_d, fv.pedestal = 0.0, float('inf'), 0.0
    vals = np.array([O.eval_voigt(x, fv) for x in xs[::4]])
    integral = np.trapezoid(vals, xs[::4])
    # Lorentzian tail beyond 40 deg on each flank: cL (H/2) (pi/2 - atan(2 X / H))
    tail = sum(fv.cl[s] * (0.5 / fv.inv_l[s]) * (math.pi / 2 - math.atan(2 * 40 * fv.inv_l[s])) for s in range(2))
    near(integral + tail, v.analytic_area, 2e-4 * v.analytic_area)
    assert w > 0


def test_tch_eta_polynomial():
    """eta from the TCH polynomial (Thompson, Cox & Hastings 1987) at a known ratio."""
    v = O.prepare_voigt(3, [0, 0, 1e-9, 10.0, 0, 0, 0], 1e-6, 30.0, 30.0)
    # pure Lorentzian input: gG at the floor, so r -> 1 and eta = 1.36603 - 0.47719 + 0.11116
    near(v.eta, 1.36603 - 0.47719 + 0.11116, 2e-3)


def test_asymmetry_area_brute_force():
    """regressions.test.cjs:1583-1591: the area normaliser includes the SL/HL asymmetry."""
    prof = [0, 0, 5, 5, 0, 0.05, 0.02]
    v = O.prepare_voigt(3, prof, 0.005, 10.0, 10.0)
    asym = O.OrVoigt()
    ctypes.pointer(asym)[0] = v
    asym.x0, asym.max_d, asym.pedestal = 0.0, float('inf'), 0.0
    sym = O.OrVoigt()
    ctypes.pointer(sym)[0] = asym
    sym.asym = 0.0
    Lw, n, diff = 3000 * v.fwhm, 200000, 0.0
    u = -1 + (np.arange(n) + 0.5) * 2 / n
    xs = Lw * u ** 3
    dx = 3 * Lw * u * u * 2 / n
    for x, w in zip(xs, dx):
        diff += (O.eval_voigt(x, asym) - O.eval_voigt(x, sym)) * w
    near(v.area, v.analytic_area + diff, 2e-4 * (v.analytic_area + diff))
    assert v.area > v.analytic_area * 1.01


def _single_line_problem(lam2, ratio):
    """Pm-3m a = 3 A, one Fe atom at the origin, 2theta 20-38 deg: the (100) line alone."""
    r, t = S.group_ops(221)
    cell = [3.0, 3.0, 3.0, 90, 90, 90]
    at = PR.Atom('Fe1', 'Fe', S.site_symmetry(r, t, [0, 0, 0], cell))
    ph = PR.Phase('p1', 221, cell, [at], 'tch_aniso')
    pb = PR.Problem(20 + 0.01 * np.arange(1801), [ph], lambda1=PR.CU_KA1, lambda2=lam2, ratio=ratio,
                    pol_k=float('nan'), n_bkg=0)
    pb.default_values()
    for n, v in zip(('U', 'V', 'W', 'X', 'Y'), (18.03, -9.02, 5.41, 5.0, 2.0)):
        pb.set('p1.' + n, v)
    pb.refine('p1.scale', 0, math.inf)
    pb.y = np.ones_like(pb.tth)
    return pb


def test_doublet_area_ratio():
    """numerics.test.cjs:186-200: the Ka2 line carries ratio x area1/area2 of the Ka1 line,
    so with Lp left out (polK not finite) the integrated single-reflection pattern with
    the doublet is (1 + ratio) times the Ka1-only one."""
    pa, pb = _single_line_problem(0.0, 0.0), _single_line_problem(1.544426, 0.5)
    ya = O.Model(pa).pattern(pa.start_vector())
    yb = O.Model(pb).pattern(pb.start_vector())
    near(yb.sum() / ya.sum(), 1.5, 2e-5)
    # the Ka2 line sits at 2 asin(sin(theta1) lambda2/lambda1)
    i1, i2 = int(np.argmax(ya)), int(np.argmax(yb - ya))
    th1 = math.radians(pa.tth[i1] / 2)
    t2 = 2 * math.degrees(math.asin(math.sin(th1) * 1.544426 / PR.CU_KA1))
    near(pb.tth[i2], t2, 0.011)


# ---- symmetry and reflections -------------------------------------------------------
def _allowed_fm3m(h, k, l):
    return (h % 2 == k % 2 == l % 2)


def _allowed_pnma(h, k, l):
    """ITA reflection conditions of Pnma (standard setting)."""
    if h == 0 and k != 0 and l != 0 and (k + l) % 2:
        return False
    if l == 0 and h != 0 and h % 2:
        return False
    if k == 0 and l == 0 and h % 2:
        return False
    if h == 0 and l == 0 and k % 2:
        return False
    if h == 0 and k == 0 and l % 2:
        return False
    if h == 0 and l == 0 and k == 0:
        return False
    if l == 0 and h == 0 and k % 2:
        return False
    return True


def _allowed_p21c(h, k, l):
    if k == 0 and l % 2:
        return False
    if h == 0 and l == 0 and k % 2:
        return False
    return True


def _orbit_size(op_r, h):
    seen = set()
    for R in op_r:
        R = np.array(R).reshape(3, 3)
        v = tuple(int(x) for x in np.array(h) @ R)
        seen.add(v)
        seen.add(tuple(-x for x in v))
    return len(seen)


def _allowed_p21(h, k, l):          # P2_1 (unique b): 0k0 k = 2n
    return not (h == 0 and l == 0 and k % 2)


def _allowed_p212121(h, k, l):      # h00 h = 2n, 0k0 k = 2n, 00l l = 2n
    if k == 0 and l == 0 and h % 2:
        return False
    if h == 0 and l == 0 and k % 2:
        return False
    if h == 0 and k == 0 and l % 2:
        return False
    return True


@pytest.mark.parametrize('group,cell,rule', [
    (225, (5.64, 5.64, 5.64, 90, 90, 90), _allowed_fm3m),
    (216, (5.41, 5.41, 5.41, 90, 90, 90), _allowed_fm3m),
    (62, (8.48, 5.40, 6.96, 90, 90, 90), _allowed_pnma),
    (14, (10.4, 18.8, 6.8, 90, 106, 90), _allowed_p21c),
    (4, (5.2, 6.3, 7.1, 90, 101, 90), _allowed_p21),
    (19, (5.2, 6.3, 7.1, 90, 90, 90), _allowed_p212121),
])
def test_reflection_list_brute_force(group, cell, rule):
    """Every canonical hkl is allowed by the ITA conditions, the multiplicities are Laue
    orbit sizes, and sum_m m + absences equals the lattice points of the sphere."""
    op_r, op_t = S.group_ops(group)
    dmin = 1.5406 / (2 * math.sin(math.radians(75)))
    rows, d0 = O.generate_reflections(cell, op_r, op_t, dmin)
    Ginv = np.linalg.inv(S.metric_tensor(cell))
    lim = 1 / dmin ** 2 * (1 + 1e-9)
    pts = allowed = 0
    hmax = [int(math.floor(c / dmin)) for c in cell[:3]]
    for h in range(-hmax[0], hmax[0] + 1):
        for k in range(-hmax[1], hmax[1] + 1):
            for l in range(-hmax[2], hmax[2] + 1):
                if h == k == l == 0:
                    continue
                v = np.array([h, k, l], float)
                if v @ Ginv @ v <= lim:
                    pts += 1
                    allowed += rule(h, k, l)
    assert sum(int(r[3]) for r in rows) == allowed
    for (h, k, l, m) in rows:
        assert rule(h, k, l), (h, k, l)
        assert m == _orbit_size(op_r, (h, k, l))
    assert np.all(np.diff(d0) <= 0)          # sorted by decreasing d
    assert pts >= allowed


def test_nacl_structure_factor_closed_form():
    """Fm-3m NaCl: F = 4 (f_Na +- f_Cl) DW; all-even hkl add, all-odd subtract."""
    pb = PR.nacl_fm3m()
    pb.y = np.ones_like(pb.tth)
    m = O.Model(pb)
    x = pb.start_vector()
    hk, f2, h, cen = m.reflections(x)
    p = pb.values(x)
    a = p[pb.index['p1.a']]
    na, cl = pb.phases[0].atoms
    Bn, Bc = p[pb.index['p1.Na1.B']], p[pb.index['p1.Cl1.B']]
    for (hh, kk, ll, mm), F2 in zip(hk, f2):
        s2 = (hh * hh + kk * kk + ll * ll) / a ** 2
        q = s2 / 4
        fna = na.ff_c + sum(ai * math.exp(-bi * q) for ai, bi in zip(na.ff_a, na.ff_b))
        fcl = cl.ff_c + sum(ai * math.exp(-bi * q) for ai, bi in zip(cl.ff_a, cl.ff_b))
        sign = 1 if hh % 2 == 0 else -1
        F = 4 * (fna * math.exp(-Bn * q) + sign * fcl * math.exp(-Bc * q))
        near(F2, F * F, 1e-9 * max(1, F * F))


@pytest.mark.parametrize('anomalous', [False, True])
def test_zincblende_structure_factor_closed_form(anomalous):
    """F-43m ZnS (no inversion centre): Zn at 0, S at 1/4 1/4 1/4, unmixed hkl, phi =
    pi (h + k + l) / 2.  The reference's |F|^2 = A^2 + B^2 + C^2 + D^2 (f'' in C, D,
    Friedel-averaged) is 16 [(f_Zn + f_S cos phi)^2 + (f_S sin phi)^2 + f''_Zn^2], with
    f_Zn = (f0 + f') DW and f''_Zn scaled by the same DW."""
    pb = PR.zincblende_f43m(anomalous=anomalous)
    pb.y = np.ones_like(pb.tth)
    m = O.Model(pb)
    x = pb.start_vector()
    hk, f2, h, cen = m.reflections(x)
    p = pb.values(x)
    a = p[pb.index['p1.a']]
    zn, su = pb.phases[0].atoms
    Bz, Bs = p[pb.index['p1.Zn1.B']], p[pb.index['p1.S1.B']]
    assert len(hk) > 5
    for (hh, kk, ll, mm), F2 in zip(hk, f2):
        q = (hh * hh + kk * kk + ll * ll) / a ** 2 / 4
        f0z = zn.ff_c + sum(ai * math.exp(-bi * q) for ai, bi in zip(zn.ff_a, zn.ff_b))
        f0s = su.ff_c + sum(ai * math.exp(-bi * q) for ai, bi in zip(su.ff_a, su.ff_b))
        dz, ds = math.exp(-Bz * q), math.exp(-Bs * q)
        fz, fs, fzz = (f0z + zn.fp) * dz, (f0s + su.fp) * ds, zn.fpp * dz
        phi = math.pi * (hh + kk + ll) / 2
        F2ref = 16 * ((fz + fs * math.cos(phi)) ** 2 + (fs * math.sin(phi)) ** 2 + fzz ** 2)
        near(F2, F2ref, 1e-9 * max(1, F2ref))


def test_cosets_centring():
    op_r, op_t = S.group_ops(225)
    idx = (ctypes.c_int32 * len(op_r))()
    nc = ctypes.c_int()
    r = np.ascontiguousarray(op_r.reshape(-1), dtype=np.int32)
    t = np.ascontiguousarray(op_t.reshape(-1))
    n = L.or_cosets(len(op_r), r.ctypes.data_as(O.P_I), t.ctypes.data_as(O.P_D), idx, ctypes.byref(nc))
    assert nc.value == 4 and n == 48


def test_envelope_contains_trial_lists():
    """regressions.test.cjs:1035-1042: a box envelope holds the lists of expanded and
    oblique trial cells."""
    op_r, op_t = S.group_ops(2)
    env, _ = O.generate_reflections((3, 4, 5, 90, 90, 90), op_r, op_t, 1.0, box=(6, 8, 10))
    keys = {tuple(r[:3]) for r in env}
    for cell in [(6, 8, 10, 90, 90, 90), (5, 7, 9, 72, 108, 115)]:
        rows, _ = O.generate_reflections(cell, op_r, op_t, 1.0)
        assert len(rows) > 0
        for r in rows:
            assert tuple(r[:3]) in keys


# ---- the cost and the pattern ---------------------------------------------------------
def test_cost_is_weighted_residual_sum():
    pb = PR.triclinic_probe('tch_aniso', simulate=O.oracle_simulate)
    m = O.Model(pb)
    x = pb.start_vector()
    y = m.pattern(x)
    chi = np.sum(np.where(pb.w > 0, pb.w * (pb.y - y) ** 2, 0.0))
    near(m.cost(x), chi, 1e-9 * chi)
    # the synthetic observation is the model plus 2 sin(.13 i) + 1 (webgpu-pt.js:33)
    i = np.arange(len(pb.tth))
    r = pb.y - y
    assert np.allclose(r, 2 * np.sin(i * .13) + 1, atol=1e-9)


def test_pattern_background_only():
    """With the scale at zero the pattern is the Chebyshev series b0 T0 + b1 T1 + ..."""
    pb = PR.nacl_fm3m()
    pb.y = np.ones_like(pb.tth)
    m = O.Model(pb)
    x = pb.start_vector()
    x[pb.var_index('p1.scale')] = 0.0
    y = m.pattern(x)
    near(float(np.max(np.abs(y - pb.background(x)))), 0.0, 1e-9)


def test_invalid_cell_cost_is_infinite():
    pb = PR.nacl_fm3m()
    pb.y = np.ones_like(pb.tth)
    m = O.Model(pb)
    x = pb.start_vector()
    x[pb.var_index('p1.a')] = -1.0
    assert math.isinf(m.cost(x))


def test_variance_guard_vertex():
    pb = PR.nacl_fm3m()
    pb.y = np.ones_like(pb.tth)
    m = O.Model(pb)
    x = pb.start_vector()
    assert m.violation(x) == 0
    # U tan^2 + V tan + W with its vertex inside the range and negative there
    x[pb.var_index('p1.U')], x[pb.var_index('p1.V')], x[pb.var_index('p1.W')] = 100.0, -60.0, 5.0
    lo, hi = math.tan(math.radians(5)), math.tan(math.radians(60))
    t = 0.3
    assert lo < t < hi
    near(m.violation(x), -(100 * t * t - 60 * t + 5), 1e-12)


def test_temper_invariants():
    """webgpu-pt.js:134-135 on the CPU PT: proposals = chains*sweeps*nv (all finite),
    swaps recorded, best <= initial, best re-scores to itself."""
    pb = PR.triclinic_probe('tch_aniso', simulate=O.oracle_simulate)
    m = O.Model(pb)
    x0 = pb.start_vector()
    init = m.cost(x0)
    best_c, best, st = m.temper(x0, pb.scale, 10, 4, 5.0, 137)
    assert st[1] == 4 * 10 * pb.nv
    assert st[3] == sum(len(range(s % 2, 3, 2)) for s in range(10))
    assert best_c <= init
    near(m.cost(best), best_c, 1e-12 * best_c)
    # seeded: the same run twice gives the same result
    c2, b2, s2 = m.temper(x0, pb.scale, 10, 4, 5.0, 137)
    assert c2 == best_c and np.array_equal(b2, best)
