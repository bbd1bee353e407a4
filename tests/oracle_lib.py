"""ctypes binding of the FP64 oracle (oracle/p6_oracle.c) -- test infrastructure only.

Builds oracle/_build/liboracle.so with make when it is missing or older than its sources.
"""
import ctypes
import os
import subprocess

import numpy as np

from powder6_amd.pt import CProblem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, 'oracle')
LIB = os.path.join(ORACLE_DIR, '_build', 'liboracle.so')
P_D = ctypes.POINTER(ctypes.c_double)
P_I = ctypes.POINTER(ctypes.c_int32)

_lib = None


def _stale():
    if not os.path.exists(LIB):
        return True
    t = os.path.getmtime(LIB)
    for f in ('p6_oracle.c', 'p6_oracle.h', 'mlp_oracle.c'):
        if os.path.getmtime(os.path.join(ORACLE_DIR, f)) > t:
            return True
    return False


def lib():
    global _lib
    if _lib is None:
        if _stale():
            subprocess.run(['make', '-s', '-C', ORACLE_DIR], check=True)
        L = ctypes.CDLL(LIB)
        for name in ('or_soft_positive', 'or_soft_floor', 'or_lp', 'or_polarization', 'or_tch_fwhm', 'or_hypot2'):
            getattr(L, name).restype = ctypes.c_double
            getattr(L, name).argtypes = [ctypes.c_double, ctypes.c_double]
        for name in ('or_erfc', 'or_lorentz'):
            getattr(L, name).restype = ctypes.c_double
            getattr(L, name).argtypes = [ctypes.c_double]
        L.or_peak_shift.restype = ctypes.c_double
        L.or_peak_shift.argtypes = [ctypes.c_double] * 3
        L.or_cell_volume.restype = ctypes.c_double
        L.or_cell_volume.argtypes = [P_D]
        L.or_inv_dsq.argtypes = [P_D, P_D]
        L.or_generate_reflections.argtypes = [P_D, ctypes.c_int, P_I, P_D, ctypes.c_double, P_I, P_I, P_D, ctypes.c_int]
        L.or_cosets.argtypes = [ctypes.c_int, P_I, P_D, P_I, ctypes.POINTER(ctypes.c_int)]
        L.or_reflection_dmin.restype = ctypes.c_double
        L.or_reflection_dmin.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.om_param_count.restype = ctypes.c_long
        L.om_param_count.argtypes = [ctypes.c_int, ctypes.c_int]
        L.om_forward.argtypes = [P_D, ctypes.c_int, ctypes.c_int, P_D, ctypes.c_int, P_D]
        L.om_loss_grad.restype = ctypes.c_double
        L.om_loss_grad.argtypes = [P_D, ctypes.c_int, ctypes.c_int, P_D, P_D, ctypes.c_void_p, ctypes.c_int, P_D]
        L.or_create.restype = ctypes.c_void_p
        L.or_create.argtypes = [ctypes.c_void_p]
        L.or_spline_slopes.argtypes = [P_D, P_D, ctypes.c_int, P_D, P_D]
        L.or_spline_values.argtypes = [P_D, ctypes.c_int, P_D, P_D, ctypes.c_int, P_D]
        L.or_destroy.argtypes = [ctypes.c_void_p]
        L.or_destroy.restype = None
        L.or_cost.restype = ctypes.c_double
        L.or_cost.argtypes = [ctypes.c_void_p, P_D]
        L.or_pattern.argtypes = [ctypes.c_void_p, P_D, P_D]
        L.or_violation.restype = ctypes.c_double
        L.or_violation.argtypes = [ctypes.c_void_p, P_D]
        L.or_reflection_table.argtypes = [ctypes.c_void_p, P_D, ctypes.c_int, P_I, P_D, P_D, P_D, ctypes.c_int]
        L.or_line_table.argtypes = [ctypes.c_void_p, P_D, ctypes.c_int, P_D, P_D, ctypes.c_int]
        L.or_normal.argtypes = [ctypes.c_void_p, P_D, P_D, P_D, P_D, P_D]
        L.or_temper.restype = ctypes.c_double
        L.or_temper.argtypes = [ctypes.c_void_p, P_D, P_D, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                ctypes.c_uint32, P_D, P_D]
        L.or_temper_parallel.restype = ctypes.c_double
        L.or_temper_parallel.argtypes = [ctypes.c_void_p, P_D, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_uint32, ctypes.c_int, P_D]
        L.or_prepare_voigt.argtypes = [ctypes.c_int, P_D, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
        L.or_stephens_basis.argtypes = [ctypes.c_int, P_I, P_I, P_D]
        L.or_axis_angle.restype = ctypes.c_double
        L.or_axis_angle.argtypes = [P_D, P_D, P_D]
        L.or_stephens_sum.restype = ctypes.c_double
        L.or_stephens_sum.argtypes = [P_D, P_D, ctypes.c_int, P_I, P_D]
        L.or_gsas_widths.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int, P_D, P_D]
        L.or_prepare_voigt_sep.argtypes = [ctypes.c_int, P_D, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
        L.or_eval_voigt.restype = ctypes.c_double
        L.or_eval_voigt.argtypes = [ctypes.c_double, ctypes.c_void_p]
        L.or_asym_area.restype = ctypes.c_double
        L.or_asym_area.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(P_D)


class OrVoigt(ctypes.Structure):
    _fields_ = [('type', ctypes.c_int), ('x0', ctypes.c_double), ('asym', ctypes.c_double), ('pedestal', ctypes.c_double),
                ('max_d', ctypes.c_double), ('window_fwhm', ctypes.c_double), ('fwhm', ctypes.c_double),
                ('eta', ctypes.c_double), ('inv_g', ctypes.c_double * 2), ('inv_l', ctypes.c_double * 2),
                ('cl', ctypes.c_double * 2), ('cg', ctypes.c_double * 2), ('analytic_area', ctypes.c_double),
                ('area', ctypes.c_double)]


def prepare_voigt(ptype, prof, min_fwhm, tth_peak, x0):
    p = np.zeros(32)
    p[:len(prof)] = prof
    v = OrVoigt()
    lib().or_prepare_voigt(ptype, p.ctypes.data_as(P_D), min_fwhm, tth_peak, x0, ctypes.byref(v))
    return v


def eval_voigt(x, v):
    return lib().or_eval_voigt(x, ctypes.byref(v))


def generate_reflections(cell, op_r, op_t, dmin, box=None):
    L = lib()
    cell_a, cell_p = d(cell)
    r = np.ascontiguousarray(np.asarray(op_r).reshape(-1), dtype=np.int32)
    t_a, t_p = d(np.asarray(op_t).reshape(-1))
    bx = None if box is None else np.ascontiguousarray(box, dtype=np.int32)
    bp = None if bx is None else bx.ctypes.data_as(P_I)
    n = L.or_generate_reflections(cell_p, len(op_r), r.ctypes.data_as(P_I), t_p, dmin, bp, None, None, 0)
    out = np.zeros((max(n, 1), 4), dtype=np.int32)
    d0 = np.zeros(max(n, 1))
    L.or_generate_reflections(cell_p, len(op_r), r.ctypes.data_as(P_I), t_p, dmin, bp, out.ctypes.data_as(P_I),
                              d0.ctypes.data_as(P_D), n)
    return out[:n], d0[:n]


class Model:
    """FP64 oracle model of a problems.Problem."""

    def __init__(self, pb):
        self.pb = pb
        self.cp = CProblem(pb)
        self.h = lib().or_create(self.cp.ref())
        if not self.h:
            raise RuntimeError('oracle rejected the problem')

    def close(self):
        if self.h:
            lib().or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cost(self, x):
        a, p = d(x)
        return lib().or_cost(self.h, p)

    def pattern(self, x):
        a, p = d(x)
        y = np.zeros(len(self.pb.tth))
        rc = lib().or_pattern(self.h, p, y.ctypes.data_as(P_D))
        if rc:
            y[:] = np.nan
        return y

    def violation(self, x):
        a, p = d(x)
        return lib().or_violation(self.h, p)

    def reflections(self, x, phase=0, cap=100000):
        a, p = d(x)
        hk = np.zeros((cap, 4), dtype=np.int32)
        f2, h, c = np.zeros(cap), np.zeros(cap), np.zeros(cap)
        n = lib().or_reflection_table(self.h, p, phase, hk.ctypes.data_as(P_I), f2.ctypes.data_as(P_D),
                                      h.ctypes.data_as(P_D), c.ctypes.data_as(P_D), cap)
        n = min(n, cap)
        return hk[:n], f2[:n], h[:n], c[:n]

    def normal(self, x):
        """normal() of one LM step at x: (A, g, chi2, J) with J[j, i] = d ycalc_i / d x_j."""
        a, p = d(x)
        nv, n = self.pb.nv, len(self.pb.tth)
        A, g, f, J = np.zeros((nv, nv)), np.zeros(nv), np.zeros(1), np.zeros((nv, n))
        rc = lib().or_normal(self.h, p, A.ctypes.data_as(P_D), g.ctypes.data_as(P_D), f.ctypes.data_as(P_D),
                             J.ctypes.data_as(P_D))
        if rc != 0:
            raise ValueError('oracle normal equations: status %d' % rc)
        return A, g, float(f[0]), J

    def lines(self, x, phase=0, cap=100000):
        """Per reflection: lines[k, pass] = (present, centre, weight, area) of the Kalpha1
        (pass 0) and Kalpha2 (pass 1) contributions, bragg[k, pass] the Bragg angles."""
        a, p = d(x)
        ln, br = np.zeros((cap, 2, 4)), np.zeros((cap, 2))
        n = lib().or_line_table(self.h, p, phase, ln.ctypes.data_as(P_D), br.ctypes.data_as(P_D), cap)
        n = min(n, cap)
        return ln[:n], br[:n]

    def temper(self, x0, scales, sweeps, chains, max_temp, seed):
        a, p = d(x0)
        s, sp = d(scales)
        best = np.zeros(self.pb.nv)
        st = np.zeros(4)
        c = lib().or_temper(self.h, p, sp, sweeps, chains, max_temp, seed, best.ctypes.data_as(P_D), st.ctypes.data_as(P_D))
        return c, best, st


def oracle_simulate(pb, x):
    m = Model(pb)
    try:
        return m.pattern(x)
    finally:
        m.close()


# ---- surrogate network (oracle/mlp_oracle.c) ----------------------------------------
def mlp_param_count(n_in, h):
    return int(lib().om_param_count(n_in, h))


def mlp_forward(params, n_in, h, x):
    p = np.ascontiguousarray(params, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, n_in)
    y = np.zeros(x.shape[0])
    lib().om_forward(p.ctypes.data_as(P_D), n_in, h, x.ctypes.data_as(P_D), x.shape[0], y.ctypes.data_as(P_D))
    return y


def mlp_loss_grad(params, n_in, h, x, t, mask=None):
    p = np.ascontiguousarray(params, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, n_in)
    t = np.ascontiguousarray(t, dtype=np.float64)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.float64)
    g = np.zeros(p.size)
    loss = lib().om_loss_grad(p.ctypes.data_as(P_D), n_in, h, x.ctypes.data_as(P_D), t.ctypes.data_as(P_D),
                              None if m is None else m.ctypes.data_as(P_D), x.shape[0], g.ctypes.data_as(P_D))
    return loss, g


class OrSep(ctypes.Structure):
    _fields_ = [('inst', ctypes.c_double * 18), ('sample', P_D), ('size_model', ctypes.c_int), ('strain_model', ctypes.c_int),
                ('size_axis', ctypes.c_double * 3), ('strain_axis', ctypes.c_double * 3), ('cell', ctypes.c_double * 6),
                ('h', ctypes.c_double * 3), ('d', ctypes.c_double), ('n_basis', ctypes.c_int), ('basis_key', P_I),
                ('basis_coeff', P_D)]


def stephens_basis(op_r):
    """The oracle's Stephens basis of a setting: [(key index, 15 coefficients)]."""
    op_r = np.ascontiguousarray(np.asarray(op_r).reshape(-1), dtype=np.int32)
    keys = np.zeros(15, dtype=np.int32)
    coeff = np.zeros(15 * 15)
    n = lib().or_stephens_basis(len(op_r) // 9, op_r.ctypes.data_as(P_I), keys.ctypes.data_as(P_I), coeff.ctypes.data_as(P_D))
    return [(int(keys[t]), coeff[15 * t:15 * t + 15].copy()) for t in range(n)]


def gsas_widths(ptype, cell, hkl, d, sample, inst=None, size_model=0, strain_model=0, size_axis=(0, 0, 1),
                strain_axis=(0, 0, 1), basis=None, lam=1.540598, min_fwhm=0.01, tth=40.0, side=1):
    """(Gaussian, Lorentzian) FWHM of a separated profile at one reflection."""
    sp = OrSep()
    inst = [2, -2, 5, 1, 0, 0, 0, 0, 0.5, 2, -2, 5, 1, 2, -2, 5, 1, 0.5] if inst is None else inst
    for i in range(18):
        sp.inst[i] = inst[i]
    sample = np.ascontiguousarray(sample, dtype=np.float64)
    sp.sample = sample.ctypes.data_as(P_D)
    sp.size_model, sp.strain_model = size_model, strain_model
    for i in range(3):
        sp.size_axis[i], sp.strain_axis[i], sp.h[i] = size_axis[i], strain_axis[i], hkl[i]
    for i in range(6):
        sp.cell[i] = cell[i]
    sp.d = d
    keys = np.zeros(15, dtype=np.int32)
    coeff = np.zeros(225)
    for t, (k, c) in enumerate(basis or []):
        keys[t], coeff[15 * t:15 * t + 15] = k, c
    sp.n_basis = len(basis or [])
    sp.basis_key, sp.basis_coeff = keys.ctypes.data_as(P_I), coeff.ctypes.data_as(P_D)
    g, l = ctypes.c_double(), ctypes.c_double()
    lib().or_gsas_widths(ptype, ctypes.byref(sp), lam, min_fwhm, tth, side, ctypes.byref(g), ctypes.byref(l))
    return g.value, l.value


def spline_values(tth, xs, ys):
    """The oracle's spline background at the points tth (anchors xs, values ys)."""
    tth = np.ascontiguousarray(tth, dtype=np.float64)
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    ys = np.ascontiguousarray(ys, dtype=np.float64)
    out = np.zeros(len(tth))
    pd = lambda a: a.ctypes.data_as(P_D)
    lib().or_spline_values(pd(tth), len(tth), pd(xs), pd(ys), len(xs), pd(out))
    return out
