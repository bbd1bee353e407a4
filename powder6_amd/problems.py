"""Problem descriptions for the PT trial evaluator, and the seeded synthetic workloads.

A `Problem` is the flat form of the PT context js/rietveld_core.js:2271-2273 hands to
the GPU backend: one dataset (2-theta axis, counts, weights, wavelengths, polarisation,
width floor, Chebyshev or spline background), its Rietveld phases (operators, atoms on Wyckoff
sites, profile family), the model parameters as an affine function of the refined
variables, the variables' boxes and PT step scales, and the PT settings.  pt.py turns it
into the C struct of include/powder6_pt.h.

Workloads (named by what they are):
  nacl_fm3m()        NaCl Fm-3m a=5.64, Cu Ka1 only, 10-120 deg step 0.02, TCH, 6-term
                     Chebyshev background, 16 replicas, 12 refined parameters
  nacl_spline()      the same pattern over a monotone-spline background (8 refined anchors)
  pbso4_pnma()       PbSO4-type Pnma a=8.48 b=5.40 c=6.96, 5 sites (24 atoms), 10-150 deg
                     step 0.02 with the Ka1/Ka2 doublet, 256 replicas x 64 trials/sweep
  uneven_batch()     48 independent cubic / tetragonal / monoclinic patterns
  molecular_p21c()   P2_1/c a=10.4 b=18.8 c=6.8 beta=106, 17 C/N/S atoms, 512 replicas
  triclinic_probe()  the P-1 one-carbon problem of the reference's WebGPU test page
                     (tests/webgpu-pt.js:11-36), for every profile family

The observed pattern is the model at seeded true parameters plus seeded Poisson noise,
w = 1/max(y, 1) (js/rietveld/rv_state.js:1122-1129); the width floor is two median
steps (rv_state.js:1167-1172).  `simulate(problem, x) -> ycalc` computes the model; the
caller passes the GPU pattern function of the library (or any other evaluator).
"""

import math

import numpy as np

from . import symmetry as S
from .scatter_wk1995 import WK1995

# model-parameter layout of include/powder6_pt.h
P_ZERO, P_DISP, P_TRANSP, P_BKG0, MAX_BKG, P_INST0 = 0, 1, 2, 3, 24, 27
BKG_CHEBYSHEV, BKG_SPLINE = 0, 1
# the dataset's instrument function of a separated profile, `inst.<name>`, with the
# defaults of js/rietveld_core.js:413, 454-459
INST_NAMES = ['U', 'V', 'W', 'X', 'Y', 'Z', 'SL', 'HL', 'eta', 'U_L', 'V_L', 'W_L', 'X_L', 'U_R', 'V_R', 'W_R', 'X_R',
              'eta_split']
INST_DEFAULTS = [2, -2, 5, 1, 0, 0, 0, 0, 0.5, 2, -2, 5, 1, 2, -2, 5, 1, 0.5]
P_PHASE0 = P_INST0 + len(INST_NAMES)
# sample slots of a separated profile (js/rietveld_core.js:414-415 SAMPLE_PROFILE order)
QUARTIC = ['400', '040', '004', '220', '202', '022', '310', '301', '130', '103', '031', '013', '211', '121', '112']
SAMPLE_SLOTS = ['size', 'sizeAxial', 'sizeB', 'sizeC', 'rotA', 'rotB', 'rotC', 'strain', 'strainAxial', 'sizeMix',
                'strainMix'] + ['S' + q for q in QUARTIC]
SAMPLE_DEFAULTS = {'size': 1, 'sizeAxial': 1, 'sizeB': 1, 'sizeC': 1, 'rotA': 0, 'rotB': 0, 'rotC': 0, 'strain': 0,
                   'strainAxial': 0, 'sizeMix': 1, 'strainMix': 1}
SIZE_MODELS = {'isotropic': 0, 'uniaxial': 1, 'ellipsoid': 2}
STRAIN_MODELS = {'isotropic': 0, 'uniaxial': 1, 'stephens': 2}
PH = {'a': 0, 'b': 1, 'c': 2, 'alpha': 3, 'beta': 4, 'gamma': 5, 'scale': 6, 'po': 7}
PH_PROF0, PH_NPROF, AT_STRIDE = 8, 32, 5
PH_ATOM0 = PH_PROF0 + PH_NPROF
AT = {'x': 0, 'y': 1, 'z': 2, 'occ': 3, 'B': 4}
PROFILE = {'simple_pvoigt': 1, 'split_pvoigt': 2, 'tch_aniso': 3, 'gsas_cw': 4}
PROFILE_SLOTS = {
    'tch_aniso': ['U', 'V', 'W', 'X', 'Y', 'SL', 'HL'],
    'simple_pvoigt': ['GU', 'GV', 'GW', 'GP', 'LX', 'eta'],
    'split_pvoigt': ['GU_L', 'GV_L', 'GW_L', 'LX_L', 'GU_R', 'GV_R', 'GW_R', 'LX_R', 'eta_split'],
}
# start values of the profile families, js/rietveld_core.js:446-451
PROFILE_DEFAULTS = {
    'tch_aniso': {'U': 2, 'V': -2, 'W': 5, 'X': 1, 'Y': 0, 'SL': 0, 'HL': 0},
    'simple_pvoigt': {'GU': 2, 'GV': -2, 'GW': 5, 'GP': 0, 'LX': 1, 'eta': 0.5},
    'split_pvoigt': {'GU_L': 2, 'GV_L': -2, 'GW_L': 5, 'LX_L': 1, 'GU_R': 2, 'GV_R': -2, 'GW_R': 5, 'LX_R': 1, 'eta_split': 0.5},
}
SYSTEMS = {'triclinic': 0, 'monoclinic': 1, 'orthorhombic': 2, 'tetragonal': 3, 'trigonal': 4, 'hexagonal': 5, 'cubic': 6}
CU_KA1, CU_KA2, CU_RATIO = 1.540598, 1.544426, 0.5        # rv_state.js:32-37


class Atom:
    def __init__(self, label, element, site, occ=1.0, B=1.0, ff=None, b_neutron=0.0, fp=0.0, fpp=0.0):
        self.label, self.element, self.site = label, element, site
        self.occ, self.B = occ, B
        if ff is None:
            ff = WK1995[element]
        self.ff_a, self.ff_b, self.ff_c = list(ff[0]), list(ff[1]), ff[2]
        self.n_gauss = len(self.ff_a)
        self.b_neutron, self.fp, self.fpp = b_neutron, fp, fpp


class Phase:
    """A Rietveld phase.  profile 'gsas_cw' (or separated=True with another family) is a
    separated profile: the dataset's instrument function plus sample size / strain
    broadening (js/rietveld_core.js:460), with `broadening` as the reference's
    phase.broadening: {'size': 'isotropic'|'uniaxial'|'ellipsoid', 'strain':
    'isotropic'|'uniaxial'|'stephens', 'sizeAxis': [h, k, l], 'strainAxis': [h, k, l]}."""

    def __init__(self, key, group, cell, atoms, profile='tch_aniso', po_dir=(0, 0, 1), unique_axis=1, ops=None,
                 separated=False, broadening=None, method='rietveld'):
        self.key = key
        # 'lebail' / 'wppf': free reflection intensities, fitted per trial (no atoms used)
        self.method = method
        self.separated = bool(separated) or profile == 'gsas_cw'
        b = dict(broadening or {})
        self.size_model = b.get('size', 'isotropic')
        self.strain_model = b.get('strain', 'isotropic')
        self.size_axis = list(b.get('sizeAxis', [0, 0, 1]))
        self.strain_axis = list(b.get('strainAxis', [0, 0, 1]))
        if ops is None:
            ops = S.group_ops(group)
        self.op_r, self.op_t = ops
        self.system = S.group_system(group) if group is not None else 'triclinic'
        self.unique_axis = unique_axis
        self.cell = list(cell)
        self.atoms = atoms
        self.profile = profile
        self.po_dir = list(po_dir)


class Problem:
    """One dataset, its phases, model parameters p = base + M x, variables, PT settings."""

    def __init__(self, tth, phases, lambda1=CU_KA1, lambda2=0.0, ratio=0.0, pol_k=1.0, min_fwhm=None,
                 n_bkg=6, radiation=0, bkg_x=None, resolution_type=None):
        self.tth = np.ascontiguousarray(tth, dtype=np.float64)
        self.y = np.zeros_like(self.tth)
        self.w = np.ones_like(self.tth)
        self.lambda1, self.lambda2, self.ratio, self.pol_k = lambda1, lambda2, ratio, pol_k
        if min_fwhm is None:
            min_fwhm = max(1e-4, 2 * float(np.median(np.diff(self.tth))))
        self.min_fwhm = min_fwhm
        self.n_bkg, self.radiation = n_bkg, radiation
        # spline background: anchors at these 2-theta (finite, sorted, as the reference's
        # dataset build keeps them, js/rietveld_core.js:654), values bga0..bga<n-1>
        self.bkg_type, self.bkg_x = BKG_CHEBYSHEV, None
        if bkg_x is not None:
            xs = np.sort(np.asarray([v for v in map(float, bkg_x) if math.isfinite(v)], dtype=np.float64))
            if len(xs) > MAX_BKG:
                raise ValueError('at most %d spline anchors' % MAX_BKG)
            self.bkg_type, self.bkg_x, self.n_bkg = BKG_SPLINE, xs, len(xs)
        # instrument family of the separated phases (the dataset's resolutionType), or None
        self.resolution_type = resolution_type
        self.phases = phases
        self.names = []
        self.index = {}
        self._layout()
        self.base = np.zeros(len(self.names))
        self.vars, self.x0, self.lo, self.hi, self.scale = [], [], [], [], []
        self.terms = []       # (param, var, coeff)
        self.chains, self.sweeps, self.steps_per_sweep, self.max_temp, self.seed = 6, 50, 0, 100.0, 12345

    # -- layout ------------------------------------------------------------------
    def _layout(self):
        stem = 'bga' if self.bkg_type == BKG_SPLINE else 'bkg'
        names = ['zero', 'disp', 'transp'] + [stem + '%d' % k for k in range(MAX_BKG)]
        names += ['inst.' + n for n in INST_NAMES]
        for ph in self.phases:
            pre = ph.key + '.'
            names += [pre + n for n in ('a', 'b', 'c', 'alpha', 'beta', 'gamma', 'scale', 'po')]
            slots = SAMPLE_SLOTS if ph.separated else PROFILE_SLOTS[ph.profile]
            names += [pre + (slots[i] if i < len(slots) else '_prof%d' % i) for i in range(PH_NPROF)]
            for at in ph.atoms:
                names += [pre + at.label + '.' + n for n in ('x', 'y', 'z', 'occ', 'B')]
        self.names = names
        self.index = {n: i for i, n in enumerate(names)}

    def phase_base(self, q):
        b = P_PHASE0
        for ph in self.phases[:q]:
            b += PH_ATOM0 + AT_STRIDE * len(ph.atoms)
        return b

    def set(self, name, value):
        self.base[self.index[name]] = value

    def values(self, x=None):
        """All model parameters at the variables x (default x0): base + M x."""
        p = self.base.copy()
        xx = np.asarray(self.x0 if x is None else x, dtype=np.float64)
        for (k, j, c) in self.terms:
            p[k] += c * xx[j]
        return p

    def refine(self, name, lo=-math.inf, hi=math.inf, scale=None):
        """Make model parameter `name` a variable, started at its current base value."""
        k = self.index[name]
        j = len(self.vars)
        v = float(self.base[k])
        self.base[k] = 0.0
        self.vars.append(name)
        self.x0.append(min(hi, max(lo, v)))
        self.lo.append(lo)
        self.hi.append(hi)
        self.scale.append(pt_scale(name, v) if scale is None else scale)
        self.terms.append((k, j, 1.0))
        return j

    def link(self, label, names, lo=-math.inf, hi=math.inf, scale=None):
        """One variable driving several model parameters with coefficient 1 (the affine
        links of js/atomic_constraints.js), started at the mean of their base values."""
        ks = [self.index[n] for n in names]
        v = float(np.mean([self.base[k] for k in ks]))
        j = len(self.vars)
        for k in ks:
            self.base[k] = 0.0
            self.terms.append((k, j, 1.0))
        self.vars.append(label)
        self.x0.append(min(hi, max(lo, v)))
        self.lo.append(lo)
        self.hi.append(hi)
        self.scale.append(pt_scale(names[0], v) if scale is None else scale)
        return j

    def constrain(self, name, offset, terms):
        """Model parameter `name` = offset + sum_(j, c) c x_j: a dependent parameter of an
        exact constraint (js/atomic_constraints.js:76-155 hold / equiv / sum, and the
        cross-dataset parameterLinks :134-148), in the form PT contexts carry it
        (ctx.constraints.links, {d, name, offset, terms}).  Recorded in `links` too."""
        k = self.index[name]
        self.base[k] = float(offset)
        for j, c in terms:
            self.terms.append((k, int(j), float(c)))
        if not hasattr(self, 'links'):
            self.links = []
        self.links.append((name, float(offset), [(int(j), float(c)) for j, c in terms]))

    @property
    def nv(self):
        return len(self.vars)

    def default_values(self):
        """Fill the parameters the reference would hold: cell, unit occupancies, phase
        scale and po = 1, profile defaults."""
        for q, ph in enumerate(self.phases):
            pre = ph.key + '.'
            for n, v in zip(('a', 'b', 'c', 'alpha', 'beta', 'gamma'), ph.cell):
                self.set(pre + n, v)
            self.set(pre + 'scale', 1.0)
            self.set(pre + 'po', 1.0)
            defaults = SAMPLE_DEFAULTS if ph.separated else PROFILE_DEFAULTS[ph.profile]
            for n, v in defaults.items():
                self.set(pre + n, v)
            if ph.separated:
                for n, v in zip(INST_NAMES, INST_DEFAULTS):
                    self.set('inst.' + n, v)
            for at in ph.atoms:
                pos = at.site['pos']
                for i, n in enumerate('xyz'):
                    self.set(pre + at.label + '.' + n, float(pos[i]))
                self.set(pre + at.label + '.occ', at.occ)
                self.set(pre + at.label + '.B', at.B)

    def observe(self, simulate, x_true, seed, peak=1e4):
        """y = Poisson(model at x_true scaled so its tallest net peak is `peak` counts)."""
        ycalc = np.asarray(simulate(self, x_true), dtype=np.float64)
        bk = self.background(x_true)
        net = float(np.max(ycalc - bk))
        if net > 0 and peak:
            f = peak / net
            for q, ph in enumerate(self.phases):
                k = self.index[ph.key + '.scale']
                self._scale_param(k, f, x_true)
            ycalc = np.asarray(simulate(self, x_true), dtype=np.float64)
        rng = np.random.default_rng(seed)
        y = rng.poisson(np.maximum(ycalc, 0.0)).astype(np.float64)
        self.y = y
        self.w = 1.0 / np.maximum(y, 1.0)
        return ycalc

    def _scale_param(self, k, f, x_true):
        hit = False
        for (kk, j, c) in self.terms:
            if kk == k:
                x_true[j] *= f
                self.x0[j] *= f
                hit = True
        if not hit:
            self.base[k] *= f

    def background(self, x):
        p = self.values(x)
        t = self.tth
        if self.bkg_type == BKG_SPLINE:
            return spline_background(t, self.bkg_x, p[P_BKG0:P_BKG0 + self.n_bkg])
        lo, hi = t[0], t[-1]
        z = 2 * (t - lo) / (hi - lo) - 1 if hi > lo else np.zeros_like(t)
        T = [np.ones_like(t), z]
        y = np.zeros_like(t)
        for k in range(self.n_bkg):
            if k >= 2:
                T.append(2 * z * T[k - 1] - T[k - 2])
            y += p[P_BKG0 + k] * T[k]
        return y

    def start_vector(self):
        return np.array(self.x0, dtype=np.float64)

    def var_index(self, name):
        return self.vars.index(name)


def classify(name):
    """Kind of a parameter name, js/rietveld_core.js:470-484 classify (a variable label
    name@d of a multi-dataset problem classifies as name)."""
    name = name.split('@')[0]
    if name in ('zero', 'disp', 'transp'):
        return name, None
    if name.startswith('bkg'):
        return 'bkg', None
    if name.startswith('bga'):
        return 'bga', None
    if name.startswith('inst.'):
        return 'instrument', name[5:]
    parts = name.split('.')
    if len(parts) == 2:
        if parts[1] == 'scale':
            return 'scale', None
        if parts[1] == 'po':
            return 'po', None
        if parts[1] in ('a', 'b', 'c', 'alpha', 'beta', 'gamma'):
            return 'cell', parts[1]
        return 'profile', parts[1]
    if len(parts) == 3:
        return 'atom', parts[2]
    return 'unknown', None


def step_cap(name, v):
    """js/rietveld_core.js:511-523 stepCap."""
    kind, comp = classify(name)
    if kind == 'cell':
        return 0.02 * abs(v) + 1e-6 if comp in ('a', 'b', 'c') else 1.0
    if kind == 'atom':
        return 0.2 if comp == 'occ' else 2.0 if comp == 'B' else 0.05
    if kind == 'zero':
        return 0.1
    if kind == 'disp':
        return 2.0
    if kind == 'transp':
        return 0.01
    if kind == 'po':
        return 0.25
    if kind == 'instrument':
        return 0.01 if comp in ('SL', 'HL') else 0.5 * max(abs(v), 1)
    if kind == 'profile':
        if comp in ('sizeMix', 'strainMix'):
            return 0.2
        if comp.startswith('size'):
            return 0.2 * max(abs(v), 0.001)
        if comp.startswith('eta'):
            return 0.25
        if comp in ('SL', 'HL'):
            return 0.05
        return 0.5 * max(abs(v), 5)
    return math.inf


def pt_scale(name, v):
    """The PT step scale of js/rietveld_core.js:2161-2167 without the curvature term:
    natural = 0.15 * stepCap, or 0.1 * max(1, |x|) for an uncapped kind."""
    cap = step_cap(name, v)
    natural = 0.15 * cap if math.isfinite(cap) else 0.1 * max(1.0, abs(v))
    return max(natural, 1e-8)


def _grid(lo, hi, step):
    n = int(round((hi - lo) / step)) + 1
    return lo + step * np.arange(n, dtype=np.float64)


def spline_background(tth, xs, ys):
    """Monotone cubic Hermite background through the anchors (xs, ys) with Fritsch-Carlson
    tangents, flat outside the anchors (js/rietveld_core.js:549-612); NaN values read as
    0 (Model.splineY, :932-936).  Host-side, for synthetic observations."""
    t = np.asarray(tth, dtype=np.float64)
    xs = np.asarray(xs, dtype=np.float64)
    ys = np.nan_to_num(np.asarray(ys, dtype=np.float64), nan=0.0)
    na = len(xs)
    if na == 0:
        return np.zeros_like(t)
    if na == 1:
        return np.full_like(t, ys[0])
    h = np.diff(xs)
    d = np.diff(ys) / h
    m = np.empty(na)
    m[0], m[-1] = d[0], d[-1]
    for i in range(1, na - 1):
        m[i] = 0.0 if d[i - 1] * d[i] <= 0 else (h[i] * d[i - 1] + h[i - 1] * d[i]) / (h[i - 1] + h[i])
    for i in range(na - 1):
        if d[i] == 0:
            m[i] = m[i + 1] = 0.0
            continue
        a, b = m[i] / d[i], m[i + 1] / d[i]
        tau = a * a + b * b
        if tau > 9:
            f = 3 / math.sqrt(tau)
            m[i], m[i + 1] = a * d[i] * f, b * d[i] * f
    s = np.clip(np.searchsorted(xs, t, side='right') - 1, 0, na - 2)
    u = (t - xs[s]) / h[s]
    u2, u3 = u * u, u * u * u
    y = (2 * u3 - 3 * u2 + 1) * ys[s] + (u3 - 2 * u2 + u) * h[s] * m[s] + (-2 * u3 + 3 * u2) * ys[s + 1] \
        + (u3 - u2) * h[s] * m[s + 1]
    y = np.where(t <= xs[0], ys[0], y)
    return np.where(t >= xs[-1], ys[-1], y)


def _cheb(problem, coeffs):
    for k, v in enumerate(coeffs):
        problem.set('bkg%d' % k, v)


# ---------------------------------------------------------------------------
#  workloads
# ---------------------------------------------------------------------------
def nacl_fm3m(simulate=None, seed=20260101, chains=16, sweeps=1000):
    """NaCl Fm-3m a=5.64, Na 4a, Cl 4b; Cu Ka1 only; 2theta 10-120 step 0.02 (5501 pts);
    TCH widths U,V,W,X,Y = 18.03, -9.02, 5.41, 5, 2 (GSAS cdeg units: FullProf 0.01,
    -0.005, 0.003 deg^2 x 1803.4 and 0.05, 0.02 deg x 100); 6-term Chebyshev; 12 refined
    parameters: a, zero, scale, U, V, W, X, Y, B_Na, B_Cl, bkg0, bkg1."""
    r, t = S.group_ops(225)
    cell = [5.64, 5.64, 5.64, 90, 90, 90]
    atoms = [Atom('Na1', 'Na', S.site_symmetry(r, t, [0, 0, 0], cell), B=1.1),
             Atom('Cl1', 'Cl', S.site_symmetry(r, t, [0.5, 0.5, 0.5], cell), B=0.9)]
    ph = Phase('p1', 225, cell, atoms, 'tch_aniso')
    pb = Problem(_grid(10.0, 120.0, 0.02), [ph], lambda1=1.5406, lambda2=0.0, ratio=0.0, pol_k=1.0)
    pb.default_values()
    for n, v in zip(('U', 'V', 'W', 'X', 'Y'), (18.03, -9.02, 5.41, 5.0, 2.0)):
        pb.set('p1.' + n, v)
    pb.set('zero', 0.012)
    _cheb(pb, [220.0, -60.0, 25.0, -12.0, 6.0, -3.0])
    pb.set('p1.scale', 1.0)
    names = ['p1.a', 'zero', 'p1.scale', 'p1.U', 'p1.V', 'p1.W', 'p1.X', 'p1.Y', 'p1.Na1.B', 'p1.Cl1.B', 'bkg0', 'bkg1']
    limits = {'p1.a': (5.64 * 0.96, 5.64 * 1.04), 'zero': (-2, 2), 'p1.scale': (0, math.inf), 'p1.X': (0, math.inf),
              'p1.Y': (0, math.inf), 'p1.Na1.B': (-2, 30), 'p1.Cl1.B': (-2, 30)}
    for n in names:
        lo, hi = limits.get(n, (-math.inf, math.inf))
        pb.refine(n, lo, hi)
    pb.chains, pb.sweeps, pb.steps_per_sweep, pb.max_temp, pb.seed = chains, sweeps, 0, 100.0, 137
    x_true = pb.start_vector()
    if simulate is not None:
        pb.observe(simulate, x_true, seed)
        x0 = _perturb(pb, x_true, seed + 1)
        pb.x0 = list(x0)
    return pb


def nacl_spline(simulate=None, seed=20260111, chains=16, sweeps=1000, n_anchors=8, atoms_only=False):
    """The NaCl Fm-3m pattern of nacl_fm3m with a spline background instead of the
    Chebyshev series: `n_anchors` anchors spread over 10-120 deg (a falling air-scatter
    hump, 260 -> 90 counts, with a bump near 30 deg), every anchor value bga<j> refined
    with a, zero, scale, U, V, W, X, Y; with atoms_only the two B factors are the only
    variables (the structure moves of the shared-profile form)."""
    r, t = S.group_ops(225)
    cell = [5.64, 5.64, 5.64, 90, 90, 90]
    atoms = [Atom('Na1', 'Na', S.site_symmetry(r, t, [0, 0, 0], cell), B=1.1),
             Atom('Cl1', 'Cl', S.site_symmetry(r, t, [0.5, 0.5, 0.5], cell), B=0.9)]
    ph = Phase('p1', 225, cell, atoms, 'tch_aniso')
    xs = np.linspace(10.0, 120.0, n_anchors)
    pb = Problem(_grid(10.0, 120.0, 0.02), [ph], lambda1=1.5406, lambda2=0.0, ratio=0.0, pol_k=1.0, bkg_x=xs)
    pb.default_values()
    for n, v in zip(('U', 'V', 'W', 'X', 'Y'), (18.03, -9.02, 5.41, 5.0, 2.0)):
        pb.set('p1.' + n, v)
    pb.set('zero', 0.012)
    ys = 90.0 + 170.0 * np.exp(-(xs - 10.0) / 45.0) + 40.0 * np.exp(-0.5 * ((xs - 30.0) / 8.0) ** 2)
    for j, v in enumerate(ys):
        pb.set('bga%d' % j, float(v))
    pb.set('p1.scale', 1.0)
    names = ['p1.a', 'zero', 'p1.scale', 'p1.U', 'p1.V', 'p1.W', 'p1.X', 'p1.Y'] + ['bga%d' % j for j in range(n_anchors)]
    if atoms_only:
        names = ['p1.Na1.B', 'p1.Cl1.B']
    limits = {'p1.a': (5.64 * 0.96, 5.64 * 1.04), 'zero': (-2, 2), 'p1.scale': (0, math.inf), 'p1.X': (0, math.inf),
              'p1.Y': (0, math.inf), 'p1.Na1.B': (-2, 30), 'p1.Cl1.B': (-2, 30)}
    for n in names:
        lo, hi = limits.get(n, (-math.inf, math.inf))
        pb.refine(n, lo, hi)
    pb.chains, pb.sweeps, pb.steps_per_sweep, pb.max_temp, pb.seed = chains, sweeps, 0, 100.0, 139
    x_true = pb.start_vector()
    if simulate is not None:
        pb.observe(simulate, x_true, seed)
        x0 = _perturb(pb, x_true, seed + 1)
        pb.x0 = list(x0)
    return pb


def zincblende_f43m(simulate=None, seed=20260505, chains=16, anomalous=False, neutron=False):
    """ZnS F-43m (non-centrosymmetric) a=5.41, Zn 4a (0,0,0), S 4c (1/4,1/4,1/4); Cu
    Ka1/Ka2; 2theta 20-130 step 0.02; optional anomalous f', f'' on Zn (-1.6, 0.68 at Cu Ka)
    or neutron lengths (Zn 5.68 fm, S 2.847 fm).  Refined: a, zero, scale, U, V, W, B_Zn,
    B_S, bkg0, bkg1."""
    r, t = S.group_ops(216)
    cell = [5.41, 5.41, 5.41, 90, 90, 90]
    fp, fpp = (-1.6, 0.68) if anomalous else (0.0, 0.0)
    atoms = [Atom('Zn1', 'Zn', S.site_symmetry(r, t, [0, 0, 0], cell), B=0.7, fp=fp, fpp=fpp, b_neutron=5.68),
             Atom('S1', 'S', S.site_symmetry(r, t, [0.25, 0.25, 0.25], cell), B=0.9, b_neutron=2.847)]
    ph = Phase('p1', 216, cell, atoms, 'tch_aniso')
    pb = Problem(_grid(20.0, 130.0, 0.02), [ph], lambda1=CU_KA1, lambda2=CU_KA2, ratio=CU_RATIO, pol_k=1.0,
                 radiation=1 if neutron else 0)
    pb.default_values()
    for n, v in zip(('U', 'V', 'W', 'X', 'Y'), (18.03, -9.02, 5.41, 5.0, 2.0)):
        pb.set('p1.' + n, v)
    _cheb(pb, [120.0, -20.0, 8.0, -3.0, 1.0, -0.5])
    limits = {'p1.a': (5.41 * 0.97, 5.41 * 1.03), 'zero': (-1, 1), 'p1.scale': (0, math.inf), 'p1.Zn1.B': (-2, 30),
              'p1.S1.B': (-2, 30)}
    for n in ('p1.a', 'zero', 'p1.scale', 'p1.U', 'p1.V', 'p1.W', 'p1.Zn1.B', 'p1.S1.B', 'bkg0', 'bkg1'):
        lo, hi = limits.get(n, (-math.inf, math.inf))
        pb.refine(n, lo, hi)
    pb.chains, pb.sweeps, pb.steps_per_sweep, pb.max_temp, pb.seed = chains, 10, 0, 100.0, 211
    if simulate is not None:
        x_true = pb.start_vector()
        pb.observe(simulate, x_true, seed)
        pb.x0 = list(_perturb(pb, x_true, seed + 1))
    return pb


def general_noncentro(group, simulate=None, seed=20260606, chains=16):
    """A non-centrosymmetric phase in general positions: P2_1 (group 4, a=5.2 b=6.3 c=7.1
    beta=101, 3 atoms) or P2_12_12_1 (group 19, a=5.2 b=6.3 c=7.1, 3 atoms); Cu Ka1/Ka2,
    2theta 10-110 step 0.02; every coordinate and B refined with the cell, zero, scale."""
    rng = np.random.default_rng(seed + group)
    r, t = S.group_ops(group)
    cell = [5.2, 6.3, 7.1, 90, 101.0 if group == 4 else 90, 90]
    atoms = []
    for j, el in enumerate(('Fe', 'O', 'C')):
        xyz = list(rng.uniform(0.05, 0.95, size=3))
        atoms.append(Atom('%s%d' % (el, j + 1), el, S.site_symmetry(r, t, xyz, cell), B=float(rng.uniform(0.5, 2.0))))
    ph = Phase('p1', group, cell, atoms, 'tch_aniso', unique_axis=1)
    pb = Problem(_grid(10.0, 110.0, 0.02), [ph], lambda1=CU_KA1, lambda2=CU_KA2, ratio=CU_RATIO, pol_k=1.0)
    pb.default_values()
    for n, v in zip(('U', 'V', 'W', 'X', 'Y'), (18.03, -9.02, 5.41, 5.0, 2.0)):
        pb.set('p1.' + n, v)
    _cheb(pb, [100.0, -10.0, 5.0])
    for n in ('a', 'b', 'c'):
        v = pb.base[pb.index['p1.' + n]]
        pb.refine('p1.' + n, 0.98 * v, 1.02 * v)
    pb.refine('zero', -1, 1)
    pb.refine('p1.scale', 0, math.inf)
    for at in atoms:
        for f in at.site['free']:
            n = 'p1.%s.%s' % (at.label, 'xyz'[f])
            v = pb.base[pb.index[n]]
            pb.refine(n, v - 0.1, v + 0.1)
        pb.refine('p1.%s.B' % at.label, -2, 30)
    pb.chains, pb.sweeps, pb.steps_per_sweep, pb.max_temp, pb.seed = chains, 10, 0, 100.0, 300 + group
    if simulate is not None:
        x_true = pb.start_vector()
        pb.observe(simulate, x_true, seed)
        pb.x0 = list(_perturb(pb, x_true, seed + 1))
    return pb


def two_phase(simulate=None, seed=20260707, chains=16):
    """NaCl (Fm-3m, a=5.64) + ZnS (F-43m, a=5.41) in one pattern, Cu Ka1/Ka2, 2theta 20-120
    step 0.02; each phase with its own cell, scale, widths and B; shared zero and background."""
    r1, t1 = S.group_ops(225)
    r2, t2 = S.group_ops(216)
    c1, c2 = [5.64, 5.64, 5.64, 90, 90, 90], [5.41, 5.41, 5.41, 90, 90, 90]
    ph1 = Phase('p1', 225, c1, [Atom('Na1', 'Na', S.site_symmetry(r1, t1, [0, 0, 0], c1), B=1.1),
                                Atom('Cl1', 'Cl', S.site_symmetry(r1, t1, [0.5, 0.5, 0.5], c1), B=0.9)], 'tch_aniso')
    ph2 = Phase('p2', 216, c2, [Atom('Zn1', 'Zn', S.site_symmetry(r2, t2, [0, 0, 0], c2), B=0.7),
                                Atom('S1', 'S', S.site_symmetry(r2, t2, [0.25, 0.25, 0.25], c2), B=0.9)], 'tch_aniso')
    pb = Problem(_grid(20.0, 120.0, 0.02), [ph1, ph2], lambda1=CU_KA1, lambda2=CU_KA2, ratio=CU_RATIO, pol_k=1.0)
    pb.default_values()
    for q in ('p1', 'p2'):
        for n, v in zip(('U', 'V', 'W', 'X', 'Y'), (18.03, -9.02, 5.41, 5.0, 2.0)):
            pb.set(q + '.' + n, v)
    pb.set('p2.scale', 0.6)
    _cheb(pb, [150.0, -30.0, 10.0])
    for n, lo, hi in (('p1.a', 5.5, 5.8), ('p2.a', 5.3, 5.5), ('zero', -1, 1), ('p1.scale', 0, math.inf),
                      ('p2.scale', 0, math.inf), ('p1.U', -math.inf, math.inf), ('p2.W', -math.inf, math.inf),
                      ('p1.Na1.B', -2, 30), ('p2.S1.B', -2, 30), ('bkg0', -math.inf, math.inf)):
        pb.refine(n, lo, hi)
    pb.chains, pb.sweeps, pb.steps_per_sweep, pb.max_temp, pb.seed = chains, 10, 0, 100.0, 401
    if simulate is not None:
        x_true = pb.start_vector()
        pb.observe(simulate, x_true, seed)
        pb.x0 = list(_perturb(pb, x_true, seed + 1))
    return pb


PBSO4_SITES = [  # anglesite-like PbSO4, Pnma (ITA standard setting), 4c x 4 + 8d
    ('Pb1', 'Pb', (0.188, 0.25, 0.167), 1.2),
    ('S1', 'S', (0.063, 0.25, 0.684), 0.8),
    ('O1', 'O', (0.908, 0.25, 0.595), 1.6),
    ('O2', 'O', (0.193, 0.25, 0.543), 1.5),
    ('O3', 'O', (0.081, 0.027, 0.809), 1.4),
]


def pbso4_pnma(simulate=None, seed=20260202, chains=256, steps_per_sweep=64, sweeps=10, jitter=0.05,
               atoms_only=False):
    """Pnma phase a~8.48 b~5.40 c~6.96, 5 independent sites (24 atoms), coordinates drawn
    within +-jitter of the seeded sites, B in U[0.5, 2]; 2theta 10-150 step 0.02 (7001
    points) with Cu Ka1/Ka2 (ratio 0.5); TCH profile; 6-term Chebyshev.  Refined: cell,
    zero, scale, U..Y, 11 free coordinates, 5 B, 6 background terms (32 variables), or
    only the coordinates, B and scale when atoms_only (the shared-profile workload)."""
    rng = np.random.default_rng(seed)
    r, t = S.group_ops(62)
    cell = [8.48, 5.40, 6.96, 90, 90, 90]
    atoms = []
    for (lab, el, xyz, B) in PBSO4_SITES:
        site = S.site_symmetry(r, t, xyz, cell)
        pos = list(site['pos'])
        for f in site['free']:
            pos[f] += rng.uniform(-jitter, jitter)
        site['pos'] = pos
        atoms.append(Atom(lab, el, site, B=float(rng.uniform(0.5, 2.0))))
    ph = Phase('p1', 62, cell, atoms, 'tch_aniso')
    pb = Problem(_grid(10.0, 150.0, 0.02), [ph], lambda1=CU_KA1, lambda2=CU_KA2, ratio=CU_RATIO, pol_k=1.0)
    pb.default_values()
    for n, v in zip(('U', 'V', 'W', 'X', 'Y'), (18.03, -9.02, 5.41, 5.0, 2.0)):
        pb.set('p1.' + n, v)
    pb.set('zero', -0.008)
    _cheb(pb, [180.0, -40.0, 18.0, -9.0, 4.0, -2.0])
    names = []
    if not atoms_only:
        names += ['p1.a', 'p1.b', 'p1.c', 'zero', 'p1.U', 'p1.V', 'p1.W', 'p1.X', 'p1.Y']
    names += ['p1.scale']
    for at in atoms:
        names += ['p1.%s.%s' % (at.label, 'xyz'[f]) for f in at.site['free']]
    for at in atoms:
        names.append('p1.%s.B' % at.label)
    if not atoms_only:
        names += ['bkg%d' % k for k in range(6)]
    for n in names:
        kind, comp = classify(n)
        if kind == 'cell':
            v = pb.base[pb.index[n]]
            pb.refine(n, 0.96 * v, 1.04 * v)
        elif n == 'zero':
            pb.refine(n, -2, 2)
        elif n == 'p1.scale':
            pb.refine(n, 0, math.inf)
        elif kind == 'profile' and comp in ('X', 'Y'):
            pb.refine(n, 0, math.inf)
        elif kind == 'atom' and comp == 'B':
            pb.refine(n, -2, 30)
        elif kind == 'atom':
            v = pb.base[pb.index[n]]
            pb.refine(n, v - 0.1, v + 0.1)
        else:
            pb.refine(n)
    pb.chains, pb.sweeps, pb.steps_per_sweep, pb.max_temp, pb.seed = chains, sweeps, steps_per_sweep, 100.0, 4242
    x_true = pb.start_vector()
    if simulate is not None:
        pb.observe(simulate, x_true, seed + 7)
        pb.x0 = list(_perturb(pb, x_true, seed + 11))
    return pb


def _perturb(pb, x_true, seed, frac=0.3):
    """A start vector away from the truth: each variable moved by +-frac of its PT scale
    times a few, clipped to its box."""
    rng = np.random.default_rng(seed)
    x = np.array(x_true, dtype=np.float64)
    for j, n in enumerate(pb.vars):
        kind, comp = classify(n)
        if kind in ('bkg', 'bga') or n.endswith('.scale'):
            continue
        x[j] += rng.uniform(-1, 1) * frac * 4 * pb.scale[j]
        x[j] = min(pb.hi[j], max(pb.lo[j], x[j]))
    return x


MOLECULE_ATOMS = [  # 17 non-H atoms, general positions (fractional, seeded layout)
    ('S1', 'S'), ('N1', 'N'), ('N2', 'N'), ('N3', 'N'), ('N4', 'N'), ('N5', 'N'), ('N6', 'N'),
    ('C1', 'C'), ('C2', 'C'), ('C3', 'C'), ('C4', 'C'), ('C5', 'C'), ('C6', 'C'), ('C7', 'C'),
    ('C8', 'C'), ('C9', 'C'), ('C10', 'C'),
]


def molecular_p21c(simulate=None, seed=20260404, chains=512, steps_per_sweep=0, sweeps=10, shared_b=False):
    """P2_1/c (unique axis b, -P 2ybc) a=10.4 b=18.8 c=6.8 beta=106; 17 C/N/S atoms in
    general positions (a seeded random molecule-like cluster); 2theta 5-145 step 0.02
    (7001 points), Cu Ka1/Ka2; refined: 51 coordinates, scale, and either 17 atomic B
    (69 variables) or, with shared_b, one Biso linked to all 17 atoms (53 variables: the
    input width of the 64-wide surrogate network)."""
    rng = np.random.default_rng(seed)
    r, t = S.group_ops(14)
    cell = [10.4, 18.8, 6.8, 90, 106.0, 90]
    centre = np.array([0.25, 0.12, 0.3])
    atoms = []
    b_common = float(rng.uniform(2.0, 4.0))
    for (lab, el) in MOLECULE_ATOMS:
        xyz = centre + rng.uniform(-0.18, 0.18, size=3) * np.array([1.0, 0.55, 1.4])
        site = S.site_symmetry(r, t, list(xyz), cell)
        b = float(rng.uniform(2.0, 4.0))
        atoms.append(Atom(lab, el, site, B=b_common if shared_b else b))
    ph = Phase('p1', 14, cell, atoms, 'tch_aniso', unique_axis=1)
    pb = Problem(_grid(5.0, 145.0, 0.02), [ph], lambda1=CU_KA1, lambda2=CU_KA2, ratio=CU_RATIO, pol_k=1.0)
    pb.default_values()
    for n, v in zip(('U', 'V', 'W', 'X', 'Y'), (18.03, -9.02, 5.41, 5.0, 2.0)):
        pb.set('p1.' + n, v)
    _cheb(pb, [150.0, -30.0, 12.0, -6.0, 3.0, -1.5])
    pb.refine('p1.scale', 0, math.inf)
    for at in atoms:
        for f in at.site['free']:
            n = 'p1.%s.%s' % (at.label, 'xyz'[f])
            v = pb.base[pb.index[n]]
            pb.refine(n, v - 0.15, v + 0.15)
    if shared_b:
        pb.link('p1.Biso', ['p1.%s.B' % at.label for at in atoms], -2, 30)
    else:
        for at in atoms:
            pb.refine('p1.%s.B' % at.label, -2, 30)
    pb.chains, pb.sweeps, pb.steps_per_sweep, pb.max_temp, pb.seed = chains, sweeps, steps_per_sweep, 100.0, 77
    x_true = pb.start_vector()
    if simulate is not None:
        pb.observe(simulate, x_true, seed + 3)
        pb.x0 = list(_perturb(pb, x_true, seed + 5))
    return pb


UNEVEN_GROUPS = [(221, 'cubic'), (225, 'cubic'), (229, 'cubic'), (123, 'tetragonal'), (139, 'tetragonal'),
                 (136, 'tetragonal'), (14, 'monoclinic'), (15, 'monoclinic'), (10, 'monoclinic')]
UNEVEN_ELEMENTS = ['O', 'Mg', 'Si', 'S', 'K', 'Ca', 'Ti', 'Fe', 'Cu', 'Zn', 'Ba', 'Pb', 'C', 'N', 'Na', 'Cl']


def uneven_pattern(index, simulate=None, seed=20260303, chains=32):
    """Pattern `index` of the uneven batch: a cubic, tetragonal or monoclinic cell with
    edges in [4, 20] A, 1-12 sites, 2000-12000 points.  The reflection count is aimed at a
    log-uniform target in [50, 3000]: the cell is scaled so that the lattice points in the
    envelope sphere, (4 pi / 3) q^3 V, divided by the operator count (Laue orbit x
    centring) meet the target, edges clipped to [4, 20] A."""
    rng = np.random.default_rng(seed + 1000 * index)
    grp, system = UNEVEN_GROUPS[index % len(UNEVEN_GROUPS)]
    r, t = S.group_ops(grp)
    n_pts = int(rng.integers(2000, 12001))
    lo = float(rng.uniform(5.0, 20.0))
    step = float(rng.choice([0.01, 0.015, 0.02]))
    step = min(step, (160.0 - lo) / n_pts)
    tth = lo + step * np.arange(n_pts, dtype=np.float64)
    target = math.exp(rng.uniform(math.log(50.0), math.log(3000.0)))
    # envelope sphere: 2theta_max + 10 deg + the zero-shift limit (2 deg), lambda = Ka1
    th = math.radians(min(179.0, tth[-1] + 12.0) / 2.0)
    q = 2.0 * math.sin(th) / CU_KA1
    if system == 'cubic':
        shape = [1.0, 1.0, 1.0]
        ang = [90, 90, 90]
    elif system == 'tetragonal':
        shape = [1.0, 1.0, rng.uniform(0.4, 2.5)]
        ang = [90, 90, 90]
    else:
        shape = [1.0, rng.uniform(0.6, 1.8), rng.uniform(0.5, 1.5)]
        ang = [90, rng.uniform(95, 115), 90]
    sin_b = math.sin(math.radians(ang[1]))
    v_shape = shape[0] * shape[1] * shape[2] * sin_b
    v_target = target * len(r) / (4.0 * math.pi / 3.0 * q ** 3)
    f = (v_target / v_shape) ** (1.0 / 3.0)
    edges = [min(20.0, max(4.0, f * e)) for e in shape]
    if system == 'cubic':
        edges = [edges[0]] * 3
    elif system == 'tetragonal':
        edges[1] = edges[0]
    cell = edges + ang
    n_sites = int(rng.integers(1, 13))
    atoms = []
    for s_i in range(n_sites):
        xyz = list(rng.uniform(0.02, 0.98, size=3))
        site = S.site_symmetry(r, t, xyz, cell)
        el = UNEVEN_ELEMENTS[int(rng.integers(0, len(UNEVEN_ELEMENTS)))]
        atoms.append(Atom('A%d' % (s_i + 1), el, site, B=float(rng.uniform(0.5, 2.0))))
    ph = Phase('p1', grp, cell, atoms, 'tch_aniso', unique_axis=1)
    pb = Problem(tth, [ph], lambda1=CU_KA1, lambda2=CU_KA2, ratio=CU_RATIO, pol_k=1.0)
    pb.default_values()
    for n, v in zip(('U', 'V', 'W', 'X', 'Y'), (18.03, -9.02, 5.41, 5.0, 2.0)):
        pb.set('p1.' + n, v)
    _cheb(pb, [150.0, -30.0, 10.0, -5.0, 2.0, -1.0])
    for n in ('a',) + (('c',) if system == 'tetragonal' else ('b', 'c') if system == 'monoclinic' else ()):
        v = pb.base[pb.index['p1.' + n]]
        pb.refine('p1.' + n, 0.98 * v, 1.02 * v)
    pb.refine('zero', -2, 2)
    pb.refine('p1.scale', 0, math.inf)
    for at in atoms:
        for f in at.site['free']:
            n = 'p1.%s.%s' % (at.label, 'xyz'[f])
            v = pb.base[pb.index[n]]
            pb.refine(n, v - 0.1, v + 0.1)
        pb.refine('p1.%s.B' % at.label, -2, 30)
    for k in range(2):
        pb.refine('bkg%d' % k)
    pb.chains, pb.sweeps, pb.steps_per_sweep, pb.max_temp, pb.seed = chains, 10, 0, 100.0, 900 + index
    x_true = pb.start_vector()
    if simulate is not None:
        pb.observe(simulate, x_true, seed + 17 * index + 1)
        pb.x0 = list(_perturb(pb, x_true, seed + 19 * index + 2))
    return pb


def uneven_batch(simulate=None, seed=20260303, count=48, chains=32):
    return [uneven_pattern(i, simulate, seed, chains) for i in range(count)]


TRICLINIC_FF = ((2.31, 1.02, 1.5886, .865), (20.8439, 10.2075, .5687, 51.6512), .2156)


SEPARATED_INST = {'U': 8.0, 'V': -4.0, 'W': 4.0, 'X': 1.5, 'Y': 0.5, 'Z': 0.2, 'SL': 0.02, 'HL': 0.01, 'eta': 0.5,
                  'U_L': 8.0, 'V_L': -4.0, 'W_L': 4.0, 'X_L': 1.5, 'U_R': 10.0, 'V_R': -5.0, 'W_R': 5.0, 'X_R': 2.0,
                  'eta_split': 0.6}
SEPARATED_SAMPLE = {'size': 0.12, 'sizeAxial': 0.06, 'sizeB': 0.09, 'sizeC': 0.15, 'rotA': 10.0, 'rotB': -20.0,
                    'rotC': 30.0, 'strain': 900.0, 'strainAxial': 1500.0, 'sizeMix': 0.8, 'strainMix': 0.6,
                    'S400': 190.0, 'S040': 1190.0, 'S004': 430.0, 'S220': 100.0, 'S202': 80.0, 'S022': 150.0}


def separated_probe(size='isotropic', strain='isotropic', family='gsas_cw', simulate=None, seed=20260808, chains=16,
                    size_axis=(0, 0, 1), strain_axis=(1, 0, 0)):
    """The Pnma phase of pbso4_pnma with a separated (GSAS) profile: the dataset's
    instrument function of `family` (gsas_cw = TCH, simple_pvoigt, split_pvoigt; inst
    U 8, V -4, W 4, X 1.5, Y 0.5, Z 0.2, SL .02, HL .01, split flanks 8/-4/4/1.5 and
    10/-5/5/2) plus sample broadening: size 0.12 um (uniaxial 0.06 along size_axis;
    ellipsoid 0.12 x 0.09 x 0.15 rotated 10, -20, 30 deg), strain 900 (uniaxial 1500 along
    strain_axis; Stephens S400 190, S040 1190, S004 430, S220 100, S202 80, S022 150),
    mixes 0.8 / 0.6.  2theta 10-100 step 0.03 with the Cu Ka doublet, 3-term Chebyshev.
    Refined: cell, zero, scale, inst U, W, X, the sample parameters of the models,
    4 coordinates, 2 B, bkg0."""
    rng = np.random.default_rng(seed)
    r, t = S.group_ops(62)
    cell = [8.48, 5.40, 6.96, 90, 90, 90]
    atoms = []
    for (lab, el, xyz, B) in PBSO4_SITES:
        site = S.site_symmetry(r, t, xyz, cell)
        pos = list(site['pos'])
        for f in site['free']:
            pos[f] += rng.uniform(-0.03, 0.03)
        site['pos'] = pos
        atoms.append(Atom(lab, el, site, B=float(rng.uniform(0.5, 2.0))))
    ph = Phase('p1', 62, cell, atoms, 'gsas_cw' if family == 'gsas_cw' else family, separated=True,
               broadening={'size': size, 'strain': strain, 'sizeAxis': list(size_axis), 'strainAxis': list(strain_axis)})
    pb = Problem(_grid(10.0, 100.0, 0.03), [ph], lambda1=CU_KA1, lambda2=CU_KA2, ratio=CU_RATIO, pol_k=1.0, n_bkg=3,
                 resolution_type=family)
    pb.default_values()
    for n, v in SEPARATED_INST.items():
        pb.set('inst.' + n, v)
    for n, v in SEPARATED_SAMPLE.items():
        pb.set('p1.' + n, v)
    pb.set('zero', 0.01)
    _cheb(pb, [150.0, -25.0, 8.0])
    names = ['p1.a', 'p1.b', 'p1.c', 'zero', 'p1.scale']
    names += ['inst.U_L', 'inst.W_R', 'inst.X_L'] if family == 'split_pvoigt' else ['inst.U', 'inst.W', 'inst.X']
    names += ['p1.size', 'p1.sizeMix', 'p1.strainMix']
    names += {'uniaxial': ['p1.sizeAxial'], 'ellipsoid': ['p1.sizeB', 'p1.sizeC', 'p1.rotC']}.get(size, [])
    names += {'isotropic': ['p1.strain'], 'uniaxial': ['p1.strain', 'p1.strainAxial'],
              'stephens': ['p1.S400', 'p1.S040', 'p1.S220']}[strain]
    names += ['p1.%s.%s' % (at.label, 'xyz'[f]) for at in atoms[:2] for f in at.site['free']][:4]
    names += ['p1.%s.B' % at.label for at in atoms[:2]] + ['bkg0']
    for n in names:
        kind, comp = classify(n)
        v = pb.base[pb.index[n]]
        if kind == 'cell':
            pb.refine(n, 0.97 * v, 1.03 * v)
        elif n == 'zero':
            pb.refine(n, -1, 1)
        elif n == 'p1.scale' or comp in ('size', 'sizeAxial', 'sizeB', 'sizeC'):
            pb.refine(n, 0.002 if comp else 0.0, math.inf)
        elif comp in ('sizeMix', 'strainMix'):
            pb.refine(n, 0, 1)
        elif comp in ('strain', 'strainAxial'):
            pb.refine(n, 0, math.inf)
        elif kind == 'atom' and comp == 'B':
            pb.refine(n, -2, 30)
        elif kind == 'atom':
            pb.refine(n, v - 0.1, v + 0.1)
        else:
            pb.refine(n)
    pb.chains, pb.sweeps, pb.steps_per_sweep, pb.max_temp, pb.seed = chains, 20, 0, 100.0, 777
    x_true = pb.start_vector()
    if simulate is not None:
        pb.observe(simulate, x_true, seed + 7)
        pb.x0 = list(_perturb(pb, x_true, seed + 11, frac=0.15))
    return pb


def link_datasets(problems, shared, own, limits=None):
    """Give several datasets one variable list (a multi-dataset PT problem,
    js/rietveld_core.js:2271-2273 with ctx.info[j].ds): first the `shared` parameter names,
    one variable each acting in every dataset, then the names own[d] of dataset d alone,
    labelled name@d.  Start values come from the first dataset holding the name; boxes
    from `limits` (name -> (lo, hi); a cell length defaults to +-3 %), PT scales from the
    reference's natural step."""
    limits = dict(limits or {})
    entries = [(n, None) for n in shared] + [(n, d) for d, names in enumerate(own) for n in names]
    labels, x0, lo, hi, sc = [], [], [], [], []
    for n, d in entries:
        src = problems[0 if d is None else d]
        v = float(src.base[src.index[n]])
        kind, comp = classify(n)
        l, h = limits.get(n, (0.97 * v, 1.03 * v) if kind == 'cell' and comp in ('a', 'b', 'c') else (-math.inf, math.inf))
        labels.append(n if d is None else '%s@%d' % (n, d))
        x0.append(min(h, max(l, v)))
        lo.append(l)
        hi.append(h)
        sc.append(pt_scale(n, v))
    for d, pb in enumerate(problems):
        pb.vars, pb.terms = list(labels), []
        for j, (n, dd) in enumerate(entries):
            if dd is None or dd == d:
                k = pb.index[n]
                pb.base[k] = 0.0
                pb.terms.append((k, j, 1.0))
        pb.x0, pb.lo, pb.hi, pb.scale = list(x0), list(lo), list(hi), list(sc)
    return problems


def xray_neutron_pair(simulate=None, seed=20260909, chains=32):
    """Two datasets of the Pnma phase of pbso4_pnma refined together: Cu Ka1/Ka2 X-ray
    10-100 deg step 0.03 (TCH U..Y 18/-9/5.4/5/2) and constant-wavelength neutron
    1.594 A 10-150 deg step 0.05 (TCH 30/-20/12/3/1, scattering lengths), each with its
    own zero, scale, U, W and 3-term background.  Shared: a, b, c, 4 coordinates, 2 B."""
    rng = np.random.default_rng(seed)
    r, t = S.group_ops(62)
    cell = [8.48, 5.40, 6.96, 90, 90, 90]
    blen = {'Pb': 9.405, 'S': 2.847, 'O': 5.803}
    atoms = []
    for (lab, el, xyz, B) in PBSO4_SITES:
        site = S.site_symmetry(r, t, xyz, cell)
        pos = list(site['pos'])
        for f in site['free']:
            pos[f] += rng.uniform(-0.03, 0.03)
        site['pos'] = pos
        atoms.append(Atom(lab, el, site, B=float(rng.uniform(0.5, 2.0)), b_neutron=blen[el]))
    xr = Problem(_grid(10.0, 100.0, 0.03), [Phase('p1', 62, cell, atoms, 'tch_aniso')], lambda1=CU_KA1, lambda2=CU_KA2,
                 ratio=CU_RATIO, pol_k=1.0, n_bkg=3)
    nt = Problem(_grid(10.0, 150.0, 0.05), [Phase('p1', 62, cell, atoms, 'tch_aniso')], lambda1=1.594, lambda2=0.0,
                 ratio=0.0, pol_k=1.0, n_bkg=3, radiation=1)
    for pb, prof, bk, zero in ((xr, (18.03, -9.02, 5.41, 5.0, 2.0), [150.0, -25.0, 8.0], 0.01),
                               (nt, (30.0, -20.0, 12.0, 3.0, 1.0), [400.0, 60.0, -20.0], -0.02)):
        pb.default_values()
        for n, v in zip(('U', 'V', 'W', 'X', 'Y'), prof):
            pb.set('p1.' + n, v)
        pb.set('zero', zero)
        _cheb(pb, bk)
    coords = ['p1.%s.%s' % (at.label, 'xyz'[f]) for at in atoms[:2] for f in at.site['free']][:4]
    shared = ['p1.a', 'p1.b', 'p1.c'] + coords + ['p1.%s.B' % at.label for at in atoms[:2]]
    own = [['zero', 'p1.scale', 'p1.U', 'p1.W', 'bkg0'] for _ in range(2)]
    limits = {'zero': (-1, 1), 'p1.scale': (0, math.inf)}
    for n in coords:
        v = float(xr.base[xr.index[n]])
        limits[n] = (v - 0.1, v + 0.1)
    for at in atoms[:2]:
        limits['p1.%s.B' % at.label] = (-2, 30)
    link_datasets([xr, nt], shared, own, limits)
    for pb in (xr, nt):
        pb.chains, pb.sweeps, pb.steps_per_sweep, pb.max_temp, pb.seed = chains, 20, 0, 100.0, 9090
    x_true = xr.start_vector()
    if simulate is not None:
        for d, pb in enumerate((xr, nt)):
            pb.x0 = list(x_true)
            pb.observe(simulate, x_true, seed + 7 + d)
            # the scale rescaling touched this dataset's own scale variable in x_true
        x_start = _perturb(xr, x_true, seed + 11, frac=0.15)
        for pb in (xr, nt):
            pb.x0 = list(x_start)
    return [xr, nt]


def lebail_probe(mixed=False, simulate=None, seed=20261010, chains=16, profile='tch_aniso', method='lebail'):
    """Free-intensity (Le Bail) refinement: NaCl Fm-3m a=5.64 over 10-120 deg step 0.02,
    Cu Ka1/Ka2, TCH, 3-term Chebyshev; the observed pattern is the Rietveld model (Na 4a,
    Cl 4b) with Poisson noise, the fitted phase a Le Bail phase whose intensities are free.
    With mixed, a second phase ZnS F-43m a=5.41 (Zn 4a, S 4c) joins as a Rietveld phase.
    Refined: a, zero, U, V, W, X, Y, bkg0, bkg1 (and the ZnS scale and a)."""
    r, t = S.group_ops(225)
    cell = [5.64, 5.64, 5.64, 90, 90, 90]
    atoms = [Atom('Na1', 'Na', S.site_symmetry(r, t, [0, 0, 0], cell), B=1.1),
             Atom('Cl1', 'Cl', S.site_symmetry(r, t, [0.5, 0.5, 0.5], cell), B=0.9)]
    phases = [Phase('p1', 225, cell, atoms, profile)]
    if mixed:
        r2, t2 = S.group_ops(216)
        c2 = [5.41, 5.41, 5.41, 90, 90, 90]
        a2 = [Atom('Zn1', 'Zn', S.site_symmetry(r2, t2, [0, 0, 0], c2), B=0.8),
              Atom('S1', 'S', S.site_symmetry(r2, t2, [0.25, 0.25, 0.25], c2), B=0.7)]
        phases.append(Phase('p2', 216, c2, a2, 'tch_aniso'))
    pb = Problem(_grid(10.0, 120.0, 0.02), phases, lambda1=CU_KA1, lambda2=CU_KA2, ratio=CU_RATIO, pol_k=1.0, n_bkg=3)
    pb.default_values()
    widths = {'tch_aniso': ('U', 'V', 'W', 'X', 'Y'), 'simple_pvoigt': ('GU', 'GV', 'GW', 'GP', 'LX'),
              'split_pvoigt': ('GU_L', 'GV_L', 'GW_L', 'GU_R', 'LX_L'), 'gsas_cw': ('inst.U', 'inst.V', 'inst.W', 'inst.X', 'inst.Y')}
    for ph in phases:
        names_w = widths[ph.profile]
        for n, v in zip(names_w, (18.03, -9.02, 5.41, 5.0 if ph.profile != 'simple_pvoigt' else 0.0, 2.0)):
            pb.set(n if n.startswith('inst.') else ph.key + '.' + n, v)
    if profile == 'gsas_cw':
        pb.set('p1.size', 0.5)
    if mixed:
        pb.set('p2.scale', 0.6)
    pb.set('zero', 0.012)
    _cheb(pb, [150.0, -30.0, 10.0])
    w1 = [n if n.startswith('inst.') else 'p1.' + n for n in widths[profile]]
    names = ['p1.a', 'zero'] + w1 + ['bkg0', 'bkg1'] + (['p2.a', 'p2.scale'] if mixed else [])
    limits = {'p1.a': (5.64 * 0.98, 5.64 * 1.02), 'p2.a': (5.41 * 0.98, 5.41 * 1.02), 'zero': (-1, 1), 'p2.scale': (0, math.inf),
              'p1.X': (0, math.inf), 'p1.Y': (0, math.inf), 'p1.LX_L': (0, math.inf), 'p1.GP': (-1, math.inf),
              'inst.X': (0, math.inf), 'inst.Y': (0, math.inf)}
    for n in names:
        lo, hi = limits.get(n, (-math.inf, math.inf))
        pb.refine(n, lo, hi)
    pb.chains, pb.sweeps, pb.steps_per_sweep, pb.max_temp, pb.seed = chains, 20, 0, 100.0, 1010
    x_true = pb.start_vector()
    if simulate is not None:
        pb.observe(simulate, x_true, seed)
        pb.x0 = list(_perturb(pb, x_true, seed + 1, frac=0.15))
    phases[0].method = method            # the observation came from its atoms; the fit frees them
    return pb


def triclinic_probe(profile='tch_aniso', simulate=None, neutron=False, sl=0.0, hl=0.0, po=0.87, n_points=160,
                    method='rietveld', broadening=None, lambda2=1.544, inst=None, span=None):
    """The one-atom P-1 problem of tests/webgpu-pt.js:11-36: a=3.7 b=4.1 c=4.3 alpha=87
    beta=94 gamma=102, C at (.13,.21,.32) with the 4-Gaussian f0 given there; 160 points
    8-63.65 deg (step 0.35), w 0.01 (point 5 excluded), lambda 1.540598/1.544 ratio .5,
    polK .83, minFwhm .07; zero .03, disp 1.5, transp 1e-5, bkg 13 + 2 T1, scale .012,
    po .87.  Refined: a, b, c (+-4%), zero, disp, transp, bkg0, scale, po, C1 x,y,z,B,occ
    and the family's width parameters."""
    ops = S.group_ops(2)
    cell = [3.7, 4.1, 4.3, 87, 94, 102]
    at = Atom('C1', 'C', S.site_symmetry(ops[0], ops[1], [.13, .21, .32], cell), B=0.5, ff=TRICLINIC_FF,
              b_neutron=6.646)
    sep = profile == 'gsas_cw' or broadening is not None
    ph = Phase('p1', 2, cell, [at], profile, po_dir=(0, 0, 1), separated=sep, broadening=broadening)
    step = .35 if span is None else span / (n_points - 1)
    tth = 8 + np.arange(n_points, dtype=np.float64) * step
    pb = Problem(tth, [ph], lambda1=1.540598, lambda2=lambda2, ratio=.5, pol_k=.83, min_fwhm=.07, n_bkg=2,
                 radiation=1 if neutron else 0)
    pb.default_values()
    for n, v in (inst or {}).items():
        pb.set('inst.' + n, v)
    pb.set('zero', .03)
    pb.set('disp', 1.5)
    pb.set('transp', .00001)
    pb.set('bkg0', 13)
    pb.set('bkg1', 2)
    pb.set('p1.scale', .012)
    pb.set('p1.po', po)
    if profile == 'tch_aniso':
        pb.set('p1.SL', sl)
        pb.set('p1.HL', hl)
    names = ['p1.a', 'p1.b', 'p1.c', 'zero', 'disp', 'transp', 'bkg0', 'p1.scale', 'p1.po', 'p1.C1.x', 'p1.C1.y',
             'p1.C1.z', 'p1.C1.B', 'p1.C1.occ']
    if sep:
        names += ['inst.W', 'inst.X', 'p1.size']
    else:
        for n in ['GU', 'LX', 'eta', 'GU_L', 'GU_R', 'LX_L', 'LX_R', 'eta_split', 'U', 'W', 'X', 'Y', 'SL', 'HL']:
            if 'p1.' + n in pb.index and n in PROFILE_SLOTS[profile]:
                names.append('p1.' + n)
    if method != 'rietveld':
        # a Le Bail or WPPF phase has no scale, orientation or atoms (rietveld_core.js:540-545)
        names = [n for n in names if n not in ('p1.scale', 'p1.po') and classify(n)[0] != 'atom']
    for n in names:
        kind, comp = classify(n)
        if kind == 'cell':
            v = pb.base[pb.index[n]]
            pb.refine(n, .96 * v, 1.04 * v)
        elif n == 'zero':
            pb.refine(n, -2, 2)
        elif n == 'p1.scale':
            pb.refine(n, 0, math.inf)
        elif n == 'p1.po':
            pb.refine(n, 0.1, 5)
        elif kind == 'atom' and comp == 'occ':
            pb.refine(n, 0, 2)
        elif kind == 'atom' and comp == 'B':
            pb.refine(n, -2, 30)
        elif kind == 'profile' and comp.startswith('eta'):
            pb.refine(n, 0, 1)
        elif kind == 'profile' and comp in ('LX', 'LX_L', 'LX_R', 'X', 'Y', 'GP'):
            pb.refine(n, 0, math.inf)
        elif n == 'p1.size':
            pb.refine(n, 0.002, math.inf)
        else:
            pb.refine(n)
    pb.chains, pb.sweeps, pb.steps_per_sweep, pb.max_temp, pb.seed = 4, 10, 0, 5.0, 137
    if simulate is not None:
        ycalc = np.asarray(simulate(pb, pb.start_vector()), dtype=np.float64)
        i = np.arange(n_points)
        pb.y = ycalc + 2 * np.sin(i * .13) + 1
        pb.w = np.full(n_points, .01)
        pb.w[5] = 0
    ph.method = method                   # observed from the atoms; a free phase then frees them
    return pb


# the sample values the reference's WebGPU test page gives a separated phase
# (tests/webgpu-pt.js:23)
PAGE_SAMPLE = {'size': .2, 'sizeAxial': .4, 'sizeB': .3, 'sizeC': .45, 'rotA': 13, 'rotB': 21, 'rotC': 33,
               'strain': 150, 'strainAxial': 230, 'S400': 4, 'S040': 5, 'S004': 6, 'sizeMix': .6, 'strainMix': .4}
PAGE_CELL = [3.7, 4.1, 4.3, 87, 94, 102]


def _page_dataset(phases, neutron=False, lambda2=1.544, ratio=.5, resolution_type=None):
    """A dataset of the page (tests/webgpu-pt.js:25): 160 points 8-63.65 deg step 0.35,
    lambda 1.540598, polK .83, minFwhm .07, 2 background terms, w = .01 with point 5 out,
    and its values: zero .03, disp 1.5, transp 1e-5, bkg 13 + 2 T1, every phase's scale
    .012, po .87 and the page's sample values."""
    tth = 8 + np.arange(160, dtype=np.float64) * .35
    pb = Problem(tth, phases, lambda1=1.540598, lambda2=lambda2, ratio=ratio, pol_k=.83, min_fwhm=.07, n_bkg=2,
                 radiation=1 if neutron else 0, resolution_type=resolution_type)
    pb.default_values()
    for n, v in (('zero', .03), ('disp', 1.5), ('transp', 1e-5), ('bkg0', 13.0), ('bkg1', 2.0)):
        pb.set(n, v)
    for ph in phases:
        pb.set(ph.key + '.scale', .012)
        pb.set(ph.key + '.po', .87)
        for n, v in PAGE_SAMPLE.items():
            pb.set(ph.key + '.' + n, v)
    pb.w = np.full(len(tth), .01)
    pb.w[5] = 0
    return pb


def _page_atoms(two=False):
    ops = S.group_ops(2)
    atoms = [Atom('C1', 'C', S.site_symmetry(ops[0], ops[1], [.13, .21, .32], PAGE_CELL), B=.5, ff=TRICLINIC_FF,
                  b_neutron=6.646, occ=.4 if two else 1.0)]
    if two:
        atoms.append(Atom('C2', 'C', S.site_symmetry(ops[0], ops[1], [.26, .33, .41], PAGE_CELL), B=.5, ff=TRICLINIC_FF,
                          b_neutron=6.646, occ=.6))
    return atoms


def _page_observe(pbs, simulate, amp=2.0):
    """y = model at the start + amp sin(0.13 j) + 1 (tests/webgpu-pt.js:95, 104)."""
    for pb in pbs:
        y = np.asarray(simulate(pb, pb.start_vector()), dtype=np.float64)
        j = np.arange(len(y))
        pb.y = y + amp * np.sin(j * .13) + 1.0


def page_joint_case(simulate=None):
    """The page's joint case (tests/webgpu-pt.js:83-99): the gsas_cw P-1 probe with a second
    atom C2 at (.26, .33, .41), and
      * atomic constraints (js/atomic_constraints.js:76-155): C1.B held; C2.x = 0.5 C1.x
        (equivalence, mult 0.5); C1.occ + C2.occ = 1 (constraint equation, C2.occ the
        dependent, as the last refined term);
      * a WPPF phase p2 (a copy of p1, free intensities) beside the Rietveld phase p1 in
        the X-ray dataset;
      * a second dataset, neutron, without the doublet and without p2 (a phase absent from
        one dataset), with its own bkg0 = 18 and zero = -0.02.
    Variables: p1 cell, disp, transp, scale, po, C1 x y z occ, inst.W, inst.X, size shared by
    both datasets; zero and bkg0 per dataset.  The observations are each dataset's model at
    the start (p2 from its atoms) + 2 sin(0.13 j) + 1."""
    p1 = Phase('p1', 2, PAGE_CELL, _page_atoms(True), 'gsas_cw', po_dir=(0, 0, 1), separated=True)
    p2 = Phase('p2', 2, PAGE_CELL, _page_atoms(True), 'gsas_cw', po_dir=(0, 0, 1), separated=True)
    xr = _page_dataset([p1, p2])
    nt = _page_dataset([Phase('p1', 2, PAGE_CELL, _page_atoms(True), 'gsas_cw', po_dir=(0, 0, 1), separated=True)],
                       neutron=True, lambda2=0.0, ratio=0.0)
    for pb in (xr, nt):
        pb.set('p1.C2.x', .065)
        pb.set('p1.C1.occ', .4)
        pb.set('p1.C2.occ', .6)
    for n in ('C1.x', 'C1.y', 'C1.z', 'C1.occ', 'C1.B', 'C2.x', 'C2.y', 'C2.z', 'C2.occ', 'C2.B'):
        xr.set('p2.' + n, xr.base[xr.index['p1.' + n]])
    nt.set('bkg0', 18.0)
    nt.set('zero', -.02)
    shared = ['p1.a', 'p1.b', 'p1.c', 'disp', 'transp', 'p1.scale', 'p1.po', 'p1.C1.x', 'p1.C1.y', 'p1.C1.z',
              'p1.C1.occ', 'inst.W', 'inst.X', 'p1.size']
    own = [['zero', 'bkg0'], ['zero', 'bkg0']]
    limits = {'zero': (-2, 2), 'p1.scale': (0, math.inf), 'p1.po': (.1, 5), 'p1.C1.occ': (0, 1), 'inst.X': (0, math.inf),
              'p1.size': (.002, math.inf)}
    for n in ('a', 'b', 'c'):
        v = xr.base[xr.index['p1.' + n]]
        limits['p1.' + n] = (.96 * v, 1.04 * v)
    for n in ('x', 'y', 'z'):
        v = xr.base[xr.index['p1.C1.' + n]]
        limits['p1.C1.' + n] = (v - .1, v + .1)
    link_datasets([xr, nt], shared, own, limits)
    jx, jo = shared.index('p1.C1.x'), shared.index('p1.C1.occ')
    for pb in (xr, nt):
        pb.constrain('p1.C2.x', 0.0, [(jx, 0.5)])          # equivalence, mult [1, .5]
        pb.constrain('p1.C2.occ', 1.0, [(jo, -1.0)])       # C1.occ + C2.occ = 1
        pb.chains, pb.sweeps, pb.steps_per_sweep, pb.max_temp, pb.seed = 4, 10, 0, 5.0, 137
    if simulate is not None:
        _page_observe([xr, nt], simulate)
    p2.method = 'wppf'                       # observed from its atoms; the fit frees its intensities
    return [xr, nt]


def page_scoped_case(simulate=None):
    """The page's scoped case (tests/webgpu-pt.js:100-106): two copies of the gsas_cw P-1
    dataset, the first with the dataset resolution function simple_pvoigt and the second
    split_pvoigt, every instrument value at its default (js/rietveld_core.js:459
    ALL_INSTRUMENT_PROFILE); variables bkg0, zero and C1.B per dataset, and the
    cross-dataset link C1.B(2) = 1.2 C1.B(1) + 0.1 (parameterLinks), so C1.B of the second
    dataset is not a variable.  Observations: model + 1 + sin(0.13 j)."""
    pbs = []
    for res in ('simple_pvoigt', 'split_pvoigt'):
        ph = Phase('p1', 2, PAGE_CELL, _page_atoms(False), 'gsas_cw', po_dir=(0, 0, 1), separated=True)
        pb = _page_dataset([ph], resolution_type=res)
        for n, v in zip(INST_NAMES, INST_DEFAULTS):
            pb.set('inst.' + n, v)
        pbs.append(pb)
    pbs[1].set('p1.C1.B', .7)
    link_datasets(pbs, [], [['bkg0', 'zero', 'p1.C1.B'], ['bkg0', 'zero']],
                  {'zero': (-2, 2), 'p1.C1.B': (-2, 30)})
    jb = pbs[0].vars.index('p1.C1.B@0')
    pbs[1].constrain('p1.C1.B', 0.1, [(jb, 1.2)])
    for pb in pbs:
        pb.chains, pb.sweeps, pb.steps_per_sweep, pb.max_temp, pb.seed = 4, 10, 0, 5.0, 137
    if simulate is not None:
        _page_observe(pbs, simulate, amp=1.0)
    return pbs
