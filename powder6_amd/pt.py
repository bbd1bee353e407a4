"""Python host of the MI355X PT trial evaluator: ctypes binding of include/powder6_pt.h
and the RVPTGPU-shaped session API of js/pt_gpu.js (`compile`, `create`, `run`).

The library is powder6_amd/_lib/libpowder6_pt.so (built by powder6_amd/build.py).  There
is no fallback: without the library, or without a gfx950 device, the calls raise
PTError with the library's message (the reference throws and its caller falls back to
CPU PT, js/rietveld_core.js:2456-2460).
"""
import ctypes
import math
import os

import numpy as np

from . import problems as PR

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, '_lib', 'libpowder6_pt.so')

c_int32, c_double, c_uint32 = ctypes.c_int32, ctypes.c_double, ctypes.c_uint32
P_D = ctypes.POINTER(c_double)
P_I = ctypes.POINTER(c_int32)

FORM_PLAIN, FORM_BATCHED, FORM_MATRIX, FORM_MATRIX_BF16 = 0, 1, 2, 3
S_COST, S_BEST, S_STEP, S_ACCEPTED, S_PROPOSALS, S_SWAPS, S_ATTEMPTS, S_VALID, S_LOGU, S_UNSOLVED, S_TRIAL, S_FEASIBLE = range(12)


class PTError(RuntimeError):
    pass


class PtAtom(ctypes.Structure):
    _fields_ = [('r0', c_double * 3), ('basis', (c_double * 3) * 3), ('free_mask', c_int32), ('n_gauss', c_int32),
                ('mult', c_double), ('ff_a', c_double * 5), ('ff_b', c_double * 5), ('ff_c', c_double),
                ('b_neutron', c_double), ('fp', c_double), ('fpp', c_double)]


class PtPhase(ctypes.Structure):
    _fields_ = [('system', c_int32), ('unique_axis', c_int32), ('profile_type', c_int32), ('n_ops', c_int32),
                ('op_r', P_I), ('op_t', P_D), ('po_dir', c_double * 3), ('n_atoms', c_int32),
                ('atoms', ctypes.POINTER(PtAtom)), ('separated', c_int32), ('size_model', c_int32),
                ('strain_model', c_int32), ('size_axis', c_double * 3), ('strain_axis', c_double * 3),
                ('free_intensities', c_int32)]


class PtProblem(ctypes.Structure):
    _fields_ = [('n_points', c_int32), ('tth', P_D), ('y_obs', P_D), ('weight', P_D),
                ('lambda1', c_double), ('lambda2', c_double), ('ratio', c_double), ('pol_k', c_double),
                ('min_fwhm', c_double), ('radiation', c_int32), ('n_bkg', c_int32),
                ('n_phases', c_int32), ('phases', ctypes.POINTER(PtPhase)),
                ('n_params', c_int32), ('param_base', P_D),
                ('n_vars', c_int32), ('x0', P_D), ('lo', P_D), ('hi', P_D), ('scale', P_D),
                ('n_terms', c_int32), ('term_param', P_I), ('term_var', P_I), ('term_coeff', P_D),
                ('chains', c_int32), ('sweeps', c_int32), ('steps_per_sweep', c_int32),
                ('max_temp', c_double), ('seed', c_uint32), ('bkg_type', c_int32), ('bkg_x', P_D),
                ('resolution_type', c_int32)]


class PtPlanInfo(ctypes.Structure):
    _fields_ = [(n, c_int32) for n in ('nv', 'chains', 'stride', 'point_count', 'reflection_count', 'line_count',
                                       'steps_per_sweep', 'n_params', 'uniform_grid', 'device_ok')]


def _dptr(a):
    return a.ctypes.data_as(P_D)


def _iptr(a):
    return a.ctypes.data_as(P_I)


class CProblem:
    """The C struct of a problems.Problem, with every array it points to kept alive."""

    def __init__(self, pb):
        keep = self._keep = []

        def d(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return _dptr(a)

        def i(a):
            a = np.ascontiguousarray(a, dtype=np.int32)
            keep.append(a)
            return _iptr(a)

        phases = (PtPhase * len(pb.phases))()
        for q, ph in enumerate(pb.phases):
            atoms = (PtAtom * max(1, len(ph.atoms)))()
            for j, at in enumerate(ph.atoms):
                A = atoms[j]
                for x in range(3):
                    A.r0[x] = float(at.site['r0'][x])
                    for y in range(3):
                        A.basis[x][y] = float(at.site['basis'][x][y])
                A.free_mask = sum(1 << f for f in at.site['free'])
                A.n_gauss = at.n_gauss
                A.mult = float(at.site['mult'])
                for g in range(at.n_gauss):
                    A.ff_a[g] = at.ff_a[g]
                    A.ff_b[g] = at.ff_b[g]
                A.ff_c = at.ff_c
                A.b_neutron, A.fp, A.fpp = at.b_neutron, at.fp, at.fpp
            keep.append(atoms)
            P = phases[q]
            P.system = PR.SYSTEMS[ph.system]
            P.unique_axis = ph.unique_axis
            P.profile_type = PR.PROFILE[ph.profile]
            P.n_ops = len(ph.op_r)
            P.op_r = i(np.asarray(ph.op_r).reshape(-1))
            P.op_t = d(np.asarray(ph.op_t).reshape(-1))
            for x in range(3):
                P.po_dir[x] = ph.po_dir[x]
            P.n_atoms = len(ph.atoms)
            P.atoms = ctypes.cast(atoms, ctypes.POINTER(PtAtom))
            P.separated = 1 if getattr(ph, 'separated', False) else 0
            P.free_intensities = 0 if getattr(ph, 'method', 'rietveld') == 'rietveld' else 1
            P.size_model = PR.SIZE_MODELS[getattr(ph, 'size_model', 'isotropic')]
            P.strain_model = PR.STRAIN_MODELS[getattr(ph, 'strain_model', 'isotropic')]
            for x in range(3):
                P.size_axis[x] = float(getattr(ph, 'size_axis', [0, 0, 1])[x])
                P.strain_axis[x] = float(getattr(ph, 'strain_axis', [0, 0, 1])[x])
        keep.append(phases)
        s = self.struct = PtProblem()
        s.n_points = len(pb.tth)
        s.tth, s.y_obs, s.weight = d(pb.tth), d(pb.y), d(pb.w)
        s.lambda1, s.lambda2, s.ratio, s.pol_k, s.min_fwhm = pb.lambda1, pb.lambda2, pb.ratio, pb.pol_k, pb.min_fwhm
        s.radiation, s.n_bkg = pb.radiation, pb.n_bkg
        s.n_phases = len(pb.phases)
        s.phases = ctypes.cast(phases, ctypes.POINTER(PtPhase))
        s.n_params = len(pb.base)
        s.param_base = d(pb.base)
        s.n_vars = pb.nv
        s.x0, s.lo, s.hi, s.scale = d(pb.x0), d(pb.lo), d(pb.hi), d(pb.scale)
        s.n_terms = len(pb.terms)
        s.term_param = i([t[0] for t in pb.terms] or [0])
        s.term_var = i([t[1] for t in pb.terms] or [0])
        s.term_coeff = d([t[2] for t in pb.terms] or [0.0])
        s.chains, s.sweeps, s.steps_per_sweep = pb.chains, pb.sweeps, pb.steps_per_sweep
        s.max_temp, s.seed = pb.max_temp, pb.seed
        s.bkg_type = getattr(pb, 'bkg_type', 0)
        bx = getattr(pb, 'bkg_x', None)
        s.bkg_x = d(np.asarray(bx if bx is not None else [0.0], dtype=np.float64))
        rt = getattr(pb, 'resolution_type', None)
        s.resolution_type = PR.PROFILE[rt] if rt else 0

    def ref(self):
        return ctypes.byref(self.struct)


_lib = None


def lib():
    """Load the HIP library (raises PTError when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PTError('MI355X PT library is not built: %s (run powder6_amd/build.py)' % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.pt_last_error.restype = ctypes.c_char_p
        L.pt_plan_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
        L.pt_plan_compile.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
        L.pt_plan_create_multi.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.pt_plan_dataset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.pt_plan_destroy.argtypes = [ctypes.c_void_p]
        L.pt_plan_destroy.restype = None
        L.pt_plan_info_get.argtypes = [ctypes.c_void_p, ctypes.POINTER(PtPlanInfo)]
        L.pt_plan_reflections.argtypes = [ctypes.c_void_p, ctypes.c_int, P_I, ctypes.c_int]
        L.pt_set_states.argtypes = [ctypes.c_void_p, P_D, ctypes.c_int]
        L.pt_costs.argtypes = [ctypes.c_void_p, P_D, ctypes.c_int, P_D]
        L.pt_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, P_D]
        L.pt_eval.argtypes = [ctypes.c_void_p, ctypes.c_int, P_D, ctypes.c_int, P_D]
        L.pt_pattern.argtypes = [ctypes.c_void_p, P_D, P_D]
        L.pt_last_timing.argtypes = [ctypes.c_void_p, P_D, P_D, P_I]
        L.pt_phase_base.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.pt_workload.argtypes = [ctypes.c_void_p, P_D, P_D]
        L.pt_normal_equations.argtypes = [ctypes.c_void_p, P_D, P_D, P_D, P_D, P_D]
        L.pt_set_kernel_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.pt_set_solve_limit.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.pt_trial_counts.argtypes = [ctypes.c_void_p, P_D]
        L.pt_matrix_pattern.argtypes = [ctypes.c_void_p, P_D, P_D]
        L.pt_matrix_timing.argtypes = [ctypes.c_void_p, P_D, P_D]
        L.pt_tiles.argtypes = [ctypes.c_void_p, ctypes.c_int, P_I, P_D]
        L.pt_hbm_copy_probe.argtypes = [ctypes.c_size_t, ctypes.c_int, P_D]
        P_F = ctypes.POINTER(ctypes.c_float)
        L.pt_screen.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.pt_screen_stats.argtypes = [ctypes.c_void_p, P_D]
        L.pt_screen_predict.argtypes = [ctypes.c_void_p, P_D, ctypes.c_int, P_F]
        L.pt_mlp_create.argtypes = [ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)]
        L.pt_mlp_destroy.argtypes = [ctypes.c_void_p]
        L.pt_mlp_destroy.restype = None
        L.pt_mlp_param_count.argtypes = []
        L.pt_mlp_get_params.argtypes = [ctypes.c_void_p, P_F]
        L.pt_mlp_set_params.argtypes = [ctypes.c_void_p, P_F]
        L.pt_mlp_forward.argtypes = [ctypes.c_void_p, P_F, ctypes.c_int, ctypes.c_int, P_F]
        L.pt_mlp_gradient.argtypes = [ctypes.c_void_p, P_F, P_F, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, P_F, P_F]
        L.pt_mlp_train.argtypes = [ctypes.c_void_p, P_F, P_F, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_float, ctypes.c_uint32, P_F]
        L.pt_group_create.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.pt_group_destroy.argtypes = [ctypes.c_void_p]
        L.pt_group_destroy.restype = None
        L.pt_group_member.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(PtPlanInfo),
                                      ctypes.POINTER(ctypes.c_int64)]
        L.pt_group_set_states.argtypes = [ctypes.c_void_p, ctypes.c_int, P_D, ctypes.c_int]
        L.pt_group_costs.argtypes = [ctypes.c_void_p, P_D]
        L.pt_group_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, P_D]
        L.pt_group_tiles.argtypes = [ctypes.c_void_p, ctypes.c_int, P_I, P_D]
        L.pt_group_last_timing.argtypes = [ctypes.c_void_p, P_D]
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise PTError(lib().pt_last_error().decode())
    return rc


class Plan:
    """A compiled PT plan (pt_plan_create, or pt_plan_compile with device=False).  A list
    of problems (datasets sharing one variable list) makes a multi-dataset plan
    (pt_plan_create_multi): its costs and PT run on the sum of the datasets' chi^2."""

    def __init__(self, pb, device=True):
        self.problem = pb
        h = ctypes.c_void_p()
        L = lib()
        if isinstance(pb, (list, tuple)):
            if not device:
                raise PTError('a multi-dataset plan needs the device')
            self.cps = [CProblem(q) for q in pb]
            arr = (PtProblem * len(pb))()
            for i, cp in enumerate(self.cps):
                arr[i] = cp.struct
            self._arr = arr
            _check(L.pt_plan_create_multi(ctypes.cast(arr, ctypes.c_void_p), len(pb), ctypes.byref(h)))
            self.cp = self.cps[0]
        else:
            self.cp = CProblem(pb)
            _check((L.pt_plan_create if device else L.pt_plan_compile)(self.cp.ref(), ctypes.byref(h)))
        self.h = h
        info = PtPlanInfo()
        _check(L.pt_plan_info_get(self.h, ctypes.byref(info)))
        self.info = {f: getattr(info, f) for f, _ in PtPlanInfo._fields_}
        self.nv, self.chains, self.stride = info.nv, info.chains, info.stride

    def close(self):
        if getattr(self, 'h', None):
            lib().pt_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def reflections(self, phase=0):
        n = _check(lib().pt_plan_reflections(self.h, phase, None, 0))
        out = np.zeros((max(n, 1), 4), dtype=np.int32)
        _check(lib().pt_plan_reflections(self.h, phase, _iptr(out), n))
        return out[:n]

    def _vectors(self, x):
        x = np.ascontiguousarray(np.atleast_2d(np.asarray(x, dtype=np.float64)))
        if x.shape[1] != self.nv:
            raise PTError('trial vectors must have %d entries' % self.nv)
        return x

    def initial(self, vectors):
        x = self._vectors(vectors)
        _check(lib().pt_set_states(self.h, _dptr(x), x.shape[0]))

    def costs(self, vectors):
        x = self._vectors(vectors)
        out = np.zeros(self.chains)
        _check(lib().pt_costs(self.h, _dptr(x), x.shape[0], _dptr(out)))
        return out

    def batch(self, count, bootstrap=False):
        """batch(count, bootstrap): the state snapshot.  A PTError from an unsolved chain
        carries the snapshot as its `state`."""
        out = np.zeros(self.chains * self.stride)
        try:
            _check(lib().pt_batch(self.h, int(count), 1 if bootstrap else 0, _dptr(out)))
        except PTError as e:
            e.state = out
            raise
        return out

    def trial_counts(self):
        """Proposals since initial(): (evaluated exactly, rejected without an evaluation)."""
        out = np.zeros(2)
        _check(lib().pt_trial_counts(self.h, _dptr(out)))
        return int(out[0]), int(out[1])

    def solve_limit(self, max_iterations):
        """Active-set iterations per free-intensity solve (0: the default 3 m + 10)."""
        _check(lib().pt_set_solve_limit(self.h, int(max_iterations)))

    def sweeps(self, count):
        """`count` sweeps without the snapshot copy (timing)."""
        _check(lib().pt_batch(self.h, int(count), 0, None))

    def eval(self, x, form=FORM_BATCHED):
        x = self._vectors(x)
        out = np.zeros(x.shape[0])
        _check(lib().pt_eval(self.h, form, _dptr(x), x.shape[0], _dptr(out)))
        return out

    def pattern(self, x):
        x = self._vectors(x)
        out = np.zeros(self.info['point_count'])
        _check(lib().pt_pattern(self.h, _dptr(x[0]), _dptr(out)))
        return out

    def dataset_pattern(self, d, x):
        """Calculated pattern of dataset d of a multi-dataset plan at one vector."""
        x = self._vectors(x)
        sub = ctypes.c_void_p()
        _check(lib().pt_plan_dataset(self.h, int(d), ctypes.byref(sub)))
        info = PtPlanInfo()
        _check(lib().pt_plan_info_get(sub, ctypes.byref(info)))
        out = np.zeros(info.point_count)
        _check(lib().pt_pattern(sub, _dptr(x[0]), _dptr(out)))
        return out

    def normal(self, x, columns=False):
        """normal() of one LM step at x (js/rietveld_core.js:1953-1983): (A, g, chi2), and
        with columns=True also J[j, i] = d ycalc_i / d x_j over the points of every dataset."""
        x = self._vectors(x)
        probs = self.problem if isinstance(self.problem, (list, tuple)) else [self.problem]
        n = sum(len(q.tth) for q in probs)
        A, g, f = np.zeros((self.nv, self.nv)), np.zeros(self.nv), c_double()
        J = np.zeros((self.nv, n)) if columns else None
        _check(lib().pt_normal_equations(self.h, _dptr(x[0]), _dptr(A), _dptr(g), ctypes.byref(f),
                                         _dptr(J) if columns else None))
        return (A, g, f.value, J) if columns else (A, g, f.value)

    def workload(self, x):
        """Per-trial work at x: lines, W_tot, N_hkl, S, NP, N_T."""
        x = self._vectors(x)
        out = np.zeros(6)
        _check(lib().pt_workload(self.h, _dptr(x[0]), _dptr(out)))
        return dict(lines=int(out[0]), w_tot=int(out[1]), n_hkl=int(out[2]), s_terms=int(out[3]), np=int(out[4]),
                    n_bkg=int(out[5]))

    def matrix_pattern(self, x):
        """Pattern of one structure-only vector through the shared profile matrix."""
        x = self._vectors(x)
        out = np.zeros(self.info['point_count'])
        _check(lib().pt_matrix_pattern(self.h, _dptr(x[0]), _dptr(out)))
        return out

    def matrix_timing(self):
        a, b = c_double(), c_double()
        _check(lib().pt_matrix_timing(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_kernel_timing(self, on):
        """Direct launches with HIP events around the dominant kernel (no sweep graph)."""
        _check(lib().pt_set_kernel_timing(self.h, 1 if on else 0))

    def timing(self):
        t, k, n = c_double(), c_double(), c_int32()
        _check(lib().pt_last_timing(self.h, ctypes.byref(t), ctypes.byref(k), ctypes.byref(n)))
        return t.value, k.value, n.value

    def screen(self, **cfg):
        """Surrogate-screened PT (pt_screen) with SCREEN_DEFAULTS overridden by cfg;
        screen(off=True) returns to plain PT."""
        if cfg.pop('off', False):
            _check(lib().pt_screen(self.h, None))
            return
        c = dict(SCREEN_DEFAULTS, **cfg)
        self._screen_cfg = PtScreenConfig(**c)
        _check(lib().pt_screen(self.h, ctypes.byref(self._screen_cfg)))

    def screen_predict(self, x):
        """log Rwp of trial vectors as the plan's surrogate network predicts it."""
        x = self._vectors(x)
        out = np.zeros(x.shape[0], dtype=np.float32)
        _check(lib().pt_screen_predict(self.h, _dptr(x), x.shape[0], _fptr(out)))
        return out

    def screen_stats(self):
        out = np.zeros(5)
        _check(lib().pt_screen_stats(self.h, _dptr(out)))
        return dict(iterations=int(out[0]), ring=int(out[1]), loss=float(out[2]), picked=int(out[3]),
                    skipped=int(out[4]))

    def tiles(self, request=0):
        """Tile shape of the profile sum: request 0 reports, -1 retunes, 1/2/4 fixes it.
        Returns (128-point tiles per wavefront in use, tuner times in ms for 1, 2, 4)."""
        t, ms = c_int32(), np.zeros(3)
        _check(lib().pt_tiles(self.h, int(request), ctypes.byref(t), _dptr(ms)))
        return t.value, ms


def hbm_copy_probe(nbytes=1 << 30, repeats=3):
    """GB/s (read + write) of a plain streaming copy of nbytes (pt_hbm_copy_probe)."""
    g = c_double()
    _check(lib().pt_hbm_copy_probe(int(nbytes), int(repeats), ctypes.byref(g)))
    return g.value


class PtScreenConfig(ctypes.Structure):
    _fields_ = [('oversample', ctypes.c_int), ('train_iters', ctypes.c_int), ('batch', ctypes.c_int),
                ('warmup_iters', ctypes.c_int), ('capacity', ctypes.c_int), ('lr', ctypes.c_float),
                ('seed', ctypes.c_uint32), ('margin', ctypes.c_float), ('cold_fraction', ctypes.c_float),
                ('skip_margin', ctypes.c_float), ('skip_eps', ctypes.c_float), ('train_every', ctypes.c_int)]


SCREEN_DEFAULTS = dict(oversample=4, train_iters=1, batch=512, warmup_iters=200, capacity=65536, lr=1e-3, seed=7,
                       margin=0.0, cold_fraction=1.0, skip_margin=0.0, skip_eps=0.0, train_every=1)


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Mlp:
    """The surrogate network alone (pt_mlp_*): 64 -> 256 -> 256 -> 256 -> 1 SiLU,
    BF16 products, FP32 master weights."""

    INPUTS, HIDDEN = 64, 256

    def __init__(self, seed=1):
        h = ctypes.c_void_p()
        _check(lib().pt_mlp_create(seed, ctypes.byref(h)))
        self.h = h
        self.n_params = lib().pt_mlp_param_count()

    def close(self):
        if getattr(self, 'h', None):
            lib().pt_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def params(self):
        out = np.zeros(self.n_params, dtype=np.float32)
        _check(lib().pt_mlp_get_params(self.h, _fptr(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().pt_mlp_set_params(self.h, _fptr(p)))

    @staticmethod
    def _rows(x):
        x = np.ascontiguousarray(np.atleast_2d(np.asarray(x, dtype=np.float32)))
        return x, x.shape[0], x.shape[1]

    def forward(self, x):
        x, n, k = self._rows(x)
        y = np.zeros(n, dtype=np.float32)
        _check(lib().pt_mlp_forward(self.h, _fptr(x), n, k, _fptr(y)))
        return y

    def gradient(self, x, t, mask=None):
        x, n, k = self._rows(x)
        t = np.ascontiguousarray(t, dtype=np.float32)
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = ctypes.c_float()
        _check(lib().pt_mlp_gradient(self.h, _fptr(x), _fptr(t), None if m is None else m.ctypes.data_as(ctypes.c_void_p),
                                     n, k, ctypes.byref(loss), _fptr(g)))
        return loss.value, g

    def train(self, x, t, iters, batch=256, lr=1e-3, seed=3):
        x, n, k = self._rows(x)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = ctypes.c_float()
        _check(lib().pt_mlp_train(self.h, _fptr(x), _fptr(t), n, k, int(iters), int(batch), float(lr), seed,
                                  ctypes.byref(loss)))
        return loss.value


class Group:
    """Independent PT runs of several problems sharing every launch (pt_group_*)."""

    def __init__(self, problems, steps_per_sweep=0):
        self.problems = list(problems)
        self.cps = [CProblem(pb) for pb in self.problems]
        arr = (PtProblem * len(self.cps))(*[cp.struct for cp in self.cps])
        self._arr = arr
        h = ctypes.c_void_p()
        _check(lib().pt_group_create(ctypes.cast(arr, ctypes.c_void_p), len(self.cps), int(steps_per_sweep),
                                     ctypes.byref(h)))
        self.h = h
        self.members = []
        for m in range(len(self.cps)):
            info, off = PtPlanInfo(), ctypes.c_int64()
            _check(lib().pt_group_member(self.h, m, ctypes.byref(info), ctypes.byref(off)))
            self.members.append(dict({f: getattr(info, f) for f, _ in PtPlanInfo._fields_}, offset=off.value))
        last = self.members[-1]
        self.state_size = last['offset'] + last['chains'] * last['stride']
        self.total_chains = sum(m['chains'] for m in self.members)

    def close(self):
        if getattr(self, 'h', None):
            lib().pt_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def initial(self, member, vectors):
        x = np.ascontiguousarray(np.atleast_2d(np.asarray(vectors, dtype=np.float64)))
        if x.shape[1] != self.members[member]['nv']:
            raise PTError('trial vectors must have %d entries' % self.members[member]['nv'])
        _check(lib().pt_group_set_states(self.h, member, _dptr(x), x.shape[0]))

    def costs(self):
        """chi^2 of every member's chains at their current vectors, per member."""
        out = np.zeros(self.total_chains)
        _check(lib().pt_group_costs(self.h, _dptr(out)))
        res, k = [], 0
        for m in self.members:
            res.append(out[k:k + m['chains']])
            k += m['chains']
        return res

    def batch(self, count, bootstrap=False, snapshot=True):
        out = np.zeros(self.state_size) if snapshot else None
        _check(lib().pt_group_batch(self.h, int(count), 1 if bootstrap else 0, _dptr(out) if snapshot else None))
        return self.split(out) if snapshot else None

    def split(self, state):
        return [state[m['offset']:m['offset'] + m['chains'] * m['stride']] for m in self.members]

    def tiles(self, request=0):
        t, ms = c_int32(), np.zeros(3)
        _check(lib().pt_group_tiles(self.h, int(request), ctypes.byref(t), _dptr(ms)))
        return t.value, ms

    def timing(self):
        t = c_double()
        _check(lib().pt_group_last_timing(self.h, ctypes.byref(t)))
        return t.value


def gpu_simulate(pb, x):
    """Model pattern of problem pb at variables x on the GPU (for synthetic data)."""
    with Plan(pb) as pl:
        return pl.pattern(x)


# ---------------------------------------------------------------------------------
#  RVPTGPU-shaped API (js/pt_gpu.js:20-424)
# ---------------------------------------------------------------------------------
def compile(ctx):
    """compile(ctx): the plan summary without a device (js/pt_gpu.js:226 fields)."""
    with Plan(ctx['problem'], device=False) as pl:
        info = dict(pl.info)
    info.update(nv=pl.nv, chains=pl.chains, stride=pl.stride)
    return info


class Session:
    """create(ctx) -> {plan, initial, costs, batch, close} (js/pt_gpu.js:371-406)."""

    def __init__(self, ctx):
        self.ctx = ctx
        self._plan = Plan(ctx['problem'])
        self.plan = dict(self._plan.info)

    def initial(self, vectors):
        self._plan.initial(vectors)

    def costs(self, vectors):
        return self._plan.costs(vectors)

    def batch(self, count, bootstrap=False):
        return self._plan.batch(count, bootstrap)

    def close(self):
        self._plan.close()


def create(ctx):
    return Session(ctx)


def _reflected(v, lo, hi):
    """js/pt_gpu.js:408 reflected()"""
    if math.isfinite(lo) and math.isfinite(hi):
        span = hi - lo
        if span <= 0:
            return lo
        folded = math.fmod(math.fmod(v - lo, 2 * span) + 2 * span, 2 * span)
        return lo + (folded if folded <= span else 2 * span - folded)
    return 2 * lo - v if v < lo else 2 * hi - v if v > hi else v


def run(ctx, tolerance=1e-3):
    """run(ctx) (js/pt_gpu.js:409-423).  ctx: problem, x (start), scales, sweeps, chains,
    max_temp, seed, cost(x) -> FP64 chi^2 (the caller's model), feasible(x), should_stop(),
    on_iter(dict), initial_cost.  Returns {x, stopped, pt: {...}}; raises PTError with
    'disagrees' when a probe cost is off the FP64 model by more than `tolerance`."""
    pb = ctx['problem']
    gpu = create(ctx)
    try:
        nv = pb.nv
        x0 = np.asarray(ctx.get('x', pb.x0), dtype=np.float64)
        probes = [x0, np.array([_reflected(v + pb.scale[j] * 0.25, pb.lo[j], pb.hi[j]) for j, v in enumerate(x0)])]
        gc = gpu.costs(probes)
        validation = 0.0
        for c in range(pb.chains):
            cpu = ctx['cost'](probes[c % 2])
            if not math.isfinite(cpu):
                continue
            err = abs(gc[c] - cpu) / max(1.0, cpu)
            validation = max(validation, err)
            if not math.isfinite(gc[c]) or err > tolerance:
                raise PTError('MI355X PT cost disagrees with the double-precision model (relative error %.3g).' % err)
        gpu.initial([x0])
        state = gpu.batch(1, True)
        best = x0.copy()
        best_cost = ctx.get('initial_cost', ctx['cost'](x0))
        completed, stopped = 0, False
        seen = set()
        stride = gpu.plan['stride']
        sweeps = ctx.get('sweeps', pb.sweeps)

        def inspect(state):
            nonlocal best, best_cost
            for c in range(pb.chains):
                b = c * stride
                cand = state[b + 2 * nv: b + 3 * nv]
                key = cand.tobytes()
                if key in seen:
                    continue
                seen.add(key)
                if not all(math.isfinite(v) and pb.lo[j] <= v <= pb.hi[j] for j, v in enumerate(cand)):
                    continue
                if ctx.get('feasible') and not ctx['feasible'](cand):
                    continue
                v = ctx['cost'](cand)
                if v < best_cost:
                    best, best_cost = cand.copy(), v

        while completed < sweeps:
            if ctx.get('should_stop') and ctx['should_stop']():
                stopped = True
                break
            count = min(max(1, math.ceil(8 / nv)), sweeps - completed)
            try:
                state = gpu.batch(count, False)
            except PTError as e:
                # js/pt_gpu.js:418: a chain whose intensity solve failed ends the GPU run
                if 'did not converge' in str(e):
                    raise PTError('MI355X PT intensity solve did not converge; restarting CPU PT.') from None
                raise
            completed += count
            inspect(state)
            if ctx.get('on_iter'):
                ctx['on_iter']({'it': completed, 'chi2': best_cost, 'stage': 'pt', 'backend': 'mi355x'})
        stats = {'accepted': 0, 'proposals': 0, 'swaps': 0, 'swapAttempts': 0}
        for c in range(pb.chains):
            b = c * stride + 3 * nv
            stats['accepted'] += int(state[b + S_ACCEPTED])
            stats['proposals'] += int(state[b + S_PROPOSALS])
            stats['swaps'] += int(state[b + S_SWAPS])
            stats['swapAttempts'] += int(state[b + S_ATTEMPTS])
        pt = {'backend': 'mi355x', 'chains': pb.chains, 'sweeps': completed, 'requestedSweeps': sweeps,
              'maxTemp': pb.max_temp, 'seed': pb.seed, 'initialChi2': ctx.get('initial_cost'), 'bestChi2': best_cost,
              'validationError': validation}
        pt.update(stats)
        return {'x': best, 'stopped': stopped, 'pt': pt}
    finally:
        gpu.close()
