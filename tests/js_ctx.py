"""A PT context shaped like the one RVCore yields (js/rietveld_core.js:2271-2273), built
from a problems.Problem, for driving powder6_amd/node/pt_hip.js from tests.  JSON has no
infinities: +-inf limits are written as null and restored by tests/js/pt_hip_probe.js."""
import json
import os
import math

import numpy as np

from powder6_amd import problems as PR


def _num(v):
    v = float(v)
    return None if not math.isfinite(v) else v


def _values(pb):
    values = {}
    p = pb.values()
    for name, k in pb.index.items():
        if '._prof' in name:
            continue
        values[name] = float(p[k])
    return values


def _dataset(pb, model_phases=None):
    """Model.datasets[d] of the context: `has` and `anom` follow the model's phase list
    (a phase may be absent from a dataset)."""
    model_phases = model_phases or pb.phases
    own = {ph.key: ph for ph in pb.phases}
    anom = [[[at.fp, at.fpp] for at in own[ph.key].atoms] if ph.key in own else [] for ph in model_phases]
    ds = {'def': {'tth': pb.tth.tolist(), 'y': pb.y.tolist(), 'w': pb.w.tolist(), 'lambda1': pb.lambda1,
                  'lambda2': pb.lambda2, 'ratio': pb.ratio, 'polK': pb.pol_k, 'minFwhm': pb.min_fwhm,
                  'radiation': 'neutron' if pb.radiation else 'xray', 'resolutionType': pb.resolution_type},
          'has': [ph.key in own for ph in model_phases], 'anom': anom}
    if getattr(pb, 'bkg_type', PR.BKG_CHEBYSHEV) == PR.BKG_SPLINE:
        ds.update(spline=True, T=[], bkgX=[float(v) for v in pb.bkg_x])
    else:
        ds.update(spline=False, T=[0] * pb.n_bkg, bkgX=[])
    return ds


def reference_ctx(pb, initial_cost=None):
    if isinstance(pb, (list, tuple)):
        # several datasets over the first one's phases: values per dataset; a variable
        # label name@d is the reference's variable `name` acting in dataset d only
        # (ctx.info[j].ds), a plain label acts in every dataset
        ctx = reference_ctx(pb[0], initial_cost)
        ctx['problem']['values'] = [_values(q) for q in pb]
        ctx['problem']['vars'] = [{'name': n.split('@')[0]} for n in pb[0].vars]
        ctx['model']['datasets'] = [_dataset(q, pb[0].phases) for q in pb]
        for j, n in enumerate(pb[0].vars):
            ctx['info'][j]['ds'] = [int(n.split('@')[1])] if '@' in n else list(range(len(pb)))
        links = [{'d': d, 'name': name, 'offset': off, 'terms': [[j, c] for j, c in terms]}
                 for d, q in enumerate(pb) for (name, off, terms) in getattr(q, 'links', [])]
        if links:
            ctx['constraints'] = {'links': links, 'names': sorted({l['name'] for l in links})}
        return ctx
    values = {}
    p = pb.values()
    for name, k in pb.index.items():
        if name.startswith('p1._prof') or '._prof' in name:
            continue
        values[name] = float(p[k])
    phases = []
    for q, ph in enumerate(pb.phases):
        atoms = []
        for at in ph.atoms:
            basis = {str(f): list(map(float, at.site['basis'][f])) for f in at.site['free']}
            atoms.append({'label': at.label, 'ff': {'xray': {'a': list(at.ff_a), 'b': list(at.ff_b), 'c': at.ff_c},
                                                    'neutron': at.b_neutron},
                          'site': {'r0': list(map(float, at.site['r0'])), 'free': list(at.site['free']),
                                   'basis': basis, 'mult': at.site['mult']}})
        ops = [{'r': [int(v) for v in r], 't': [float(v) for v in t]} for r, t in zip(ph.op_r, ph.op_t)]
        pdef = {'key': ph.key, 'system': ph.system, 'uniqueAxis': 'abc'[ph.unique_axis],
                'profileType': ph.profile, 'ops': ops, 'poDir': list(ph.po_dir), 'atoms': atoms}
        if getattr(ph, 'separated', False):
            pdef['separatedProfile'] = True
            pdef['broadening'] = {'size': ph.size_model, 'strain': ph.strain_model,
                                  'sizeAxis': list(ph.size_axis), 'strainAxis': list(ph.strain_axis)}
        phases.append({'def': pdef, 'method': getattr(ph, 'method', 'rietveld'), 'prefix': ph.key + '.'})
    anom = [[[at.fp, at.fpp] for at in ph.atoms] for ph in pb.phases]
    ds = {'def': {'tth': pb.tth.tolist(), 'y': pb.y.tolist(), 'w': pb.w.tolist(), 'lambda1': pb.lambda1,
                  'lambda2': pb.lambda2, 'ratio': pb.ratio, 'polK': pb.pol_k, 'minFwhm': pb.min_fwhm,
                  'radiation': 'neutron' if pb.radiation else 'xray', 'resolutionType': pb.resolution_type},
          'has': [True] * len(pb.phases), 'anom': anom}
    if getattr(pb, 'bkg_type', PR.BKG_CHEBYSHEV) == PR.BKG_SPLINE:
        ds.update(spline=True, T=[], bkgX=[float(v) for v in pb.bkg_x])
    else:
        ds.update(spline=False, T=[0] * pb.n_bkg, bkgX=[])
    return {
        'problem': {'values': [values], 'vars': [{'name': n} for n in pb.vars]},
        'model': {'datasets': [ds], 'phases': phases},
        'x': [float(v) for v in pb.x0],
        'info': [{'lo': _num(lo), 'hi': _num(hi)} for lo, hi in zip(pb.lo, pb.hi)],
        'setup': {'scales': [float(s) for s in pb.scale], 'chains': pb.chains, 'sweeps': pb.sweeps,
                  'maxTemp': pb.max_temp, 'seed': pb.seed, 'initialCost': initial_cost},
        'constraints': ({'links': [{'d': 0, 'name': n, 'offset': o, 'terms': [[j, c] for j, c in t]}
                                   for (n, o, t) in pb.links]} if getattr(pb, 'links', None) else None),
        'maxIter': pb.sweeps,
    }


class OracleCostPipe:
    """ctx.cost for tests/js/pt_hip_probe.js: two FIFOs in `directory` and a thread of the
    test process that answers every vector the Node side writes (one JSON array per line)
    with the FP64 oracle's cost (`model.cost`), one number per line.  Use as a context
    manager; `paths` goes into the context as ctx.costPipe."""

    def __init__(self, model, directory):
        import threading
        self.model = model
        self.req = os.path.join(str(directory), 'cost_req')
        self.resp = os.path.join(str(directory), 'cost_resp')
        os.mkfifo(self.req)
        os.mkfifo(self.resp)
        self.calls = 0
        self.thread = threading.Thread(target=self._serve, daemon=True)
        self.thread.start()

    @property
    def paths(self):
        return {'req': self.req, 'resp': self.resp}

    def _serve(self):
        # the same open order as the Node side: requests first, then answers
        with open(self.req, 'r') as fr, open(self.resp, 'w') as fw:
            for line in fr:
                c = float(self.model.cost(np.asarray(json.loads(line), dtype=np.float64)))
                self.calls += 1
                fw.write(('Infinity' if c == math.inf else 'NaN' if c != c else repr(c)) + '\n')
                fw.flush()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.thread.join(timeout=5)
        return False


def write_ctx(pb, path, **extra):
    ctx = reference_ctx(pb, extra.pop('initial_cost', None))
    ctx.update(extra)
    with open(path, 'w') as f:
        json.dump(ctx, f)
    return path
