"""PT trial-pattern throughput on one MI355X.

Workload (the batched FP32 VALU configuration): Pnma PbSO4-type phase a=8.48 b=5.40 c=6.96,
5 sites / 24 atoms, 2theta 10-150 deg step 0.02 (7001 points), Cu Ka1/Ka2, TCH profile,
6-term Chebyshev background, 32 refined variables; 256 replicas x 64 trials per sweep.
A "step" is one PT sweep: 256 x 64 = 16384 trial patterns, each a full calculated pattern
and its chi^2, proposed, evaluated and accepted on the GPU, then one replica exchange.

Prints one JSON line.  `value` is the exactly evaluated trial patterns per second of the
chosen form (the staged PT step pipeline replayed as one HIP graph per sweep; proposals
rejected without an evaluation are not counted).  The same line carries, for every
workload of the benchmark table (the NaCl CPU-path case, the Pnma pipeline, the shared-
profile MFMA form, the uneven batch, the learned screen, free intensities, the LM normal
equations): the GPU rate, the plain single-trial rate, the FP64 CPU rate where the
restatement runs it, the chi^2 error against the FP64 oracle, and the roofline of the
dominant kernel, k_stage_accum (FP32 VALU), with its HBM traffic.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'PT trial patterns/s on 1 MI355X, plain vs chosen form, χ² rel err ≤1e-5'
PEAK_FP32 = 157.3e12      # MI355X FP32 vector peak, FLOP/s
PEAK_FP16_MFMA = 2.5e15   # MI355X dense FP16/BF16 MFMA, FLOP/s
PEAK_FP64 = 78.6e12       # MI355X FP64 vector peak, FLOP/s
PEAK_BF16_MFMA = 2.5e15
MLP_FLOP_FWD = 2 * (64 * 256 + 2 * 256 * 256 + 256)   # one row through 64-256-256-256-1
LEGS = ('nacl', 'matrix', 'uneven', 'learned', 'free', 'normal')
EXTRA_LEGS = ('training',)      # run alone with --leg (inside the learned leg otherwise)


def flop_per_trial(w):
    """Algorithmic FLOP of one trial: 8 N_hkl S + 10 W_tot + 4 NP + 2 N_T NP."""
    return 8 * w['n_hkl'] * w['s_terms'] + 10 * w['w_tot'] + 4 * w['np'] + 2 * w['n_bkg'] * w['np']


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, 'tests'))
    import oracle_lib as O
    return O


def chi2_error(pb, xs, costs):
    """Largest relative error of GPU chi^2 against the FP64 oracle over (x, cost) pairs."""
    O = _oracle()
    m = O.Model(pb)
    err = 0.0
    for x, g in zip(xs, costs):
        c = m.cost(np.asarray(x))
        if np.isfinite(c) and c < 3e38:
            err = max(err, abs(float(g) - c) / max(1.0, c))
    m.close()
    return err


def host_info():
    model = None
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    hwc = os.cpu_count() or 1
    return {'cpu_model': model, 'hardware_concurrency': hwc, 'threads': max(1, min(16, hwc))}


def cpu_rates(pb, single_sweeps, single_chains, par_sweeps, par_chains):
    """FP64 CPU PT (the oracle's temper() + cost()) on this host: one thread, and one
    independent temper() per thread on every thread of the box (capped at 16)."""
    O = _oracle()
    from powder6_amd.pt import CProblem
    m = O.Model(pb)
    t = time.perf_counter()
    _, _, st = m.temper(pb.start_vector(), pb.scale, single_sweeps, single_chains, pb.max_temp, 99)
    single = st[1] / (time.perf_counter() - t)
    m.close()
    hi = host_info()
    cp = CProblem(pb)
    sc = np.ascontiguousarray(pb.scale, dtype=np.float64)
    secs = O.ctypes.c_double()
    rate = O.lib().or_temper_parallel(O.ctypes.addressof(cp.struct), sc.ctypes.data_as(O.P_D), par_sweeps, par_chains,
                                      pb.max_temp, 7, hi['threads'], O.ctypes.byref(secs))
    return single, rate, secs.value, hi


def plain_rate(pt, plan, pb, n, seed=1):
    rng = np.random.default_rng(seed)
    x0 = pb.start_vector()
    X = x0 + rng.uniform(-1, 1, size=(n, pb.nv)) * np.array(pb.scale)
    X = np.minimum(np.array(pb.hi), np.maximum(np.array(pb.lo), X))
    plan.eval(X[:4], pt.FORM_PLAIN)
    t = time.perf_counter()
    chi = plan.eval(X, pt.FORM_PLAIN)
    return len(X) / (time.perf_counter() - t), X, chi


def pt_states(plan, pb, every):
    S = plan.batch(0).reshape(plan.chains, -1)
    pick = list(range(0, plan.chains, max(1, plan.chains // every)))
    return [S[c, :pb.nv].copy() for c in pick], [S[c, 3 * pb.nv] for c in pick]


def nacl_cpu_path(pt, PR, no_cpu, sweeps=1000):
    """NaCl Fm-3m, 5501 points, 16 replicas x 1000 sweeps x 12 parameters
    (192,000 trials), the reference's CPU-path case: the whole PT run on the GPU, the
    plain form, and the FP64 CPU restatement on a bounded sample."""
    pb = PR.nacl_fm3m(simulate=pt.gpu_simulate)
    out = {'workload': 'NaCl Fm-3m, 5,501 pts Cu Ka1, TCH, 6 Chebyshev terms, 16 replicas x 1000 sweeps x 12 params'}
    with pt.Plan(pb) as pl:
        pl.initial([pb.start_vector()])
        pl.batch(1, bootstrap=True)
        pl.sweeps(5)                                       # graph capture, tile tuning
        e0, s0 = pl.trial_counts()
        t = time.perf_counter()
        pl.sweeps(sweeps)
        dt = time.perf_counter() - t
        e1, s1 = pl.trial_counts()
        ev, skipped = e1 - e0, s1 - s0
        xs, gc = pt_states(pl, pb, 16)
        plain, X, chi = plain_rate(pt, pl, pb, 256)
    out['chosen'] = {'value': ev / dt, 'unit': 'trial patterns/s', 'sweeps': sweeps, 'us_per_step': dt / (sweeps * pb.nv) * 1e6,
                     'seconds_for_192000_proposals': dt * 1000 / sweeps,
                     'evaluated': ev, 'rejected_unevaluated': skipped}
    out['plain'] = {'value': plain, 'unit': 'trial patterns/s'}
    out['chi2_rel_err_vs_fp64'] = max(chi2_error(pb, xs, gc), chi2_error(pb, X[:16], chi[:16]))
    if not no_cpu:
        single, rate, secs, hi = cpu_rates(pb, 20, 16, 40, 4)
        out['cpu_fp64'] = {'single_thread': single, 'all_threads': rate, 'threads': hi['threads'], 'unit': 'trial patterns/s',
                           'sample': 'single thread: 16 replicas x 20 sweeps; %d threads x (4 replicas x 40 sweeps), %.1f s'
                                     % (hi['threads'], secs),
                           'seconds_for_192000_single_thread': 192000 / single}
        out['chosen_over_cpu_single'] = out['chosen']['value'] / single
    out['chosen_over_plain'] = out['chosen']['value'] / plain
    return out


def matrix_form(pt, PR, trials=32768):
    """The shared-profile MFMA form on the atom-only variant of the same Pnma workload:
    P (7001 x N_hkl) fixed, I[hkl, trial] per trial, Y = P I as FP16 hi/lo on MFMA."""
    pb = PR.pbso4_pnma(simulate=pt.gpu_simulate, atoms_only=True)
    x0 = pb.start_vector()
    rng = np.random.default_rng(5)
    X = x0 + rng.uniform(-1, 1, size=(trials, pb.nv)) * np.array(pb.scale)
    X = np.minimum(np.array(pb.hi), np.maximum(np.array(pb.lo), X))
    with pt.Plan(pb) as plan:
        plan.eval(X[:4096], pt.FORM_MATRIX)                 # builds P, warms up
        t = time.perf_counter()
        chi2 = plan.eval(X, pt.FORM_MATRIX)
        dt = time.perf_counter() - t
        dev_ms = plan.timing()[0]          # HIP events around the device work (vectors already uploaded)
        ti_ms, tp_ms = plan.matrix_timing()
        n_hkl = plan.workload(x0)['n_hkl']
        chi_bf16 = plan.eval(X[:8], pt.FORM_MATRIX_BF16)     # the BF16-pair precision comparison
        plain, _, _ = plain_rate(pt, plan, pb, 256)
    dense = 2.0 * len(pb.tth) * n_hkl                      # FLOP per trial of the dense product
    return {'value': trials / (dev_ms / 1e3), 'unit': 'trial patterns/s', 'trials': trials, 'n_vars': pb.nv,
            'value_with_host_transfers': trials / dt, 'plain': plain,
            'form': 'shared profile matrix, Y = P I on MFMA (v_mfma_f32_16x16x32_f16, hi/lo, 3 products), fused chi^2',
            'intensity_ms': ti_ms, 'product_ms': tp_ms,
            'product_dense_tflops': dense * trials / (tp_ms / 1e3) / 1e12,
            'product_frac_of_fp16_peak': dense * trials / (tp_ms / 1e3) / PEAK_FP16_MFMA,
            'roofline': {'bound': 'mfma', 'kernel': 'k_matrix_chi2', 'achieved': dense * trials / (tp_ms / 1e3) / 1e12,
                         'peak': PEAK_FP16_MFMA / 1e12, 'unit': 'TFLOP/s',
                         'frac': dense * trials / (tp_ms / 1e3) / PEAK_FP16_MFMA,
                         'flop_per_trial_dense': dense}}, pb, X[:8], chi2[:8], chi_bf16


def uneven_batch(pt, PR, no_cpu, sweeps=2):
    """48 independent patterns (cubic / tetragonal / monoclinic, 2000-12000 points,
    1-12 sites), 32 replicas each, every PT stage one launch over all of them, one graph
    launch per sweep of all 48 PT runs."""
    pbs = PR.uneven_batch(simulate=pt.gpu_simulate)
    with pt.Group(pbs) as g:
        for m, pb in enumerate(pbs):
            g.initial(m, [pb.start_vector()])
        costs = [c[0] for c in g.costs()]
        g.batch(1, bootstrap=True)
        g.batch(1, snapshot=False)                       # tunes the tile shape, captures the graph
        t = time.perf_counter()
        g.batch(sweeps, snapshot=False)
        dt = time.perf_counter() - t
        steps = g.members[0]['steps_per_sweep']
        tiles, tune_ms = g.tiles()
        refl = [m['reflection_count'] for m in g.members]
        pts = [m['point_count'] for m in g.members]
        trials = g.total_chains * steps * sweeps
    # the work of one trial of every member (at its start vector), the plain form on a
    # sample of members (the reference fits them one after another)
    flop_step, plain_t, plain_n = 0.0, 0.0, 0
    for i, pb in enumerate(pbs):
        with pt.Plan(pb) as pl:
            flop_step += flop_per_trial(pl.workload(pb.start_vector())) * pb.chains
            if i % 6 == 0:
                r, _, _ = plain_rate(pt, pl, pb, 48, seed=i)
                plain_t += 48 / r
                plain_n += 48
    sweep_s = dt / sweeps
    out = {'value': trials / dt, 'unit': 'trial patterns/s', 'patterns': len(pbs), 'replicas_each': 32,
           'trials_per_step': 32 * len(pbs), 'steps_per_sweep': steps, 'ms_per_sweep': 1e3 * sweep_s,
           'reflections_min_max': [min(refl), max(refl)], 'points_min_max': [min(pts), max(pts)],
           'tiles_per_wave': tiles, 'tuner_accum_ms': [round(v, 4) for v in tune_ms],
           'plain': plain_n / plain_t,
           'roofline_whole_step': {'bound': 'valu', 'achieved': flop_step * steps / sweep_s / 1e12, 'peak': PEAK_FP32 / 1e12,
                                   'unit': 'TFLOP/s', 'frac': flop_step * steps / sweep_s / PEAK_FP32,
                                   'note': 'all stages of a sweep (setup, lines, accumulate, finish, exchange) in the denominator'},
           'form': 'one launch per stage over all 48 PT runs, one graph per sweep'}
    out['chi2_rel_err_vs_fp64'] = max(chi2_error(pb, [pb.start_vector()], [c]) for pb, c in zip(pbs, costs))
    if not no_cpu:
        # FP64 CPU PT of three members of increasing size, one thread each
        rates = []
        for i in (0, 17, 34):
            pb = pbs[i]
            O = _oracle()
            m = O.Model(pb)
            t = time.perf_counter()
            _, _, st = m.temper(pb.start_vector(), pb.scale, 2, 4, pb.max_temp, 5)
            rates.append(st[1] / (time.perf_counter() - t))
            m.close()
        out['cpu_fp64_single_thread'] = {'value': float(np.mean(rates)), 'unit': 'trial patterns/s',
                                         'sample': 'members 0, 17, 34: 4 replicas x 2 sweeps each, mean rate'}
    out['chosen_over_plain'] = out['value'] / out['plain']
    return out


def learned_screening(pt, PR, no_cpu, warm=8, sweeps=16):
    """The P2_1/c molecular problem (17 atoms, 53 variables with one Biso), 512 replicas,
    7001 points, 24 sweeps from the same start (the last 16 timed): plain PT against
    surrogate-screened PT (4 candidates per chain and step, the network trained every
    step), against skipping PT (proposals predicted hopeless are rejected unevaluated) with
    the network trained every step, and with the skip test inside the PT step (one proposal
    per chain, 12 Adam steps before each sweep); best Rwp after the 24 sweeps for each, and
    the network's training throughput."""
    pb = PR.molecular_p21c(simulate=pt.gpu_simulate, chains=512, shared_b=True)
    x0 = pb.start_vector()
    wss = float(np.sum(np.array(pb.w) * np.array(pb.y) ** 2))
    out = {'workload': 'P2_1/c, 17 C/N/S atoms, 53 variables, 7001 pts Ka1/Ka2, 512 replicas, 24 sweeps',
           'network': '64-256-256-256-1 SiLU, BF16 MFMA, FP32 accumulation, Adam'}
    skip = dict(oversample=1, train_iters=1, batch=512, warmup_iters=100, cold_fraction=1.0, skip_eps=1e-2, skip_margin=0.02)
    runs = (('plain', None),
            ('screened', dict(oversample=4, train_iters=1, batch=512, warmup_iters=100, margin=0.002, cold_fraction=0.5)),
            ('skipping', skip),
            ('skipping_in_step', dict(oversample=1, train_iters=12, train_every=pb.nv, batch=512, warmup_iters=48,
                                      cold_fraction=1.0, skip_eps=0.1, skip_margin=0.005)))
    errs = []
    for label, screen in runs:
        with pt.Plan(pb) as pl:
            if screen:
                pl.screen(**screen)
            pl.initial([x0])
            pl.batch(1, bootstrap=True)
            pl.sweeps(warm)                                  # graph capture, network warm-up
            e0, s0 = pl.trial_counts()
            t = time.perf_counter()
            pl.sweeps(sweeps)
            dt = time.perf_counter() - t
            e1, s1 = pl.trial_counts()
            spw = pl.info['steps_per_sweep']
            S = pl.batch(0).reshape(pb.chains, -1)
            out[label] = {'value': (e1 - e0) / dt, 'unit': 'exact trial patterns/s', 'ms_per_sweep': 1e3 * dt / sweeps,
                          'proposals_per_s': pb.chains * spw * sweeps / dt,
                          'exact_evaluation_fraction': (e1 - e0) / max(1, (e1 - e0) + (s1 - s0)),
                          'best_rwp_after_24_sweeps': float(np.sqrt(np.min(S[:, 3 * pb.nv + 1]) / wss))}
            if screen:
                out[label]['proposals_screened_per_s'] = screen['oversample'] * pb.chains * spw * sweeps / dt
                out[label]['stats'] = pl.screen_stats()
                out[label]['config'] = screen
            xs, gc = pt_states(pl, pb, 8)
            errs.append(chi2_error(pb, xs, gc))
            if label == 'plain':
                roof_w = pl.workload(x0)
                pl.set_kernel_timing(True)
                pl.sweeps(1)
                _, kms, nl = pl.timing()
                pl.set_kernel_timing(False)
                f = flop_per_trial(roof_w)
                out['roofline'] = {'bound': 'valu', 'kernel': 'k_stage_accum', 'achieved': f * pb.chains / (kms / nl / 1e3) / 1e12,
                                   'peak': PEAK_FP32 / 1e12, 'unit': 'TFLOP/s',
                                   'frac': f * pb.chains / (kms / nl / 1e3) / PEAK_FP32, 'avg_launch_ms': kms / nl,
                                   'flop_per_trial': f, 'workload': roof_w}
                plain, X, chi = plain_rate(pt, pl, pb, 128)
                out['plain_single_trial'] = {'value': plain, 'unit': 'trial patterns/s'}
                errs.append(chi2_error(pb, X[:8], chi[:8]))
    out['chi2_rel_err_vs_fp64'] = max(errs)
    out['screen_overhead'] = out['screened']['ms_per_sweep'] / out['plain']['ms_per_sweep'] - 1.0
    out['skipping_overhead'] = out['skipping']['ms_per_sweep'] / out['plain']['ms_per_sweep'] - 1.0
    out['skipping_in_step_over_plain_time'] = out['skipping_in_step']['ms_per_sweep'] / out['plain']['ms_per_sweep']
    if not no_cpu:
        O = _oracle()
        m = O.Model(pb)
        t = time.perf_counter()
        _, _, st = m.temper(x0, pb.scale, 1, 2, pb.max_temp, 3)
        out['cpu_fp64_single_thread'] = {'value': st[1] / (time.perf_counter() - t), 'unit': 'trial patterns/s',
                                         'sample': '2 replicas x 1 sweep (106 trials)'}
        m.close()
    out['training'] = network_training(pt)
    return out


def network_training(pt):
    """Training throughput of the surrogate network alone: 400 Adam steps on minibatches of
    512 rows drawn from 65,536 stored rows (53 inputs), and one inference of 4096 rows."""
    rng = np.random.default_rng(4)
    Xn = rng.uniform(-1, 1, size=(65536, 53)).astype(np.float32)
    T = rng.normal(size=65536).astype(np.float32)
    with pt.Mlp(seed=2) as net:
        net.train(Xn, T, iters=20, batch=512)
        t = time.perf_counter()
        net.train(Xn, T, iters=400, batch=512)
        dt = time.perf_counter() - t
        net.forward(Xn[:4096])
        t = time.perf_counter()
        net.forward(Xn[:4096])
        di = time.perf_counter() - t
    flop = 3 * MLP_FLOP_FWD * 512 * 400
    return {'rows_per_s': 512 * 400 / dt, 'tflops': flop / dt / 1e12, 'us_per_step': dt / 400 * 1e6,
            'frac_of_bf16_peak': flop / dt / PEAK_BF16_MFMA, 'batch': 512,
            'inference_rows_per_s_with_host_copies': 4096 / di,
            'note': 'wall time of 400 Adam steps (3 launches each: the fused step, the gradient sum with Adam, the step count)'}


def free_intensities(pt, PR, sweeps=5):
    """PT over the NaCl pattern (5,501 points, 256 replicas) with the NaCl phase as a Le Bail
    phase (free intensities solved per trial) and as a Rietveld phase, alone and next to a
    Rietveld ZnS phase: trial patterns/s of each, from `sweeps` timed sweeps."""
    out = {'workload': 'NaCl Fm-3m 10-120 deg step 0.02 (5,501 pts), TCH, 256 replicas; '
                       'with_zns adds a Rietveld ZnS phase'}
    errs = []
    for mixed in (False, True):
        for method in ('lebail', 'rietveld'):
            pb = PR.lebail_probe(mixed=mixed, simulate=pt.gpu_simulate, chains=256)
            pb.phases[0].method = method
            with pt.Plan(pb) as pl:
                pl.initial([pb.start_vector()])
                pl.batch(1, bootstrap=True)
                pl.sweeps(2)
                e0, _ = pl.trial_counts()
                t0 = time.perf_counter()
                pl.sweeps(sweeps)
                dt = (time.perf_counter() - t0) / sweeps
                e1, _ = pl.trial_counts()
                xs, gc = pt_states(pl, pb, 8)
                errs.append(chi2_error(pb, xs, gc))
            out[('with_zns_' if mixed else '') + method] = {'value': (e1 - e0) / (dt * sweeps), 'unit': 'trial patterns/s',
                                                            'ms_per_sweep': dt * 1e3}
    out['chi2_rel_err_vs_fp64'] = max(errs)
    return out


def normal_equations(pt, PR, no_cpu, reps=10):
    """The FP64 normal equations of one LM step (js/rietveld_core.js:1953-1983) of the Pnma
    workload on the GPU, against the oracle's FP64 restatement on one host thread."""
    O = _oracle()
    pb = PR.pbso4_pnma(simulate=pt.gpu_simulate, chains=16)
    x = pb.start_vector()
    with pt.Plan(pb) as pl:
        pl.normal(x)
        t = time.perf_counter()
        for _ in range(reps):
            A, g, f = pl.normal(x)
        wall = (time.perf_counter() - t) / reps
        dev_ms = pl.timing()[0]
    m = O.Model(pb)
    t = time.perf_counter()
    A0, g0, f0, _ = m.normal(x)
    cpu = time.perf_counter() - t
    m.close()
    d = np.sqrt(np.diag(A0))
    out = {'workload': 'Pnma bench plan, 32 variables, 7001 points', 'ms_per_build': wall * 1e3, 'device_ms_per_build': dev_ms,
           'a_rel_err_vs_fp64': float(np.max(np.abs(A - A0) / np.outer(d, d))),
           'g_rel_err_vs_fp64': float(np.max(np.abs(g - g0) / (d * np.sqrt(f0)))),
           'chi2_rel_err_vs_fp64': abs(f - f0) / f0}
    if not no_cpu:
        out['cpu_fp64_single_thread_ms'] = cpu * 1e3
        out['gpu_over_cpu'] = cpu / wall
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--plain-trials', type=int, default=512)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--nacl-sweeps', type=int, default=1000,
                    help='sweeps of the NaCl case (fewer for a kernel trace of that leg alone)')
    ap.add_argument('--main-only', action='store_true',
                    help='only the chosen PT form (for counter passes under a profiler)')
    ap.add_argument('--tiles', type=int, default=0, choices=(0, 1, 2, 4),
                    help='accumulate tile shape (128-point tiles per wavefront); 0 lets the autotuner pick')
    ap.add_argument('--copy-probe', action='store_true',
                    help='also run a plain 1 GiB device copy (HBM counter calibration)')
    ap.add_argument('--leg', choices=LEGS + EXTRA_LEGS, default=None,
                    help='run only this workload of the table and print its JSON (for counter passes under a profiler)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one MI355X: this build runs on a single GPU (replica sharding across '
              'GPUs is a design note only); run with --gpus 1.', file=sys.stderr)
        sys.exit(2)

    from powder6_amd import problems as PR
    from powder6_amd import pt

    legs_fn = {'nacl': lambda: nacl_cpu_path(pt, PR, args.no_cpu, args.nacl_sweeps), 'matrix': lambda: matrix_form(pt, PR),
               'uneven': lambda: uneven_batch(pt, PR, args.no_cpu), 'learned': lambda: learned_screening(pt, PR, args.no_cpu),
               'free': lambda: free_intensities(pt, PR), 'normal': lambda: normal_equations(pt, PR, args.no_cpu),
               'training': lambda: network_training(pt)}
    if args.leg:
        r = legs_fn[args.leg]()
        if args.leg == 'matrix':
            r = r[0]
        if args.copy_probe:
            r['hbm_copy_gbps'] = pt.hbm_copy_probe(1 << 30, 3)
        print(json.dumps({'metric': METRIC, 'leg': args.leg, 'result': r}))
        return

    pb = PR.pbso4_pnma(simulate=pt.gpu_simulate)          # 256 replicas x 64 trials/sweep
    x0 = pb.start_vector()
    with pt.Plan(pb) as plan:
        chains, spw = plan.chains, plan.info['steps_per_sweep']
        work = plan.workload(x0)
        if args.tiles:
            plan.tiles(args.tiles)
        plan.initial([x0])
        plan.batch(1, bootstrap=True)
        for _ in range(args.warmup):
            plan.sweeps(1)
        e0, s0 = plan.trial_counts()
        t0 = time.perf_counter()
        plan.sweeps(args.steps)                               # graph replays; synchronises its stream
        t1 = time.perf_counter()
        e1, s1 = plan.trial_counts()
        xs, gc = pt_states(plan, pb, 16)
        tiles, tune_ms = plan.tiles()
        # dominant kernel (banded profile sum) timed with HIP events on the plan's stream
        plan.set_kernel_timing(True)
        plan.sweeps(2)
        total_ms, kernel_ms, launches = plan.timing()
        plan.set_kernel_timing(False)
        proposals = chains * spw * args.steps
        evaluated = e1 - e0
        value = evaluated / (t1 - t0)
        ms_per_step = 1e3 * (t1 - t0) / args.steps
        # plain single-trial form on a bounded sample of trial vectors
        plain, X, chi_plain = plain_rate(pt, plan, pb, args.plain_trials)
        # the staged stages on 16384 independent vectors (one pass): device time between
        # HIP events with the vectors in HBM, and wall time with the host copies
        XB = np.repeat(X, (16384 + len(X) - 1) // len(X), axis=0)[:16384]
        plan.eval(XB, pt.FORM_BATCHED)
        t = time.perf_counter()
        gb = plan.eval(XB, pt.FORM_BATCHED)
        batched_wall = len(XB) / (time.perf_counter() - t)
        batched = len(XB) / (plan.timing()[0] / 1e3)

    copy_gbps = None
    if args.copy_probe:
        copy_gbps = pt.hbm_copy_probe(1 << 30, 3)       # 1 + 3 launches of k_copy_probe, 1 GiB each
        print('copy probe: k_copy_probe 1 GiB, %.0f GB/s' % copy_gbps, file=sys.stderr)
    if args.main_only:
        print(json.dumps({'metric': METRIC, 'value': value, 'unit': 'trial patterns/s', 'main_only': True,
                          'plain': plain, 'batched': batched, 'hbm_copy_gbps': copy_gbps}))
        return
    f_trial = flop_per_trial(work)
    launch_s = kernel_ms / max(1, launches) / 1e3
    achieved = f_trial * chains / launch_s           # one launch = one PT step of every chain
    # HBM bytes per launch of the tuned tile shape, from the committed FETCH_SIZE / WRITE_SIZE
    # passes (scaled by the k_copy_probe calibration)
    traffic, traffic_src = None, None
    tfile = os.path.join(ROOT, 'profiles', 'k_stage_accum_traffic.json')
    if os.path.exists(tfile):
        with open(tfile) as f:
            tr = json.load(f)
        ent = tr.get('per_tiles_per_wave', {}).get(str(tiles))
        if ent:
            traffic, traffic_src = ent['bytes_per_launch'], 'profiles/k_stage_accum_traffic.json (tiles per wave %d)' % tiles
    roof = {'bound': 'valu', 'kernel': 'k_stage_accum', 'achieved': achieved / 1e12, 'peak': PEAK_FP32 / 1e12,
            'unit': 'TFLOP/s', 'frac': achieved / PEAK_FP32, 'traffic': traffic, 'traffic_source': traffic_src,
            'hbm_gbps_at_traffic': (traffic / launch_s / 1e9) if traffic else None,
            'flop_per_trial': f_trial, 'avg_launch_ms': launch_s * 1e3, 'workload': work,
            'whole_step_frac': f_trial * evaluated / (t1 - t0) / PEAK_FP32}

    legs = {}
    for name in LEGS:
        t = time.perf_counter()
        legs[name] = legs_fn[name]()
        print('bench leg %s done in %.1f s' % (name, time.perf_counter() - t), file=sys.stderr, flush=True)
    nacl, uneven, learned, free, normal = legs['nacl'], legs['uneven'], legs['learned'], legs['free'], legs['normal']
    mat, mpb, mX, mchi, mchi_bf16 = legs['matrix']

    cpu = None
    err_chosen = chi2_error(pb, xs, gc)
    if not args.no_cpu:
        single, rate, secs, hi = cpu_rates(pb, 20, 4, 40, 2)
        cpu = {'value': rate, 'unit': 'trial patterns/s', 'cores': hi['threads'], 'kind': 'port',
               'sample': ('oracle FP64 temper() + cost() on the same Pnma problem: %d threads x (2 replicas x 40 sweeps x %d '
                          'trials), %.1f s wall; single thread 4 replicas x 20 sweeps' % (hi['threads'], pb.nv, secs)),
               'single_thread': single, 'cpu_model': hi['cpu_model'], 'hardware_concurrency': hi['hardware_concurrency'],
               'threads_used': hi['threads'],
               'note': ('one independent PT run per thread; the reference pool runs at most min(8, hardware_concurrency - 1) '
                        'fits at once (js/rietveld/rv_commands.js:14), so %d threads is at least its parallelism'
                        % hi['threads'])}
        O = _oracle()
        mm = O.Model(mpb)
        mat['chi2_rel_err_vs_fp64'] = max(abs(g - mm.cost(x)) / max(1.0, mm.cost(x)) for x, g in zip(mX, mchi))
        mat['chi2_rel_err_vs_fp64_with_bf16_pairs'] = max(abs(g - mm.cost(x)) / max(1.0, mm.cost(x))
                                                          for x, g in zip(mX, mchi_bf16))
        mm.close()
    out = {
        'metric': METRIC, 'value': value, 'unit': 'trial patterns/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': ms_per_step, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'fp32', 'data': 'synthetic (seeded Poisson pattern of the model at seeded true parameters)',
        'config': {'workload': 'Pnma PbSO4-type, 24 atoms, 7001 pts Ka1/Ka2, TCH, 256 replicas x 64 trials/sweep',
                   'replicas': chains, 'trials_per_sweep': spw, 'n_vars': pb.nv, 'points': work['np'],
                   'reflections': work['n_hkl'], 'lines': work['lines'], 'form': 'staged PT step pipeline, one graph per sweep',
                   'tiles_per_wave': tiles, 'tuner_accum_ms': [round(v, 4) for v in tune_ms]},
        'proposals_per_s': proposals / (t1 - t0), 'evaluated_fraction': evaluated / max(1, evaluated + (s1 - s0)),
        'chi2_rel_err_vs_fp64': err_chosen,
        'plain': {'value': plain, 'unit': 'trial patterns/s', 'form': 'one trial per launch sequence',
                  'chi2_rel_err_vs_fp64': chi2_error(pb, X[:16], chi_plain[:16])},
        'batched_eval': {'value': batched, 'value_with_host_transfers': batched_wall, 'unit': 'trial patterns/s',
                         'form': 'staged pipeline, 16384 independent vectors',
                         'chi2_rel_err_vs_fp64': chi2_error(pb, XB[:16], gb[:16])},
        'nacl_cpu_path_case': nacl,
        'matrix_form': mat,
        'uneven_batch': uneven,
        'learned_screening': learned,
        'free_intensities': free,
        'lm_normal_equations': normal,
        'chosen_over_plain': value / plain,
        'roofline': roof,
        'cpu_baseline': cpu,
    }
    if cpu:
        out['chosen_over_cpu'] = value / cpu['value']
        out['chosen_over_cpu_single_thread'] = value / cpu['single_thread']
    print(json.dumps(out))


if __name__ == '__main__':
    main()
