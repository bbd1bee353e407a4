"""PT trial-pattern throughput on one MI355X.

Workload (the batched FP32 VALU configuration): Pnma PbSO4-type phase a=8.48 b=5.40 c=6.96,
5 sites / 24 atoms, 2theta 10-150 deg step 0.02 (7001 points), Cu Ka1/Ka2, TCH profile,
6-term Chebyshev background, 32 refined variables; 256 replicas x 64 trials per sweep.
A "step" is one PT sweep: 256 x 64 = 16384 trial patterns, each a full calculated pattern
and its chi^2, proposed, evaluated and accepted on the GPU, then one replica exchange.

Prints one JSON line (value = trial patterns/s of the chosen form, the staged PT
step pipeline replayed as one HIP graph per sweep), with the plain single-trial form, the roofline of the dominant kernel
(k_stage_accum, FP32 VALU) and the FP64 CPU PT baseline (the oracle, timed on this host) in
the same line.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'PT trial patterns/s on 1 MI355X, plain vs chosen form, χ² rel err ≤1e-5'
PEAK_FP32 = 157.3e12      # MI355X FP32 vector peak, FLOP/s


def flop_per_trial(w):
    """Algorithmic FLOP of one trial: 8 N_hkl S + 10 W_tot + 4 NP + 2 N_T NP."""
    return 8 * w['n_hkl'] * w['s_terms'] + 10 * w['w_tot'] + 4 * w['np'] + 2 * w['n_bkg'] * w['np']


def cpu_baseline(pb, x_states, gpu_costs):
    """The FP64 CPU PT path (oracle restatement of temper() + cost()) on this host, on a
    bounded sample; also checks chi^2 of the GPU chain states against the FP64 oracle."""
    sys.path.insert(0, os.path.join(ROOT, 'tests'))
    import oracle_lib as O
    m = O.Model(pb)
    errs = []
    for x, g in zip(x_states, gpu_costs):
        c = m.cost(x)
        errs.append(abs(g - c) / max(1.0, c))
    # single thread: one temper() run of 4 replicas x 20 sweeps (about 10 s)
    x0 = pb.start_vector()
    t = time.perf_counter()
    _, _, st = m.temper(x0, pb.scale, 20, 4, pb.max_temp, 99)
    dt = time.perf_counter() - t
    single = st[1] / dt
    # all host threads: one independent temper() per thread (the reference runs one PT fit
    # per worker, js/rietveld/rv_commands.js:11-100)
    threads = max(1, min(16, os.cpu_count() or 1))
    from powder6_amd.pt import CProblem
    cp = CProblem(pb)
    sc = np.ascontiguousarray(pb.scale, dtype=np.float64)
    secs = O.ctypes.c_double()
    rate = O.lib().or_temper_parallel(O.ctypes.addressof(cp.struct), sc.ctypes.data_as(O.P_D), 40, 2, pb.max_temp, 7,
                                      threads, O.ctypes.byref(secs))
    m.close()
    return {
        'value': rate, 'unit': 'trial patterns/s', 'cores': threads, 'kind': 'port',
        'sample': ('oracle FP64 temper() on the same problem: %d threads x (2 replicas x 40 sweeps x %d trials), '
                   '%.1f s wall; single thread %.1f trials/s (4 replicas x 20 sweeps)' % (threads, pb.nv, secs.value, single)),
        'single_thread': single,
        'chi2_rel_err_vs_fp64': max(errs) if errs else None,
    }


PEAK_FP16_MFMA = 2.5e15   # MI355X dense FP16/BF16 MFMA, FLOP/s


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
    dense = 2.0 * len(pb.tth) * n_hkl                      # FLOP per trial of the dense product
    return {'value': trials / (dev_ms / 1e3), 'unit': 'trial patterns/s', 'trials': trials, 'n_vars': pb.nv,
            'value_with_host_transfers': trials / dt,
            'form': 'shared profile matrix, Y = P I on MFMA (FP16 hi/lo, 3 products), fused chi^2',
            'intensity_ms': ti_ms, 'product_ms': tp_ms,
            'product_dense_tflops': dense * trials / (tp_ms / 1e3) / 1e12,
            'product_frac_of_fp16_peak': dense * trials / (tp_ms / 1e3) / PEAK_FP16_MFMA}, pb, X[:8], chi2[:8], chi_bf16


def uneven_batch(pt, PR, sweeps=2):
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
    return {'value': trials / dt, 'unit': 'trial patterns/s', 'patterns': len(pbs), 'replicas_each': 32,
            'trials_per_step': 32 * len(pbs), 'steps_per_sweep': steps, 'ms_per_sweep': 1e3 * dt / sweeps,
            'reflections_min_max': [min(refl), max(refl)], 'points_min_max': [min(pts), max(pts)],
            'tiles_per_wave': tiles, 'tuner_accum_ms': [round(v, 4) for v in tune_ms],
            'form': 'one launch per stage over all 48 PT runs, one graph per sweep'}, pbs, costs


MLP_FLOP_FWD = 2 * (64 * 256 + 2 * 256 * 256 + 256)   # one row through 64-256-256-256-1
PEAK_BF16_MFMA = 2.5e15


def free_intensities(pt, PR, sweeps=5):
    """PT over the NaCl pattern (5,501 points, 256 replicas) with the NaCl phase as a Le Bail
    phase (free intensities solved per trial) and as a Rietveld phase, alone and next to a
    Rietveld ZnS phase: trial patterns/s of each, from `sweeps` timed sweeps."""
    out = {'workload': 'NaCl Fm-3m 10-120 deg step 0.02 (5,501 pts), TCH, 256 replicas; '
                       'with_zns adds a Rietveld ZnS phase'}
    for mixed in (False, True):
        for method in ('lebail', 'rietveld'):
            pb = PR.lebail_probe(mixed=mixed, simulate=pt.gpu_simulate, chains=256)
            pb.phases[0].method = method
            with pt.Plan(pb) as pl:
                pl.initial([pb.start_vector()])
                pl.batch(1, bootstrap=True)
                pl.sweeps(2)
                t0 = time.perf_counter()
                pl.sweeps(sweeps)
                dt = (time.perf_counter() - t0) / sweeps
                rate = pb.chains * pl.info['steps_per_sweep'] / dt
            out[('with_zns_' if mixed else '') + method] = {'value': rate, 'unit': 'trial patterns/s',
                                                            'ms_per_sweep': dt * 1e3}
    return out


def learned_screening(pt, PR, sweeps=2):
    """The P2_1/c molecular problem (17 atoms, 53 variables with one Biso), 512 replicas,
    7001 points: plain PT against surrogate-screened PT (4 candidates per chain and step,
    one Adam step of 512 rows per PT step), against skipping PT (proposals predicted
    hopeless are rejected unevaluated), and the network's training throughput."""
    pb = PR.molecular_p21c(simulate=pt.gpu_simulate, chains=512, shared_b=True)
    x0 = pb.start_vector()
    out = {'workload': 'P2_1/c, 17 C/N/S atoms, 53 variables, 7001 pts Ka1/Ka2, 512 replicas',
           'network': '64-256-256-256-1 SiLU, BF16 MFMA, FP32 accumulation, Adam'}
    runs = (('plain', None),
            ('screened', dict(oversample=4, train_iters=1, batch=512, warmup_iters=100, margin=0.002, cold_fraction=0.5)),
            ('skipping', dict(oversample=1, train_iters=1, batch=512, warmup_iters=100, cold_fraction=1.0,
                              skip_eps=1e-2, skip_margin=0.02)))
    for label, screen in runs:
        with pt.Plan(pb) as pl:
            if screen:
                pl.screen(**screen)
            pl.initial([x0])
            pl.batch(1, bootstrap=True)
            pl.sweeps(2)                                     # warm: graph capture, network warm-up
            t = time.perf_counter()
            pl.sweeps(sweeps)
            dt = time.perf_counter() - t
            spw = pl.info['steps_per_sweep']
            out[label] = {'value': pb.chains * spw * sweeps / dt, 'unit': 'exact trial patterns/s',
                          'ms_per_sweep': 1e3 * dt / sweeps}
            if screen:
                out[label]['proposals_screened_per_s'] = screen['oversample'] * pb.chains * spw * sweeps / dt
                out[label]['stats'] = st = pl.screen_stats()
                if screen.get('skip_eps'):
                    proposals = float(np.sum(pl.batch(0).reshape(pb.chains, -1)[:, 3 * pb.nv + 4]))
                    out[label]['value'] *= 1.0 - st['skipped'] / proposals
                    out[label]['exact_evaluation_fraction'] = 1.0 - st['skipped'] / proposals
    out['screen_overhead'] = out['screened']['ms_per_sweep'] / out['plain']['ms_per_sweep'] - 1.0
    # training throughput of the network alone (minibatches of 512 rows)
    rng = np.random.default_rng(4)
    X = rng.uniform(-1, 1, size=(65536, 53)).astype(np.float32)
    T = rng.normal(size=65536).astype(np.float32)
    with pt.Mlp(seed=2) as net:
        net.train(X, T, iters=20, batch=512)
        t = time.perf_counter()
        net.train(X, T, iters=400, batch=512)
        dt = time.perf_counter() - t
    flop = 3 * MLP_FLOP_FWD * 512 * 400
    out['training'] = {'rows_per_s': 512 * 400 / dt, 'tflops': flop / dt / 1e12,
                       'frac_of_bf16_peak': flop / dt / PEAK_BF16_MFMA, 'batch': 512,
                       'note': 'wall time of 400 Adam steps (13 launches each) incl. launch gaps'}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--plain-trials', type=int, default=512)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--main-only', action='store_true',
                    help='only the chosen PT form (for counter passes under a profiler)')
    ap.add_argument('--tiles', type=int, default=0, choices=(0, 1, 2, 4),
                    help='accumulate tile shape (128-point tiles per wavefront); 0 lets the autotuner pick')
    ap.add_argument('--copy-probe', action='store_true',
                    help='also run a plain 1 GiB device copy (HBM counter calibration)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one MI355X: this build runs on a single GPU (replica sharding across '
              'GPUs is a design note only); run with --gpus 1.', file=sys.stderr)
        sys.exit(2)

    from powder6_amd import problems as PR
    from powder6_amd import pt

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
        t0 = time.perf_counter()
        plan.sweeps(args.steps)                               # graph replays; synchronises its stream
        t1 = time.perf_counter()
        state = plan.batch(0)
        tiles, tune_ms = plan.tiles()
        # dominant kernel (banded profile sum) timed with HIP events on the plan's stream
        plan.set_kernel_timing(True)
        plan.sweeps(2)
        total_ms, kernel_ms, launches = plan.timing()
        plan.set_kernel_timing(False)
        trials = chains * spw * args.steps
        value = trials / (t1 - t0)
        ms_per_step = 1e3 * (t1 - t0) / args.steps
        # plain single-trial form on a bounded sample of trial vectors
        rng = np.random.default_rng(1)
        X = x0 + rng.uniform(-1, 1, size=(args.plain_trials, pb.nv)) * np.array(pb.scale)
        X = np.minimum(np.array(pb.hi), np.maximum(np.array(pb.lo), X))
        plan.eval(X[:8], pt.FORM_PLAIN)
        t = time.perf_counter()
        plan.eval(X, pt.FORM_PLAIN)
        plain = len(X) / (time.perf_counter() - t)
        # batched (one launch, one workgroup per trial) on 16384 independent vectors
        XB = np.repeat(X, (16384 + len(X) - 1) // len(X), axis=0)[:16384]
        plan.eval(XB, pt.FORM_BATCHED)
        t = time.perf_counter()
        gb = plan.eval(XB, pt.FORM_BATCHED)
        batched = len(XB) / (time.perf_counter() - t)

    copy_gbps = None
    if args.copy_probe:
        copy_gbps = pt.hbm_copy_probe(1 << 30, 3)       # 1 + 3 launches of k_copy_probe, 1 GiB each
        print('copy probe: k_copy_probe 1 GiB, %.0f GB/s' % copy_gbps, file=sys.stderr)
    if args.main_only:
        print(json.dumps({'metric': METRIC, 'value': value, 'unit': 'trial patterns/s', 'main_only': True,
                          'plain': plain, 'batched': batched, 'hbm_copy_gbps': copy_gbps}))
        return
    mat, mpb, mX, mchi, mchi_bf16 = matrix_form(pt, PR)
    uneven, upbs, ucosts = uneven_batch(pt, PR)
    learned = learned_screening(pt, PR)
    free = free_intensities(pt, PR)

    S = state.reshape(chains, -1)
    nv = pb.nv
    pick = list(range(0, chains, max(1, chains // 16)))
    xs = [S[c, :nv].copy() for c in pick]
    gc = [S[c, 3 * nv + pt.S_COST] for c in pick]

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
            'flop_per_trial': f_trial, 'avg_launch_ms': launch_s * 1e3, 'workload': work}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pb, xs, gc)
        # the matrix form's chi^2 against the FP64 oracle on a sample of its trials
        sys.path.insert(0, os.path.join(ROOT, 'tests'))
        import oracle_lib as O
        mm = O.Model(mpb)
        mat['chi2_rel_err_vs_fp64'] = max(abs(g - mm.cost(x)) / max(1.0, mm.cost(x)) for x, g in zip(mX, mchi))
        mat['chi2_rel_err_vs_fp64_with_bf16_pairs'] = max(abs(g - mm.cost(x)) / max(1.0, mm.cost(x))
                                                          for x, g in zip(mX, mchi_bf16))
        mm.close()
        err = 0.0
        for upb, g in zip(upbs, ucosts):
            um = O.Model(upb)
            ref = um.cost(upb.start_vector())
            um.close()
            err = max(err, abs(g - ref) / max(1.0, ref))
        uneven['chi2_rel_err_vs_fp64'] = err
    out = {
        'metric': METRIC, 'value': value, 'unit': 'trial patterns/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': ms_per_step, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'fp32', 'data': 'synthetic (seeded Poisson pattern of the model at seeded true parameters)',
        'config': {'workload': 'Pnma PbSO4-type, 24 atoms, 7001 pts Ka1/Ka2, TCH, 256 replicas x 64 trials/sweep',
                   'replicas': chains, 'trials_per_sweep': spw, 'n_vars': nv, 'points': work['np'],
                   'reflections': work['n_hkl'], 'lines': work['lines'], 'form': 'staged PT step pipeline, one graph per sweep',
                   'tiles_per_wave': tiles, 'tuner_accum_ms': [round(v, 4) for v in tune_ms]},
        'plain': {'value': plain, 'unit': 'trial patterns/s', 'form': 'one trial per launch sequence'},
        'batched_eval': {'value': batched, 'unit': 'trial patterns/s', 'form': 'staged pipeline, 16384 independent vectors'},
        'matrix_form': mat,
        'uneven_batch': uneven,
        'learned_screening': learned,
        'free_intensities': free,
        'chosen_over_plain': value / plain,
        'roofline': roof,
        'cpu_baseline': cpu,
    }
    if cpu:
        out['chosen_over_cpu'] = value / cpu['value']
    print(json.dumps(out))


if __name__ == '__main__':
    main()
