/* pt_hip.js -- MI355X parallel-tempering backend with the RVPTGPU surface of
 * js/pt_gpu.js:424 ({compile, create, run}), over the N-API addon pt_addon.node.
 *
 * Written in Node 12 syntax (no ?? or ?.), loadable with require() or importScripts-like
 * evaluation.  rietveld_worker.js would load it where it loads pt_gpu.js
 * (js/rietveld_worker.js:15-17) and RVCore.refineAsync would call run(ctx)
 * (js/rietveld_core.js:2455-2459); any throw there makes the caller fall back to CPU PT.
 *
 * ctx is the context RVCore yields (js/rietveld_core.js:2271-2273): {problem, model, x,
 * info, setup, constraints, cost, feasible, shouldStop, onIter, maxIter}.  fromContext()
 * flattens it into the C-ABI problem (include/powder6_pt.h); a ctx may instead carry
 * that flat problem directly as ctx.flat.
 */
'use strict';
var path = require('path');

var addon = null;
function native() {
    if (!addon) {
        try {
            addon = require(path.join(__dirname, '..', '_lib', 'pt_addon.node'));
        } catch (e) {
            throw new Error('MI355X PT addon is unavailable in this process or context: ' + e.message);
        }
    }
    return addon;
}

var SYSTEM = { triclinic: 0, monoclinic: 1, orthorhombic: 2, tetragonal: 3, trigonal: 4, hexagonal: 5, rhombohedral: 5, cubic: 6 };
var PROFILE = { simple_pvoigt: 1, split_pvoigt: 2, tch_aniso: 3, gsas_cw: 4 };
/* the dataset's instrument slots of separated profiles and their defaults
 * (js/rietveld_core.js:413, 454-459), and the sample slots of a separated phase
 * (:414-415 SAMPLE_PROFILE order) */
var INST = ['U', 'V', 'W', 'X', 'Y', 'Z', 'SL', 'HL', 'eta', 'U_L', 'V_L', 'W_L', 'X_L', 'U_R', 'V_R', 'W_R', 'X_R', 'eta_split'];
var INST_DEFAULT = [2, -2, 5, 1, 0, 0, 0, 0, 0.5, 2, -2, 5, 1, 2, -2, 5, 1, 0.5];
var SAMPLE = ['size', 'sizeAxial', 'sizeB', 'sizeC', 'rotA', 'rotB', 'rotC', 'strain', 'strainAxial', 'sizeMix', 'strainMix'].concat(
    ['400', '040', '004', '220', '202', '022', '310', '301', '130', '103', '031', '013', '211', '121', '112'].map(function (k) { return 'S' + k; }));
var SAMPLE_DEFAULT = { size: 1, sizeAxial: 1, sizeB: 1, sizeC: 1, sizeMix: 1, strainMix: 1 };
var SIZE_MODEL = { isotropic: 0, uniaxial: 1, ellipsoid: 2 }, STRAIN_MODEL = { isotropic: 0, uniaxial: 1, stephens: 2 };
var SLOTS = {
    tch_aniso: ['U', 'V', 'W', 'X', 'Y', 'SL', 'HL'],
    simple_pvoigt: ['GU', 'GV', 'GW', 'GP', 'LX', 'eta'],
    split_pvoigt: ['GU_L', 'GV_L', 'GW_L', 'LX_L', 'GU_R', 'GV_R', 'GW_R', 'LX_R', 'eta_split']
};
var CELL = ['a', 'b', 'c', 'alpha', 'beta', 'gamma'];
var P_INST0 = 27, P_PHASE0 = 45, PH_PROF0 = 8, PH_ATOM0 = 40, AT_STRIDE = 5;

/* Parameter index of a reference value name in the C-ABI layout, or -1.  The 24
 * background slots hold bkg<k> (Chebyshev) or bga<j> (spline anchor values,
 * js/rietveld_core.js:932-936), as the dataset's background form says. */
function paramIndex(layout, name, spline) {
    if (name === 'zero') return 0;
    if (name === 'disp') return 1;
    if (name === 'transp') return 2;
    var m = (spline ? /^bga(\d+)$/ : /^bkg(\d+)$/).exec(name);
    if (m) return +m[1] < 24 ? 3 + (+m[1]) : -1;
    if (name.indexOf('inst.') === 0) { var ii = INST.indexOf(name.slice(5)); return ii >= 0 ? P_INST0 + ii : -1; }
    for (var q = 0; q < layout.length; q++) {
        var L = layout[q];
        if (name.indexOf(L.prefix) !== 0) continue;
        var rest = name.slice(L.prefix.length);
        var c = CELL.indexOf(rest);
        if (c >= 0) return L.base + c;
        if (rest === 'scale') return L.base + 6;
        if (rest === 'po') return L.base + 7;
        var s = L.slots.indexOf(rest);
        if (s >= 0) return L.base + PH_PROF0 + s;
        var am = /^(.+)\.(x|y|z|occ|B)$/.exec(rest);
        if (am) {
            var j = L.labels.indexOf(am[1]);
            if (j >= 0) return L.base + PH_ATOM0 + AT_STRIDE * j + ['x', 'y', 'z', 'occ', 'B'].indexOf(am[2]);
        }
    }
    return -1;
}

/* Flatten a reference PT context (Rietveld phases, Chebyshev or spline background, TCH /
 * simple / split pseudo-Voigt, separated GSAS profiles) into the C-ABI problem: one flat
 * problem, or one per dataset (all with the context's variable list) when the context
 * holds several datasets. */
function fromContext(ctx) {
    if (ctx.flat) return ctx.flat;
    var model = ctx.model;
    if (!model || !model.datasets.length || model.datasets.length > 8) throw new Error('MI355X PT backend takes 1 to 8 datasets per fit; using the CPU path.');
    if (model.datasets.length === 1) return flattenDataset(ctx, 0);
    return model.datasets.map(function (_, d) { return flattenDataset(ctx, d); });
}

/* dataset d of the context: its phases (D.has), its value map problem.values[d], and the
 * variables whose dataset list ctx.info[j].ds names d (js/pt_gpu.js:25 slots) */
function flattenDataset(ctx, d) {
    var model = ctx.model, problem = ctx.problem;
    var D = model.datasets[d], def = D.def, multi = model.datasets.length > 1;
    var spline = !!D.spline;
    if (spline && D.bkgX.length > 24) throw new Error('MI355X PT backend takes at most 24 spline anchors; using the CPU path.');
    var values = problem.values[d], layout = [], base = P_PHASE0, phases = [];
    model.phases.forEach(function (P, q) {
        if (!D.has[q]) {
            if (!multi) throw new Error('MI355X PT backend needs every phase in the dataset.');
            return;
        }
        // Le Bail and WPPF phases carry free reflection intensities, solved per trial
        if (P.method !== 'rietveld' && P.method !== 'lebail' && P.method !== 'wppf') throw new Error('MI355X PT backend cannot evaluate a ' + P.method + ' phase; using the CPU path.');
        var type = P.def.profileType;
        // rietveld_core.js:460 usesSeparatedProfile, :416-421 broadeningConfig
        var sep = type === 'gsas_cw' || P.def.separatedProfile === true, b = P.def.broadening || {};
        var axis = function (a) { return Array.isArray(a) && a.length === 3 && a.every(isFinite) && a.some(function (v) { return v !== 0; }) ? a.map(Number) : [0, 0, 1]; };
        if (!PROFILE[type]) throw new Error('unknown profile type ' + type);
        var atoms = P.def.atoms || [];
        layout.push({ prefix: P.prefix, base: base, slots: sep ? SAMPLE : SLOTS[type], sep: sep, labels: atoms.map(function (a) { return a.label; }) });
        var opR = [], opT = [];
        P.def.ops.forEach(function (o) { for (var i = 0; i < 9; i++) opR.push(o.r[i]); for (var k = 0; k < 3; k++) opT.push(o.t[k]); });
        var anom = (D.anom && D.anom[q]) || [];
        phases.push({
            system: SYSTEM[String(P.def.system || 'triclinic').toLowerCase()] || 0,
            uniqueAxis: { a: 0, b: 1, c: 2 }[P.def.uniqueAxis || 'b'],
            profileType: PROFILE[type], opR: Int32Array.from(opR), opT: Float64Array.from(opT),
            separated: sep ? 1 : 0, sizeModel: SIZE_MODEL[b.size] || 0, strainModel: STRAIN_MODEL[b.strain] || 0,
            freeIntensities: P.method === 'rietveld' ? 0 : 1,
            sizeAxis: axis(b.sizeAxis), strainAxis: axis(b.strainAxis),
            poDir: P.def.poDir || [0, 0, 1],
            atoms: atoms.map(function (at, j) {
                var site = at.site, basis = [0, 0, 0, 0, 0, 0, 0, 0, 0], mask = 0;
                (site.free || []).forEach(function (f) { mask |= 1 << f; for (var i = 0; i < 3; i++) basis[3 * f + i] = site.basis[f][i]; });
                var xr = (at.ff && at.ff.xray) || { a: [], b: [], c: 0 }, an = anom[j] || [0, 0];
                return { r0: site.r0, basis: basis, freeMask: mask, mult: site.mult || 1, ffA: xr.a, ffB: xr.b, ffC: xr.c || 0,
                    bNeutron: (at.ff && at.ff.neutron) || 0, fp: an[0], fpp: an[1] };
            })
        });
        base += PH_ATOM0 + AT_STRIDE * atoms.length;
    });
    var nParams = base, paramBase = new Float64Array(nParams);
    // defaults where the reference reads a missing or non-finite value as one
    // (rietveld_core.js:700-701 instrument values, profile.js:389 value(n, d) sample values)
    for (var ii = 0; ii < INST.length; ii++) paramBase[P_INST0 + ii] = INST_DEFAULT[ii];
    layout.forEach(function (L) {
        if (L.sep) SAMPLE.forEach(function (n, j) { if (SAMPLE_DEFAULT[n]) paramBase[L.base + PH_PROF0 + j] = SAMPLE_DEFAULT[n]; });
    });
    Object.keys(values).forEach(function (name) {
        var k = paramIndex(layout, name, spline);
        if (k >= 0 && Number.isFinite(values[name])) paramBase[k] = values[name];
    });
    // a monoclinic or triclinic cell reads all six slots; default the unused angles to 90
    layout.forEach(function (L) { for (var c = 3; c < 6; c++) if (!paramBase[L.base + c]) paramBase[L.base + c] = 90; });
    var termParam = [], termVar = [], termCoeff = [];
    problem.vars.forEach(function (vr, j) {
        var ds = ctx.info[j] && ctx.info[j].ds;
        if (multi && ds && ds.indexOf(d) < 0) return;          // acts in other datasets only
        var k = paramIndex(layout, vr.name, spline);
        if (k < 0) throw new Error('MI355X PT backend cannot refine ' + vr.name + '; using the CPU path.');
        paramBase[k] = 0; termParam.push(k); termVar.push(j); termCoeff.push(1);
    });
    if (ctx.constraints) ctx.constraints.links.forEach(function (link) {
        if (multi && link.d !== undefined && link.d !== d) return;
        var k = paramIndex(layout, link.name, spline);
        if (k < 0) throw new Error('MI355X PT backend cannot link ' + link.name + '; using the CPU path.');
        paramBase[k] = link.offset;
        link.terms.forEach(function (t) { termParam.push(k); termVar.push(t[0]); termCoeff.push(t[1]); });
    });
    return {
        tth: Float64Array.from(def.tth), y: Float64Array.from(def.y), w: Float64Array.from(def.w),
        lambda1: def.lambda1, lambda2: def.lambda2 || 0, ratio: def.ratio || 0, polK: def.polK, minFwhm: def.minFwhm,
        radiation: def.radiation === 'neutron' ? 1 : 0, nBkg: spline ? D.bkgX.length : D.T.length,
        bkgType: spline ? 1 : 0, bkgX: Float64Array.from(spline ? D.bkgX : []), phases: phases, paramBase: paramBase,
        resolutionType: PROFILE[def.resolutionType] || 0,
        x0: Float64Array.from(ctx.x), lo: Float64Array.from(ctx.info.map(function (i) { return i.lo; })),
        hi: Float64Array.from(ctx.info.map(function (i) { return i.hi; })), scale: Float64Array.from(ctx.setup.scales),
        termParam: Int32Array.from(termParam), termVar: Int32Array.from(termVar), termCoeff: Float64Array.from(termCoeff),
        chains: ctx.setup.chains, sweeps: ctx.setup.sweeps, stepsPerSweep: 0, maxTemp: ctx.setup.maxTemp, seed: ctx.setup.seed
    };
}

/* compile(ctx): the plan summary without a device (js/pt_gpu.js:226 fields; counts summed
 * over the datasets) */
function compile(ctx) {
    var flat = fromContext(ctx), list = Array.isArray(flat) ? flat : [flat], out = null;
    list.forEach(function (f) {
        var r = native().planCreate(f, false), i = r.info;
        native().planDestroy(r.handle);
        if (!out) out = { nv: i.nv, chains: i.chains, stride: i.stride, pointCount: 0, reflectionCount: 0, columnCount: 0,
                          workCount: 0, datasets: list.length, paramCount: 0 };
        out.pointCount += i.pointCount; out.reflectionCount += i.reflectionCount; out.columnCount += i.lineCount;
        out.workCount += i.pointCount; out.paramCount += i.paramCount;
    });
    return out;
}

function flatten(vectors, nv) {
    var out = new Float64Array(vectors.length * nv);
    vectors.forEach(function (v, c) { for (var j = 0; j < nv; j++) out[c * nv + j] = v[j]; });
    return out;
}

/* create(ctx) -> Promise<session> (js/pt_gpu.js:371-406) */
function create(ctx) {
    return new Promise(function (resolve) {
        var A = native(), r = A.planCreate(fromContext(ctx), true), h = r.handle, plan = r.info, closed = false;
        resolve({
            plan: plan,
            initial: function (vectors) { return A.initial(h, flatten(vectors, plan.nv)); },
            costs: function (vectors) { return A.costs(h, flatten(vectors, plan.nv)).then(function (c) { return Array.from(c); }); },
            batch: function (count, bootstrap) { return A.batch(h, count, !!bootstrap); },
            evalTrials: function (form, vectors) { return A.evalTrials(h, form, flatten(vectors, plan.nv)); },
            pattern: function (x) { return A.pattern(h, Float64Array.from(x)); },
            // normal equations of one LM step at x (js/rietveld_core.js:1953-1983 normal):
            // {A: nv*nv row major, g, f}, FP64 derivative columns on the GPU
            normal: function (x) {
                return A.normal(h, Float64Array.from(x)).then(function (o) {
                    var nv = plan.nv;
                    return { A: o.subarray(0, nv * nv), g: o.subarray(nv * nv, nv * nv + nv), f: o[nv * nv + nv] };
                });
            },
            reflections: function (phase) { return A.reflections(h, phase || 0); },
            close: function () { if (!closed) { closed = true; A.planDestroy(h); } }
        });
    });
}

function reflected(value, lo, hi) {
    if (Number.isFinite(lo) && Number.isFinite(hi)) {
        var span = hi - lo;
        if (span <= 0) return lo;
        var folded = ((value - lo) % (2 * span) + 2 * span) % (2 * span);
        return lo + (folded <= span ? folded : 2 * span - folded);
    }
    return value < lo ? 2 * lo - value : value > hi ? 2 * hi - value : value;
}

/* run(ctx) (js/pt_gpu.js:409-423): validate against ctx.cost (FP64), PT sweeps on the GPU,
 * FP64 re-scoring of every new best, best-so-far kept; resolves {x, stopped, pt}. */
function run(ctx, options) {
    options = options || {};
    var tolerance = options.tolerance || 1e-3, gpu = null;
    return create(ctx).then(function (g) {
        gpu = g;
        var nv = ctx.x.length, chains = ctx.setup.chains;
        var probes = [Array.from(ctx.x), Array.from(ctx.x, function (v, j) { return reflected(v + ctx.setup.scales[j] * 0.25, ctx.info[j].lo, ctx.info[j].hi); })];
        return gpu.costs(probes).then(function (gc) {
            var validationError = 0;
            for (var c = 0; c < chains; c++) {
                var cpu = ctx.cost(probes[c % probes.length]);
                if (!Number.isFinite(cpu)) continue;
                var error = Math.abs(gc[c] - cpu) / Math.max(1, cpu);
                validationError = Math.max(validationError, error);
                if (!Number.isFinite(gc[c]) || error > tolerance)
                    throw new Error('MI355X PT cost disagrees with the double-precision model (relative error ' + error.toPrecision(3) + ').');
            }
            var best = Float64Array.from(ctx.x), bestCost = ctx.setup.initialCost, completed = 0, stopped = false, seen = new Set(), state = null;
            var stride = gpu.plan.stride;
            function inspect() {
                for (var c = 0; c < chains; c++) {
                    var b = c * stride;
                    // js/pt_gpu.js:418: a chain whose intensity solve failed ends the GPU run
                    if (state[b + 3 * nv + 9]) throw new Error('MI355X PT intensity solve did not converge; restarting CPU PT.');
                    var candidate = Array.from(state.subarray(b + 2 * nv, b + 3 * nv)), key = candidate.join(',');
                    if (seen.has(key)) continue;
                    seen.add(key);
                    if (!candidate.every(function (v, j) { return Number.isFinite(v) && v >= ctx.info[j].lo && v <= ctx.info[j].hi; })) continue;
                    if (ctx.feasible && !ctx.feasible(candidate)) continue;
                    var value = ctx.cost(candidate);
                    if (value < bestCost) { best = Float64Array.from(candidate); bestCost = value; }
                }
            }
            function loop() {
                if (completed >= ctx.setup.sweeps) return Promise.resolve();
                if (ctx.shouldStop && ctx.shouldStop()) { stopped = true; return Promise.resolve(); }
                var count = Math.min(Math.max(1, Math.ceil(8 / nv)), ctx.setup.sweeps - completed);
                return gpu.batch(count, false).then(function (s) {
                    state = s; completed += count; inspect();
                    if (ctx.onIter) ctx.onIter({ it: completed, chi2: bestCost, maxIter: ctx.maxIter, stage: 'pt', backend: 'mi355x' });
                    return loop();
                }, function (e) {
                    if (/did not converge/.test(e.message)) throw new Error('MI355X PT intensity solve did not converge; restarting CPU PT.');
                    throw e;
                });
            }
            return gpu.initial([ctx.x]).then(function () { return gpu.batch(1, true); }).then(function (s) { state = s; return loop(); }).then(function () {
                var stats = { accepted: 0, proposals: 0, swaps: 0, swapAttempts: 0 };
                for (var c = 0; c < chains; c++) {
                    var b = c * stride + 3 * nv;
                    stats.accepted += state[b + 3]; stats.proposals += state[b + 4]; stats.swaps += state[b + 5]; stats.swapAttempts += state[b + 6];
                }
                var pt = { backend: 'mi355x', chains: chains, sweeps: completed, requestedSweeps: ctx.setup.sweeps, maxTemp: ctx.setup.maxTemp,
                           seed: ctx.setup.seed, initialChi2: ctx.setup.initialCost, bestChi2: bestCost, validationError: validationError };
                Object.keys(stats).forEach(function (k) { pt[k] = stats[k]; });
                return { x: best, stopped: stopped, pt: pt };
            });
        });
    }).then(function (result) { if (gpu) gpu.close(); return result; }, function (e) { if (gpu) gpu.close(); throw e; });
}

var RVPTGPU = { compile: compile, create: create, run: run, fromContext: fromContext, FORM_PLAIN: 0, FORM_BATCHED: 1 };
if (typeof module !== 'undefined' && module.exports) module.exports = RVPTGPU;
if (typeof self !== 'undefined') self.RVPTGPU = RVPTGPU;
