// Drives powder6_amd/node/pt_hip.js from a JSON PT context (tests/js_ctx.py); prints one
// JSON line.  Usage: node pt_hip_probe.js ctx.json compile|costs|run
//
// ctx.cost (the caller's FP64 model) is, when ctx.costPipe = {req, resp} names two FIFOs,
// a synchronous request to the test process, which answers with the FP64 oracle's cost
// of the vector; else a lookup in ctx.expectedCosts ([...x, cost] rows, NaN otherwise).
// ctx.costOffset is added to every cost (the reference page's forced disagreement,
// tests/webgpu-pt.js:143-148); ctx.stopAfter makes shouldStop() true from its n-th call.
'use strict';
var fs = require('fs'), path = require('path');
var RVPTGPU = require(path.join(__dirname, '..', '..', 'powder6_amd', 'node', 'pt_hip.js'));
var ctx = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
var mode = process.argv[3] || 'compile';
ctx.info.forEach(function (i) { if (i.lo === null) i.lo = -Infinity; if (i.hi === null) i.hi = Infinity; });
var table = ctx.expectedCosts || [], offset = ctx.costOffset || 0, calls = 0;
var pipe = null;
if (ctx.costPipe) {
    // same open order as the server: requests first, then answers
    pipe = { req: fs.openSync(ctx.costPipe.req, 'w'), resp: fs.openSync(ctx.costPipe.resp, 'r'), buf: '' };
}
function pipeCost(x) {
    fs.writeSync(pipe.req, JSON.stringify(Array.from(x)) + '\n');
    var chunk = Buffer.alloc(256);
    while (pipe.buf.indexOf('\n') < 0) {
        var n = fs.readSync(pipe.resp, chunk, 0, chunk.length, null);
        if (n <= 0) throw new Error('cost server closed');
        pipe.buf += chunk.toString('utf8', 0, n);
    }
    var k = pipe.buf.indexOf('\n'), line = pipe.buf.slice(0, k);
    pipe.buf = pipe.buf.slice(k + 1);
    return parseFloat(line);
}
ctx.cost = function (x) {
    calls++;
    if (pipe) return pipeCost(x) + offset;
    for (var r = 0; r < table.length; r++) {
        var row = table[r], same = row.length === x.length + 1;
        for (var j = 0; same && j < x.length; j++) same = row[j] === x[j];
        if (same) return row[x.length] + offset;
    }
    return NaN;
};
if (ctx.stopAfter) {
    var asked = 0;
    ctx.shouldStop = function () { asked++; return asked >= ctx.stopAfter; };
}
function out(o) {
    if (pipe) { fs.closeSync(pipe.req); fs.closeSync(pipe.resp); pipe = null; }
    process.stdout.write(JSON.stringify(o) + '\n');
}
if (mode === 'compile') {
    out({ compile: RVPTGPU.compile(ctx) });
} else if (mode === 'flat') {
    // the C-ABI problems the host flattens the context into (model-parameter terms)
    var flat = RVPTGPU.fromContext(ctx), list = Array.isArray(flat) ? flat : [flat];
    out({ flat: list.map(function (f) {
        return { paramBase: Array.from(f.paramBase), termParam: Array.from(f.termParam), termVar: Array.from(f.termVar),
                 termCoeff: Array.from(f.termCoeff), resolutionType: f.resolutionType, radiation: f.radiation,
                 phases: f.phases.map(function (p) { return { free: p.freeIntensities, atoms: p.atoms.length }; }) };
    }) });
} else if (mode === 'cost') {
    // ctx.cost of the start vector and of the probes, without a device
    out({ cost: [ctx.x].concat(ctx.probes || []).map(function (v) { return ctx.cost(v); }) });
} else if (mode === 'costs') {
    RVPTGPU.create(ctx).then(function (g) {
        return g.costs(ctx.probes).then(function (c) { g.close(); out({ costs: c, plan: g.plan }); },
                                        function (e) { g.close(); throw e; });
    }).catch(function (e) { out({ error: e.message }); process.exitCode = 3; });
} else if (mode === 'normal') {
    RVPTGPU.create(ctx).then(function (g) {
        return g.normal(ctx.x).then(function (r) { g.close(); out({ A: Array.from(r.A), g: Array.from(r.g), f: r.f }); },
                                    function (e) { g.close(); throw e; });
    }).catch(function (e) { out({ error: e.message }); process.exitCode = 3; });
} else if (mode === 'run') {
    var iters = 0;
    ctx.onIter = function () { iters++; };
    RVPTGPU.run(ctx, { tolerance: ctx.tolerance || 1e-3 }).then(function (r) {
        out({ x: Array.from(r.x), stopped: r.stopped, pt: r.pt, iters: iters, costCalls: calls });
    }).catch(function (e) { out({ error: e.message, costCalls: calls }); process.exitCode = 3; });
}
