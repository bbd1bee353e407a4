// Drives powder6_amd/node/pt_hip.js from a JSON PT context (tests/js_ctx.py); prints one
// JSON line.  Usage: node pt_hip_probe.js ctx.json compile|costs|run
'use strict';
var fs = require('fs'), path = require('path');
var RVPTGPU = require(path.join(__dirname, '..', '..', 'powder6_amd', 'node', 'pt_hip.js'));
var ctx = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
var mode = process.argv[3] || 'compile';
ctx.info.forEach(function (i) { if (i.lo === null) i.lo = -Infinity; if (i.hi === null) i.hi = Infinity; });
var table = ctx.expectedCosts || [];
ctx.cost = function (x) {
    // FP64 costs precomputed by the oracle for known vectors ([...x, cost] rows); NaN
    // (skipped by run()) for any other vector
    for (var r = 0; r < table.length; r++) {
        var row = table[r], same = row.length === x.length + 1;
        for (var j = 0; same && j < x.length; j++) same = row[j] === x[j];
        if (same) return row[x.length];
    }
    return NaN;
};
function out(o) { process.stdout.write(JSON.stringify(o) + '\n'); }
if (mode === 'compile') {
    out({ compile: RVPTGPU.compile(ctx) });
} else if (mode === 'costs') {
    RVPTGPU.create(ctx).then(function (g) {
        return g.costs(ctx.probes).then(function (c) { g.close(); out({ costs: c, plan: g.plan }); });
    }).catch(function (e) { out({ error: e.message }); process.exitCode = 3; });
} else if (mode === 'run') {
    var iters = 0;
    ctx.onIter = function () { iters++; };
    RVPTGPU.run(ctx, { tolerance: ctx.tolerance || 1e-3 }).then(function (r) {
        out({ x: Array.from(r.x), stopped: r.stopped, pt: r.pt, iters: iters });
    }).catch(function (e) { out({ error: e.message }); process.exitCode = 3; });
}
