'use strict';
// Drives the Node capture loop (procedural-fire_amd/node/driver.js) over a small campfire with sparks and a
// compressed cue sheet (a hero pop, a log shift, a gust, a crackle burst in the first 0.3 s), then writes the
// fields and sparks for the parity test (tests/test_driver.py).  Usage: node run_capture.js OUTDIR [FRAMES]
// With OUTDIR "-" it prints the cue-sheet functions on a time grid as JSON instead.
const fs = require('fs');
const path = require('path');
const root = path.join(__dirname, '..', '..', 'procedural-fire_amd', 'node');
const D = require(path.join(root, 'driver.js'));

const out = process.argv[2];
if (out === '-') {
  const ts = [];
  for (let k = -1200; k <= 4900; k++) ts.push(k / 240);
  const rows = ts.map(t => { const p = D.simAt(t); const o = D.logOffsetsAt(t);
    return [t, p.intensity, p.burstAmt, p.burstX, p.burstY, p.burstZ, p.windX, p.windZ, o ? o.s : -1]; });
  const logs = D.applyLogOffsets(require(path.join(root, 'firesim.js')).buildCampfireLogs(7), D.logOffsetsAt(9.06));
  fs.writeSync(1, JSON.stringify({ events: D.EVENTS, rows, logs: Array.from(logs) }) + '\n');   // whole, then exit
  process.exit(0);
}
const { FireSim, Sparks, buildCampfireLogs } = require(path.join(root, 'firesim.js'));
const frames = Number(process.argv[3] || 6);
const events = [
  { t: 0.05, type: 'pop', strength: 1.0, pos: [0.06, 0.14, 0.12], aim: [0.5, 0, 0.86] },
  { t: 0.1, type: 'log-shift', strength: 1.0, pos: [0.05, 0.12, 0.0], log: 3 },
  { t: 0.0, type: 'gust', strength: 1.0, pos: [0, 0.3, 0], dur: 0.2 },
  { t: 0.15, type: 'crackle-burst', strength: 0.5, pos: [-0.10, 0.08, 0.06] },
  { t: 0.2, type: 'pop', strength: 0.4, pos: [0.02, 0.1, -0.05] },
].sort((a, b) => a.t - b.t);
const logs = buildCampfireLogs(7);
const sim = new FireSim({ res: [32, 48, 32], size: [1.2, 1.8, 1.2], logs, noiseN: 32, useGraph: true });
const sparks = new Sparks(sim, { count: 1024 });
const cap = new D.Capture(sim, sparks, logs, { simDt: 1 / 240, events });
cap.begin({ preroll: 0.1 });
for (let i = 0; i < frames; i++) cap.frame(i);
for (const f of ['vel', 'S', 'S2', 'geo']) fs.writeFileSync(path.join(out, `${f}.bin`), Buffer.from(sim.read(f, true).buffer));
fs.writeFileSync(path.join(out, 'sparks.bin'), Buffer.from(sparks.read().buffer));
fs.writeFileSync(path.join(out, 'meta.json'), JSON.stringify({ n: cap.n, simDt: cap.simDt, logUploads: cap.logUploads,
  frame: sim.frame }));
console.log('node capture done', sim.frame);
