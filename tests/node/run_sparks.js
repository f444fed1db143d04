'use strict';
// Drives the Node Sparks class: a small campfire and 1024 sparks with an ignite-like burst and an aimed
// pop, stepped together like src/main.js:70-76; writes the particle array for the parity test
// (tests/test_node_addon.py).  Usage: node run_sparks.js OUTDIR NSTEPS
const fs = require('fs');
const path = require('path');
const { FireSim, Sparks } = require(path.join(__dirname, '..', '..', 'procedural-fire_amd', 'node', 'firesim.js'));

const out = process.argv[2];
const n = Number(process.argv[3] || 20);
const sim = new FireSim({ res: [32, 48, 32], size: [1.2, 1.8, 1.2], noiseN: 32, useGraph: true });
const sparks = new Sparks(sim, { count: 1024 });
sparks.rate = 300;
sparks.setParams({ turb: 1.2 });
let threw = false;
try { sparks.setParams({ nope: 1 }); } catch (e) { threw = /unknown spark parameter/.test(String(e.message)); }
for (let k = 0; k < n; k++) {
  if (k === 0) sparks.burst({ n: 140, pos: [0, 0.15, 0], speed: 3.2, radius: 0.12, dur: 0.5 });
  if (k === 7) sparks.burst({ n: 40, pos: [0.05, 0.1, 0], speed: 5, radius: 0.03, dur: 0.05, fan: 0.9, size: 1.4, aim: [0.3, 0, -0.2] });
  sim.step(1 / 90);
  sparks.step(1 / 90);
}
const a = sparks.read();
fs.writeFileSync(path.join(out, 'sparks.bin'), Buffer.from(a.buffer));
let alive = 0;
for (let i = 0; i < sparks.count; i++) if (a[16 * i + 14] > 0.5) alive++;
fs.writeFileSync(path.join(out, 'meta.json'), JSON.stringify({ alive, threwOnUnknown: threw }));
console.log('node sparks done', alive);
