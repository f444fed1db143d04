'use strict';
// Drives the Node Render class: a small campfire stepped 40 times, then two frames of the render-volume
// passes with the final-render look, a thin-lens camera, jitter and a surface depth buffer; writes the
// volume, aux and lights buffers for the parity test (tests/test_node_addon.py).
// Usage: node run_render.js OUTDIR
const fs = require('fs');
const path = require('path');
const { FireSim, Render, orbit, buildCampfireLogs } = require(path.join(__dirname, '..', '..', 'procedural-fire_amd', 'node', 'firesim.js'));

const out = process.argv[2];
const sim = new FireSim({ res: [32, 48, 32], size: [0.6, 0.9, 0.6], logs: buildCampfireLogs(7), noiseN: 32 });
sim.step(1 / 90, 40);
const r = new Render(sim, { width: 96, height: 128, volScale: 0.5 });
r.setParams({ cubic: 1, flameSharp: 1, detailAmt: 0.3, sootEdge: 0.12, lightK: 0.6, focus: 2.2, lensX: 0.01 });
let threw = false;
try { r.setParams({ nope: 1 }); } catch (e) { threw = /unknown render parameter/.test(String(e.message)); }
r.setCamera(orbit({ yaw: 0.35, pitch: 0.12, dist: 2.2, target: [0, 0.5, 0], fov: 38 }));
const depth = new Float32Array(20 * 30);
for (let i = 0; i < depth.length; i++) depth[i] = 1.8 + (i % 7) * 0.1;
r.setDepth(depth, 20, 30);
for (let f = 0; f < 2; f++) {
  r.frame(sim.time, 5 + f, [0.25, -0.125]);
  sim.step(1 / 90);
}
fs.writeFileSync(path.join(out, 'vol.bin'), Buffer.from(r.readVol().buffer));
fs.writeFileSync(path.join(out, 'aux.bin'), Buffer.from(r.readAux().buffer));
fs.writeFileSync(path.join(out, 'lights.bin'), Buffer.from(r.readLights().buffer));
fs.writeFileSync(path.join(out, 'meta.json'), JSON.stringify({ vw: r.vw, vh: r.vh, threwOnUnknown: threw }));
console.log('node render done', r.vw, r.vh);
