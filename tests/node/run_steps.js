'use strict';
// Drives the Node host API: steps a small campfire and writes raw fields + diagnostics for the
// parity test (tests/test_node_addon.py).  Usage: node run_steps.js OUTDIR NSTEPS [SPOTS_JSON]
// With SPOTS_JSON ({spots: [[x, y, z, r, flux, temp], ...], face: [...] | null, params: {...}, dt}) the
// grid has no logs and is fed by the spots and the floor flux instead.
const fs = require('fs');
const path = require('path');
const { FireSim } = require(path.join(__dirname, '..', '..', 'procedural-fire_amd', 'node', 'firesim.js'));

async function main() {
  const out = process.argv[2];
  const n = Number(process.argv[3] || 10);
  const scene = process.argv[4] ? JSON.parse(fs.readFileSync(process.argv[4], 'utf8')) : null;
  if (scene) {
    const sim = new FireSim({ res: scene.res, size: scene.size, logs: new Float32Array(0), noiseN: 32, useGraph: true });
    sim.setParams(scene.params);
    const norms = sim.setSpots(scene.spots.map(r => ({ pos: r.slice(0, 3), radius: r[3], flux: r[4], temp: r[5] })));
    if (scene.face) sim.setFaceFlux(Float32Array.from(scene.face));
    await sim.stepAsync(n, scene.dt);
    for (const f of ['vel', 'S', 'S2']) fs.writeFileSync(path.join(out, `${f}.bin`), Buffer.from(sim.read(f, true).buffer));
    fs.writeFileSync(path.join(out, 'meta.json'), JSON.stringify({ frame: sim.frame, norms: Array.from(norms) }));
    console.log('node spot steps done', sim.frame);
    return;
  }
  const sim = new FireSim({ res: [32, 48, 32], size: [1.2, 1.8, 1.2], noiseN: 32, useGraph: true });
  sim.setParams({ vortEps: 4, intensity: 1.2 });
  let threw = false;
  try { sim.setParams({ notAParam: 1 }); } catch (e) { threw = /unknown uniform/.test(String(e.message)); }
  await sim.stepAsync(n, 1 / 90);
  for (const f of ['vel', 'S', 'S2']) fs.writeFileSync(path.join(out, `${f}.bin`), Buffer.from(sim.read(f, true).buffer));
  const meta = { time: sim.time, frame: sim.frame, threwOnUnknown: threw, stats: sim.stats('S'), profile: sim.profile(),
    res: sim.res, dx: sim.dx, levels: sim.levels.length, coarseIters: sim.coarseIters };
  fs.writeFileSync(path.join(out, 'meta.json'), JSON.stringify(meta));
  console.log('node steps done', meta.frame);
}
main().catch(e => { console.error(e); process.exit(1); });
