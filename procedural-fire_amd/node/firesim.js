'use strict';
// Node.js host API of the gfx950 fire simulation: a drop-in for the reference's WebGPU
// `FireSim` (src/sim.js:48-649) and the sim-facing part of `window.FIRE` (src/main.js:242-355),
// over the C ABI (include/firesim.h) through the N-API addon lib/firesim.node.
//
// Differences from the browser class, by construction of the boundary:
//   * no GPUDevice / GPUCommandEncoder arguments: the handle owns its HIP stream;
//     step(dt) enqueues (like encoding + submit), sync() waits (onSubmittedWorkDone);
//   * the noise/curl textures are baked by the handle from `noiseSeed` (src/main.js:36-37);
//   * textures are read back as typed arrays in the same texel order (x fastest, RGBA).
// Written for Node >= 12 (no optional chaining).
const path = require('path');

const addon = require(path.join(__dirname, '..', 'lib', 'firesim.node'));

const SIM_FIELDS = addon.paramNames();                       // src/sim.js:14-23
const GRID_KEYS = new Set(['nx', 'ny', 'nz', 'dx', 'dt', 'time', 'frame', 'ox', 'oy', 'oz', 'nLogs']);
const SIM_DEFAULTS = (() => {                                 // src/sim.js:25-44
  const d = addon.paramDefaults();
  const o = {};
  SIM_FIELDS.forEach((k, i) => { if (!GRID_KEYS.has(k) && k !== 'seed') o[k] = d[i]; });
  return o;
})();

const PRECISION = { fp32: 0, fp16: 1, bf16: 2 };
const SOLVE = { vcycles: 0, refine64: 1, tol32: 2, spectral64: 3 };
const FIELD = { vel: 0, S: 1, S2: 2, vort: 3, geo: 4, velX: 5, SX: 6, S2X: 7, velY: 8, SY: 9, S2Y: 10,
  p: 11, b: 12, noise: 13, curl: 14 };
const STAGE_LABELS = ['adv1', 'adv2', 'curl', 'forces', 'div', 'mg', 'project'];

// buildCampfire(seed).logs packed by packLogs (src/scene.js:21-42, 92-98): 16 records of 8 floats.
function buildCampfireLogs(seed = 7, max = 16) { return addon.buildCampfire(seed, max); }

function fieldId(name) {
  if (typeof name === 'number') return name;
  if (!(name in FIELD)) throw new Error(`unknown field ${name}`);
  return FIELD[name];
}

class FireSim {
  constructor({ res = [80, 120, 80], size = [1.2, 1.8, 1.2], logs, vcycles = 2, maxLevels = 0,
    precision = 'fp16', solve = 'vcycles', divTol = 1e-5, maxVcycles = 8, noiseN = 128, noiseSeed = 1,
    device = 0, useGraph = true, adaptive = true, omega = 0 } = {}) {
    if (!logs) logs = buildCampfireLogs(7);
    this.h = addon.create({ res, size, logs, vcycles, maxLevels, precision: PRECISION[precision], solve: SOLVE[solve],
      divTol, maxVcycles, noiseN, noiseSeed, device, useGraph: useGraph ? 1 : 0, adaptive: adaptive ? 1 : 0, omega });
    const info = addon.info(this.h);
    this.res = info.res;
    this.size = size;
    this.dx = info.dx;                                         // src/sim.js:53
    this.origin = info.origin;                                 // src/sim.js:54
    this.vcycles = vcycles;
    this.levels = info.levels.map(dims => ({ dims }));
    this.coarseIters = info.coarseIters;                       // src/sim.js:88-92
    this.profileStages = false;
  }

  get time() { return addon.info(this.h).time; }
  get frame() { return addon.info(this.h).frame; }

  // the current parameter block as an object (src/sim.js:59)
  get params() {
    const v = addon.getParams(this.h);
    const o = {};
    SIM_FIELDS.forEach((k, i) => { o[k] = v[i]; });
    return o;
  }

  // setParams(p) (src/sim.js:601-604): unknown keys throw and change nothing
  setParams(p) {
    const names = Object.keys(p);
    addon.setParams(this.h, names, names.map(k => Number(p[k])));
  }

  // the log buffer rewrite + geoReady = false of applyLogs (src/main.js:147-149)
  setLogs(logs) { addon.setLogs(this.h, logs); }

  // pyrolysis spots (build extension): [{pos: [x, y, z], radius, flux (kg/s), temp}], [] removes them;
  // returns each spot's mass normalisation 1 / (rho0 * sum(w) * dx^3)
  setSpots(spots) {
    const f = new Float32Array(spots.length * 6);
    spots.forEach((s, i) => { f.set([s.pos[0], s.pos[1], s.pos[2], s.radius, s.flux, s.temp], i * 6); });
    return addon.setSpots(this.h, f);
  }

  // floor fuel mass flux (kg/m^2/s) as a Float32Array of nz*nx (x fastest); null removes it
  setFaceFlux(flux) { addon.setFaceFlux(this.h, flux); }

  reset() { addon.reset(this.h); }                             // src/sim.js:606-618

  // step(dt): enqueue one step (src/sim.js:627-648); n > 1 enqueues n steps
  step(dt, n = 1) {
    if (this.profileStages) addon.profile(this.h, true);
    addon.step(this.h, dt, n);
  }

  sync() { addon.sync(this.h); }

  // FIRE.step(n, dt) (src/main.js:244): n steps, resolved once the device has finished
  stepAsync(n = 1, dt = 1 / 90) { return addon.stepAsync(this.h, dt, n); }

  read(name, native = false) { return addon.readField(this.h, fieldId(name), native); }
  write(name, arr) { addon.writeField(this.h, fieldId(name), arr); }

  residual() { return addon.residual(this.h); }

  // per-kernel device ms of the last profiled step (Profiler labels, src/sim.js:629-644)
  timings() {
    const t = addon.timings(this.h);
    const o = {};
    STAGE_LABELS.forEach((k, i) => { o[k] = t[i]; });
    o.adv = o.adv1 + o.adv2;
    return o;
  }

  // FIRE.stats(name) (src/main.js:268-294), reduced on the device: per-channel
  // min/max/mean/bad/argmax with the reference's toPrecision(4) display rounding
  stats(name = 'S') {
    const [x, y] = this.res;
    const a = addon.stats(this.h, fieldId(name));
    const out = [];
    for (let c = 0; c < 4; c++) {
      const [mn, mx, mean, bad, arg] = a.subarray(c * 5, c * 5 + 5);
      const i = arg < 0 ? null : [arg % x, Math.floor(arg / x) % y, Math.floor(arg / (x * y))];
      out.push({ min: +mn.toPrecision(4), max: +mx.toPrecision(4), mean: +mean.toPrecision(4), bad, argmax: i });
    }
    return out;
  }

  // FIRE.profile() (src/main.js:322-355), reduced on the device: heat release (kW) and
  // centreline/height rows, rounded as the reference rounds them
  profile() {
    const a = addon.profileRows(this.h);
    const rows = [];
    for (let r = 0; 1 + r * 6 < a.length; r++) {
      const o = 1 + r * 6;
      rows.push({ h: +a[o].toFixed(2), Tc: Math.round(a[o + 1]), Tmax: Math.round(a[o + 2]), soot: +a[o + 3].toFixed(2),
        fuel: +a[o + 4].toFixed(2), hotCells: a[o + 5] });
    }
    return { Q_kW: +(a[0] / 1000).toFixed(1), rows };
  }

  // per-step diagnostics (time, heat release W, max T, flame tip m, axis T per y-plane)
  diagEnable(capacity = 4096, tFlame = 1000) { this._diagCap = capacity; addon.diagEnable(this.h, capacity, tFlame); }
  diagSeries() {
    const ny = this.res[1], rec = 4 + ny;
    const a = addon.diagRead(this.h, this._diagCap || 0);
    const n = a.length / rec;
    const o = { time: [], Q: [], Tmax: [], tip: [], axisT: [] };
    for (let k = 0; k < n; k++) {
      o.time.push(a[k * rec]); o.Q.push(a[k * rec + 1]); o.Tmax.push(a[k * rec + 2]); o.tip.push(a[k * rec + 3]);
      o.axisT.push(Array.from(a.subarray(k * rec + 4, (k + 1) * rec)));
    }
    return o;
  }
}

// Spark particles carried by a FireSim's velocity: the reference's Sparks (src/sparks.js:17-271) minus
// drawing.  step(dt) right after the sim's step with the same dt; read() returns count x 16 floats
// (position+age, velocity+burn time, previous position+T, diameter/seed/alive/flicker).
class Sparks {
  constructor(sim, { count = 8192 } = {}) {
    this.sim = sim;              // kept alive while the sparks are
    this.count = count;
    this.h = addon.sparksCreate(sim.h, count);
    this._params = { rate: 3.0, turb: 1.0, lifeK: 1.0, sizeK: 1.0, upK: 1.0, brightness: 13.0 };
  }
  get rate() { return this._params.rate; }
  set rate(v) { this.setParams({ rate: v }); }
  get params() { return { ...this._params }; }
  setParams(p) {
    for (const [k, v] of Object.entries(p)) { addon.sparksSet(this.h, k, v); this._params[k] = v; }
  }
  reset() { addon.sparksReset(this.h); }                                  // src/sparks.js:38-41
  burst(opts = {}) { addon.sparksBurst(this.h, opts); }                    // src/sparks.js:44-46
  step(dt) { addon.sparksStep(this.h, dt); }                               // src/sparks.js:241-271
  read() { return addon.sparksRead(this.h, this.count); }
}

// Look-at camera (src/camera.js:1-29), for Render.setCamera.
const vsub = (a, b) => [a[0] - b[0], a[1] - b[1], a[2] - b[2]];
const vcross = (a, b) => [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]];
const vnorm = (a) => { const l = Math.hypot(a[0], a[1], a[2]) || 1; return [a[0] / l, a[1] / l, a[2] / l]; };
function lookAt(pos, target, fovYdeg = 40, roll = 0) {
  const fwd = vnorm(vsub(target, pos));
  let right = vnorm(vcross(fwd, [0, 1, 0]));
  let up = vcross(right, fwd);
  if (roll) {
    const c = Math.cos(roll), s = Math.sin(roll);
    const r2 = right.map((v, i) => v * c + up[i] * s);
    const u2 = up.map((v, i) => v * c - right[i] * s);
    right = r2; up = u2;
  }
  return { pos, target, fwd, right, up, fovY: (fovYdeg * Math.PI) / 180 };
}
function orbit({ yaw = 0, pitch = 0.15, dist = 2.4, target = [0, 0.45, 0], fov = 40 } = {}) {
  const pos = [target[0] + dist * Math.cos(pitch) * Math.sin(yaw), target[1] + dist * Math.sin(pitch),
    target[2] + dist * Math.cos(pitch) * Math.cos(yaw)];
  return lookAt(pos, target, fov);
}

// The Renderer passes that read a FireSim's S / S2 (src/renderer.js:104-326): fire-light extraction,
// flame-field prefilter and the volume ray-march.  params / setParams mirror `renderer.params`;
// frame(time, frame, jitter) is uploadParams (src/renderer.js:634-648) followed by the lights,
// prefilter and volume passes of renderHDR (src/renderer.js:668-690).  The surface pass, composite
// and post stay with the caller: pass its depth with setDepth, take vol (L.rgb, Tr) and aux
// (haze, depth) back with readVol / readAux.
const RP_COUNT = 88;
const RENDER_BUF = { vol: 0, aux: 1, lights: 2, sharp: 3, lut: 4 };
class Render {
  constructor(sim, { width = 960, height = 1080, volScale = 0.5 } = {}) {
    this.sim = sim;              // kept alive while the renderer is
    this.h = addon.renderCreate(sim.h, width, height, volScale);
    [this.vw, this.vh] = addon.renderSize(this.h);
    this._params = {};
  }
  get params() { return { ...this._params }; }
  setParams(p) {
    for (const [k, v] of Object.entries(p)) { addon.renderSet(this.h, k, v); this._params[k] = v; }
  }
  setCamera(cam) { addon.renderSetCamera(this.h, cam); }                  // src/renderer.js:632
  setWhiteBalance(twb) { addon.renderSetWhiteBalance(this.h, twb); }      // src/renderer.js:57-63
  setDepth(depth, dw, dh) { addon.renderSetDepth(this.h, depth || null, dw || 0, dh || 0); }
  frame(time, frame, jitter = [0, 0]) { addon.renderFrame(this.h, time, frame, jitter[0], jitter[1]); }
  uniforms() { return addon.renderUniforms(this.h); }
  readVol() { return addon.renderRead(this.h, RENDER_BUF.vol, this.vw * this.vh * 4); }
  readAux() { return addon.renderRead(this.h, RENDER_BUF.aux, this.vw * this.vh * 2); }
  readLights() { return addon.renderRead(this.h, RENDER_BUF.lights, 64); }
  readSharp() { const [x, y, z] = this.sim.res; return addon.renderRead(this.h, RENDER_BUF.sharp, x * y * z * 4); }
}

module.exports = { FireSim, Sparks, Render, lookAt, orbit, RP_COUNT, SIM_FIELDS, SIM_DEFAULTS, PRECISION, SOLVE, FIELD,
  STAGE_LABELS, buildCampfireLogs, addon };
