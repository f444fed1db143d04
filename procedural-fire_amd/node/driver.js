'use strict';
// Per-step driver inputs of the offline capture: the cue sheet's simulation overrides
// (src/timeline.js:96-157), the log-shift geometry update (src/main.js:135-150), the mapping of
// scripted events to spark bursts (src/main.js:77-87) and the deterministic capture step loop
// (src/main.js:153-182), over the native FireSim and Sparks.  Everything here runs on the host
// between steps; it decides the parameters, log buffer and bursts the next step sees.

const FPS = 30;
const DURATION = 20;
const PREROLL = 5.0;   // seconds simulated before t = 0

// fire intensity keyframes [t, value], smoothstep between neighbours (src/timeline.js:33-37)
const INTENSITY = [
  [-PREROLL, 1.0], [0.0, 1.0], [1.0, 1.0], [2.0, 1.12], [6.0, 1.12], [8.6, 1.0], [9.0, 1.2], [9.6, 1.05],
  [11.0, 1.08], [12.6, 1.08], [14.0, 1.05], [18.3, 1.05], [18.5, 1.15], [19.2, 1.0], [20.0, 1.0],
];

// scripted events of the cinematic (src/timeline.js:40-53)
const SCRIPTED = [
  { t: 0.30, type: 'pop', strength: 1.0, pos: [0.06, 0.14, 0.12], aim: [0.5, 0, 0.86] },
  { t: 2.433, type: 'pop', strength: 0.5, pos: [-0.08, 0.12, 0.10] },
  { t: 4.6, type: 'pop', strength: 0.6, pos: [0.10, 0.16, 0.02] },
  { t: 5.633, type: 'pop', strength: 0.45, pos: [-0.05, 0.10, 0.12] },
  { t: 7.2, type: 'crackle-burst', strength: 0.5, pos: [-0.10, 0.08, 0.06] },
  { t: 8.433, type: 'pop', strength: 0.5, pos: [0.06, 0.12, 0.10] },
  { t: 9.0, type: 'log-shift', strength: 1.0, pos: [0.05, 0.12, 0.0], log: 3 },
  { t: 11.0, type: 'gust', strength: 1.0, pos: [0, 0.3, 0], dur: 1.6 },
  { t: 13.433, type: 'pop', strength: 0.5, pos: [0.08, 0.13, -0.05] },
  { t: 15.2, type: 'pop', strength: 0.85, pos: [-0.06, 0.15, 0.08] },
  { t: 16.633, type: 'pop', strength: 0.5, pos: [0.02, 0.12, 0.12] },
  { t: 18.5, type: 'pop', strength: 1.0, pos: [0.06, 0.14, 0.12], aim: [0.5, 0, 0.86] },
];

// Background resin pops (src/timeline.js:57-77): exponential gaps of mean 1/0.8 s from the 32-bit
// LCG of src/scene.js:13-16, strength 0.12 + 0.45 u^2.2 capped at 0.6, position on the 0.2 m bed,
// a half-strength aftershock one time in four; times rounded to ms, strengths to 0.01; pops within
// 0.35 s of a scripted event are dropped.
function randomPops(seed = 20261025, scripted = SCRIPTED) {
  let state = seed >>> 0;
  const u = () => { state = (state * 1664525 + 1013904223) >>> 0; return state / 4294967296; };
  const pops = [];
  for (let t = 0.6; ;) {
    t += -Math.log(1 - u()) / 0.8;
    if (t >= DURATION - 0.3) break;
    const strength = Math.min(0.6, 0.12 + 0.45 * Math.pow(u(), 2.2));
    const ang = u() * Math.PI * 2;
    const rad = Math.sqrt(u()) * 0.2;
    pops.push({ t: +t.toFixed(3), type: 'pop', strength: +strength.toFixed(2),
      pos: [Math.cos(ang) * rad, 0.07 + u() * 0.15, Math.sin(ang) * rad] });
    if (u() < 0.25) {
      const t2 = t + 0.05 + u() * 0.25;
      pops.push({ t: +t2.toFixed(3), type: 'pop', strength: +(strength * 0.5).toFixed(2),
        pos: [Math.cos(ang) * rad, 0.1, Math.sin(ang) * rad] });
    }
  }
  return pops.filter(p => scripted.every(s => Math.abs(s.t - p.t) >= 0.35));
}

const EVENTS = [...SCRIPTED, ...randomPops()].sort((a, b) => a.t - b.t);

const orDefault = (v, d) => (v === undefined || v === null ? d : v);   // the reference's `??`

function intensityAt(t) {
  const k = INTENSITY;
  if (t <= k[0][0]) return k[0][1];
  for (let i = 1; i < k.length; i++) {
    if (t <= k[i][0]) {
      const x = (t - k[i - 1][0]) / (k[i][0] - k[i - 1][0]);
      return k[i - 1][1] + (k[i][1] - k[i - 1][1]) * (x * x * (3 - 2 * x));
    }
  }
  return k[k.length - 1][1];
}

// The log-shift: one teepee log's foot slides out and its top drops, eased over 0.12 s.
function logOffsetsAt(t, events = EVENTS) {
  const e = events.find(x => x.type === 'log-shift');
  if (!e || t < e.t) return null;
  const x = Math.min((t - e.t) / 0.12, 1);
  return { index: orDefault(e.log, 3), s: x * x * (3 - 2 * x), footOut: 0.02, topDrop: 0.04, topOut: 0.012 };
}

// Parameter overrides for the step at time t: intensity, the strongest live burst, a wandering
// draft and the scripted gust.
function simAt(t, events = EVENTS) {
  const p = { intensity: intensityAt(t), burstAmt: 0, burstX: 0, burstY: 0, burstZ: 0, windX: 0, windZ: 0 };
  for (const e of events) {
    const age = t - e.t;
    if (age < 0 || age > 0.4) continue;
    const shift = e.type === 'log-shift';
    if (shift || e.type === 'pop' || e.type === 'crackle-burst') {
      const w = (shift ? 1.0 : 0.6) * orDefault(e.strength, 1) * Math.exp(-age / (shift ? 0.12 : 0.04));
      if (w > p.burstAmt) { p.burstAmt = w; [p.burstX, p.burstY, p.burstZ] = e.pos; }
    }
  }
  p.windX = 0.28 * Math.sin(t * 0.23 + 1.0) + 0.14 * Math.sin(t * 0.61 + 0.4);
  p.windZ = 0.18 * Math.sin(t * 0.17 + 2.3);
  const g = events.find(e => e.type === 'gust');
  if (g) {
    const age = t - g.t, dur = orDefault(g.dur, 1.6);
    const w = age > 0 && age < dur + 1.2 ? Math.sin(Math.PI * Math.min(age / (dur + 0.6), 1)) ** 1.5 * g.strength : 0;
    p.windX += 2.8 * w;
  }
  return p;
}

// The packed log buffer with a log-shift applied (8 floats per log: a.xyz, r, b.xyz, burn);
// element updates are computed in double and stored to float32 like a Float32Array.
function applyLogOffsets(baseLogs, o) {
  const arr = Float32Array.from(baseLogs);
  if (!o) return arr;
  const i = o.index * 8;
  const ox = arr[i], oz = arr[i + 2];
  const l = Math.hypot(ox, oz) || 1;
  const dx = ox / l, dz = oz / l;
  arr[i] += dx * o.footOut * o.s; arr[i + 2] += dz * o.footOut * o.s; arr[i + 1] -= 0.004 * o.s;
  arr[i + 4] += dx * o.topOut * o.s; arr[i + 6] += dz * o.topOut * o.s; arr[i + 5] -= o.topDrop * o.s;
  return arr;
}

// Spark bursts of the events (src/main.js:79-87); null for events without sparks (gusts).
function sparkBurstFor(e) {
  const s = orDefault(e.strength, 1);
  const pos = e.pos || [0, 0.12, 0];
  switch (e.type) {
    case 'ignite': return { n: 140, pos, speed: 3.2, radius: 0.12, dur: 0.5 };
    case 'pop':
      if (s >= 0.8) return { n: Math.round(110 * s), pos, speed: 3.5 + 2.5 * s, radius: 0.03, dur: 0.05, fan: 0.9, size: 1.4, aim: e.aim || null };
      return { n: Math.round(3 + 12 * s), pos, speed: 3 + 3 * s, radius: 0.02, dur: 0.04, fan: 0.8, size: 1.4 };
    case 'log-shift': return { n: Math.round(190 * s), pos, speed: 4.5, radius: 0.12, dur: 0.15, fan: 0.7, size: 1.5 };
    case 'crackle-burst': return { n: Math.round(12 * s), pos, speed: 3, radius: 0.08, dur: 0.3, fan: 0.6 };
    default: return null;
  }
}

// The capture loop: step j is at t = j / (fps n) with n steps per frame, so event windows
// [t, t + dt) of adjacent steps tile the timeline exactly.
class Capture {
  constructor(sim, sparks, baseLogs, { simDt = 1 / 240, fps = FPS, events = EVENTS } = {}) {
    this.sim = sim;
    this.sparks = sparks;
    this.baseLogs = Float32Array.from(baseLogs);
    this.fps = fps;
    this.events = events;
    this.n = Math.max(1, Math.round(1 / (fps * simDt)));
    this.simDt = 1 / (fps * this.n);
    this.logKey = null;
    this.logUploads = 0;
  }
  stepT(j) { return j / (this.fps * this.n); }
  applyLogs(o) {
    const key = o ? `${o.index}:${o.s.toFixed(3)}` : '';
    if (key === this.logKey) return;
    this.logKey = key;
    this.sim.setLogs(applyLogOffsets(this.baseLogs, o));   // the emitter is rebuilt by the next step
    this.logUploads++;
  }
  step(j) {
    const t = this.stepT(j), tNext = this.stepT(j + 1);
    const p = simAt(t, this.events);
    this.sim.setParams(p);
    this.applyLogs(logOffsetsAt(t, this.events));
    for (const e of this.events) {
      if (e.t >= t && e.t < tNext) {
        const b = sparkBurstFor(e);
        if (b && this.sparks) this.sparks.burst(b);
      }
    }
    this.sim.step(this.simDt);
    if (this.sparks) {
      this.sparks.rate = 1.0 * orDefault(p.intensity, 1);   // a calm fire only drizzles sparks
      this.sparks.step(this.simDt);
    }
  }
  begin({ preroll = PREROLL } = {}) {
    this.sim.reset();
    if (this.sparks) this.sparks.reset();
    this.logKey = null;
    this.applyLogs(null);
    const P = Math.round(preroll * this.fps) * this.n;
    for (let j = -P; j < 0; j++) this.step(j);
  }
  frame(i) { for (let k = 0; k < this.n; k++) this.step(i * this.n + k); }
}

module.exports = { FPS, DURATION, PREROLL, INTENSITY, SCRIPTED, EVENTS, randomPops, intensityAt, logOffsetsAt, simAt,
  applyLogOffsets, sparkBurstFor, Capture };
