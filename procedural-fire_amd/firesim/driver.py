"""Per-step driver inputs of the offline capture, for Python hosts (the Node counterpart is
procedural-fire_amd/node/driver.js; both restate the same pieces of the reference):

  * the cue sheet's simulation overrides: intensity keyframes, scripted events and the seeded
    background pops, burst / draft / gust parameters (src/timeline.js:33-77, 96-157);
  * the log-shift geometry update, uploaded only when its 3-decimal key changes
    (src/main.js:135-150);
  * the mapping of events to spark bursts (src/main.js:77-87);
  * the deterministic capture step loop: step j at t = j / (fps n), n steps per frame, events
    fired in [t, t + dt) (src/main.js:153-182).
"""
import math
from decimal import ROUND_HALF_UP, Decimal

import numpy as np

FPS = 30
DURATION = 20
PREROLL = 5.0

INTENSITY = [
    (-PREROLL, 1.0), (0.0, 1.0), (1.0, 1.0), (2.0, 1.12), (6.0, 1.12), (8.6, 1.0), (9.0, 1.2), (9.6, 1.05),
    (11.0, 1.08), (12.6, 1.08), (14.0, 1.05), (18.3, 1.05), (18.5, 1.15), (19.2, 1.0), (20.0, 1.0),
]

SCRIPTED = [
    dict(t=0.30, type="pop", strength=1.0, pos=(0.06, 0.14, 0.12), aim=(0.5, 0, 0.86)),
    dict(t=2.433, type="pop", strength=0.5, pos=(-0.08, 0.12, 0.10)),
    dict(t=4.6, type="pop", strength=0.6, pos=(0.10, 0.16, 0.02)),
    dict(t=5.633, type="pop", strength=0.45, pos=(-0.05, 0.10, 0.12)),
    dict(t=7.2, type="crackle-burst", strength=0.5, pos=(-0.10, 0.08, 0.06)),
    dict(t=8.433, type="pop", strength=0.5, pos=(0.06, 0.12, 0.10)),
    dict(t=9.0, type="log-shift", strength=1.0, pos=(0.05, 0.12, 0.0), log=3),
    dict(t=11.0, type="gust", strength=1.0, pos=(0, 0.3, 0), dur=1.6),
    dict(t=13.433, type="pop", strength=0.5, pos=(0.08, 0.13, -0.05)),
    dict(t=15.2, type="pop", strength=0.85, pos=(-0.06, 0.15, 0.08)),
    dict(t=16.633, type="pop", strength=0.5, pos=(0.02, 0.12, 0.12)),
    dict(t=18.5, type="pop", strength=1.0, pos=(0.06, 0.14, 0.12), aim=(0.5, 0, 0.86)),
]


def _fixed(x, digits):
    """JavaScript's x.toFixed(digits) for x >= 0: the exact binary value rounded half up."""
    return str(Decimal(x).quantize(Decimal(1).scaleb(-digits), rounding=ROUND_HALF_UP))


def random_pops(seed=20261025, scripted=SCRIPTED):
    """Background resin pops: exponential gaps of mean 1/0.8 s from the LCG of src/scene.js:13-16,
    strength 0.12 + 0.45 u^2.2 (cap 0.6), position on the 0.2 m bed, a half-strength aftershock one
    time in four; pops within 0.35 s of a scripted event dropped."""
    state = seed & 0xFFFFFFFF

    def u():
        nonlocal state
        state = (state * 1664525 + 1013904223) & 0xFFFFFFFF
        return state / 4294967296

    pops = []
    t = 0.6
    while True:
        t += -math.log(1 - u()) / 0.8
        if t >= DURATION - 0.3:
            break
        strength = min(0.6, 0.12 + 0.45 * math.pow(u(), 2.2))
        ang = u() * math.pi * 2
        rad = math.sqrt(u()) * 0.2
        pops.append(dict(t=float(_fixed(t, 3)), type="pop", strength=float(_fixed(strength, 2)),
                         pos=(math.cos(ang) * rad, 0.07 + u() * 0.15, math.sin(ang) * rad)))
        if u() < 0.25:
            t2 = t + 0.05 + u() * 0.25
            pops.append(dict(t=float(_fixed(t2, 3)), type="pop", strength=float(_fixed(strength * 0.5, 2)),
                             pos=(math.cos(ang) * rad, 0.1, math.sin(ang) * rad)))
    return [p for p in pops if all(abs(s["t"] - p["t"]) >= 0.35 for s in scripted)]


EVENTS = sorted(SCRIPTED + random_pops(), key=lambda e: e["t"])


def intensity_at(t):
    k = INTENSITY
    if t <= k[0][0]:
        return k[0][1]
    for i in range(1, len(k)):
        if t <= k[i][0]:
            x = (t - k[i - 1][0]) / (k[i][0] - k[i - 1][0])
            return k[i - 1][1] + (k[i][1] - k[i - 1][1]) * (x * x * (3 - 2 * x))
    return k[-1][1]


def log_offsets_at(t, events=EVENTS):
    e = next((x for x in events if x["type"] == "log-shift"), None)
    if e is None or t < e["t"]:
        return None
    x = min((t - e["t"]) / 0.12, 1)
    return dict(index=e.get("log", 3), s=x * x * (3 - 2 * x), footOut=0.02, topDrop=0.04, topOut=0.012)


def sim_at(t, events=EVENTS):
    p = dict(intensity=intensity_at(t), burstAmt=0.0, burstX=0.0, burstY=0.0, burstZ=0.0, windX=0.0, windZ=0.0)
    for e in events:
        age = t - e["t"]
        if age < 0 or age > 0.4:
            continue
        shift = e["type"] == "log-shift"
        if shift or e["type"] in ("pop", "crackle-burst"):
            w = (1.0 if shift else 0.6) * e.get("strength", 1) * math.exp(-age / (0.12 if shift else 0.04))
            if w > p["burstAmt"]:
                p["burstAmt"] = w
                p["burstX"], p["burstY"], p["burstZ"] = e["pos"]
    p["windX"] = 0.28 * math.sin(t * 0.23 + 1.0) + 0.14 * math.sin(t * 0.61 + 0.4)
    p["windZ"] = 0.18 * math.sin(t * 0.17 + 2.3)
    g = next((x for x in events if x["type"] == "gust"), None)
    if g is not None:
        age, dur = t - g["t"], g.get("dur", 1.6)
        w = math.sin(math.pi * min(age / (dur + 0.6), 1)) ** 1.5 * g["strength"] if 0 < age < dur + 1.2 else 0
        p["windX"] += 2.8 * w
    return p


def apply_log_offsets(base_logs, o):
    """Packed logs (8 floats per log) with a log-shift applied; each update computed in double and
    stored to float32."""
    arr = np.array(base_logs, np.float32)
    if o is None:
        return arr
    i = o["index"] * 8
    ox, oz = float(arr[i]), float(arr[i + 2])
    l = math.hypot(ox, oz) or 1
    dx, dz = ox / l, oz / l
    s = o["s"]
    arr[i] = np.float32(float(arr[i]) + dx * o["footOut"] * s)
    arr[i + 2] = np.float32(float(arr[i + 2]) + dz * o["footOut"] * s)
    arr[i + 1] = np.float32(float(arr[i + 1]) - 0.004 * s)
    arr[i + 4] = np.float32(float(arr[i + 4]) + dx * o["topOut"] * s)
    arr[i + 6] = np.float32(float(arr[i + 6]) + dz * o["topOut"] * s)
    arr[i + 5] = np.float32(float(arr[i + 5]) - o["topDrop"] * s)
    return arr


def _js_round(x):
    return math.floor(x + 0.5)


def spark_burst_for(e):
    s = e.get("strength", 1)
    pos = e.get("pos") or (0, 0.12, 0)
    ty = e["type"]
    if ty == "ignite":
        return dict(n=140, pos=pos, speed=3.2, radius=0.12, dur=0.5)
    if ty == "pop" and s >= 0.8:
        return dict(n=_js_round(110 * s), pos=pos, speed=3.5 + 2.5 * s, radius=0.03, dur=0.05, fan=0.9, size=1.4,
                    aim=e.get("aim"))
    if ty == "pop":
        return dict(n=_js_round(3 + 12 * s), pos=pos, speed=3 + 3 * s, radius=0.02, dur=0.04, fan=0.8, size=1.4)
    if ty == "log-shift":
        return dict(n=_js_round(190 * s), pos=pos, speed=4.5, radius=0.12, dur=0.15, fan=0.7, size=1.5)
    if ty == "crackle-burst":
        return dict(n=_js_round(12 * s), pos=pos, speed=3, radius=0.08, dur=0.3, fan=0.6)
    return None


class Capture:
    """The capture loop over a FireSim (and optionally Sparks): begin() resets and pre-rolls,
    frame(i) runs the n steps of frame i."""

    def __init__(self, sim, sparks, base_logs, sim_dt=1 / 240, fps=FPS, events=EVENTS):
        self.sim, self.sparks = sim, sparks
        self.base_logs = np.array(base_logs, np.float32)
        self.fps, self.events = fps, events
        self.n = max(1, _js_round(1 / (fps * sim_dt)))
        self.sim_dt = 1 / (fps * self.n)
        self.log_key = None
        self.log_uploads = 0

    def step_t(self, j):
        return j / (self.fps * self.n)

    def apply_logs(self, o):
        key = f"{o['index']}:{_fixed(o['s'], 3)}" if o else ""
        if key == self.log_key:
            return
        self.log_key = key
        self.sim.set_logs(apply_log_offsets(self.base_logs, o))
        self.log_uploads += 1

    def step(self, j):
        t, t_next = self.step_t(j), self.step_t(j + 1)
        p = sim_at(t, self.events)
        self.sim.set_params(p)
        self.apply_logs(log_offsets_at(t, self.events))
        for e in self.events:
            if t <= e["t"] < t_next:
                b = spark_burst_for(e)
                if b and self.sparks is not None:
                    self.sparks.burst(**b)
        self.sim.step(self.sim_dt)
        if self.sparks is not None:
            self.sparks.set(rate=1.0 * p.get("intensity", 1))
            self.sparks.step(self.sim_dt)

    def begin(self, preroll=PREROLL):
        self.sim.reset()
        if self.sparks is not None:
            self.sparks.reset()
        self.log_key = None
        self.apply_logs(None)
        P = _js_round(preroll * self.fps) * self.n
        for j in range(-P, 0):
            self.step(j)

    def frame(self, i):
        for k in range(self.n):
            self.step(i * self.n + k)
