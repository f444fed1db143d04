"""The per-step driver inputs (cue sheet, log shift, event bursts, capture loop): Node
(procedural-fire_amd/node/driver.js) and Python (firesim.driver) against each other and
against what the cue sheet fixes.

The reference's timeline module (src/timeline.js) uses syntax this image's Node (v12) does
not parse (`??`), so it is not run: the restatements are pinned by closed forms of the cue
sheet (keyframe values, smoothstep midpoints, burst decay constants, gust peak, log-shift
easing and geometry) and by agreeing with each other bit for bit.
"""
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import firesim as F
from firesim import driver as D

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
RUN = os.path.join(ROOT, "tests", "node", "run_capture.js")
needs_node = pytest.mark.skipif(NODE is None, reason="node is missing")


def test_intensity_keyframes_and_midpoints():
    for t, v in D.INTENSITY:
        assert D.intensity_at(t) == v
    for (t0, v0), (t1, v1) in zip(D.INTENSITY, D.INTENSITY[1:]):
        assert D.intensity_at(0.5 * (t0 + t1)) == pytest.approx(0.5 * (v0 + v1), abs=1e-12)
        q = t0 + 0.25 * (t1 - t0)      # smoothstep(0.25) = 0.15625
        assert D.intensity_at(q) == pytest.approx(v0 + (v1 - v0) * 0.15625, abs=1e-12)
    assert D.intensity_at(-10.0) == 1.0 and D.intensity_at(30.0) == 1.0


def test_bursts_draft_gust_and_log_shift():
    hero = D.SCRIPTED[0]
    p = D.sim_at(hero["t"])
    assert p["burstAmt"] == pytest.approx(0.6) and (p["burstX"], p["burstY"], p["burstZ"]) == hero["pos"]
    assert D.sim_at(hero["t"] + 0.04)["burstAmt"] == pytest.approx(0.6 * math.exp(-1.0))
    assert D.sim_at(hero["t"] + 0.41)["burstAmt"] == 0.0 or D.sim_at(hero["t"] + 0.41)["burstX"] != hero["pos"][0]
    shift = next(e for e in D.EVENTS if e["type"] == "log-shift")
    assert D.sim_at(shift["t"] + 0.12)["burstAmt"] == pytest.approx(math.exp(-1.0))
    assert D.log_offsets_at(shift["t"] - 1e-9) is None
    assert D.log_offsets_at(shift["t"] + 0.06)["s"] == pytest.approx(0.5)
    assert D.log_offsets_at(shift["t"] + 1.0)["s"] == 1.0
    gust = next(e for e in D.EVENTS if e["type"] == "gust")
    peak = gust["t"] + 0.5 * (gust["dur"] + 0.6)
    base = D.sim_at(gust["t"] - 1.0, [e for e in D.EVENTS if e["type"] != "gust"])
    with_g = D.sim_at(peak)
    without = D.sim_at(peak, [e for e in D.EVENTS if e["type"] != "gust"])
    assert with_g["windX"] - without["windX"] == pytest.approx(2.8)
    assert base["windZ"] == pytest.approx(0.18 * math.sin((gust["t"] - 1.0) * 0.17 + 2.3))


def test_event_list_and_pops():
    scripted = [e for e in D.EVENTS if e in D.SCRIPTED]
    assert len(scripted) == len(D.SCRIPTED)
    pops = [e for e in D.EVENTS if e not in D.SCRIPTED]
    assert all(e["type"] == "pop" and 0.6 < e["t"] < 19.7 + 0.3 for e in pops)
    assert all(0.0 <= e["strength"] <= 0.6 for e in pops)
    assert all(math.hypot(e["pos"][0], e["pos"][2]) <= 0.2 + 1e-12 for e in pops)
    assert all(min(abs(e["t"] - s["t"]) for s in D.SCRIPTED) >= 0.35 for e in pops)
    assert [e["t"] for e in D.EVENTS] == sorted(e["t"] for e in D.EVENTS)
    assert 6 <= len(pops) <= 20


def test_log_shift_geometry():
    logs, n = F.build_campfire(7)
    o = D.log_offsets_at(9.0 + 1.0)
    moved = D.apply_log_offsets(logs, o)
    i = o["index"] * 8
    assert np.array_equal(np.delete(moved, range(i, i + 8)), np.delete(logs, range(i, i + 8)))
    foot0, foot1 = np.hypot(logs[i], logs[i + 2]), np.hypot(moved[i], moved[i + 2])
    assert foot1 - foot0 == pytest.approx(0.02, abs=1e-6)
    assert logs[i + 5] - moved[i + 5] == pytest.approx(0.04, abs=1e-6)
    assert logs[i + 1] - moved[i + 1] == pytest.approx(0.004, abs=1e-6)


def test_capture_log_uploads_follow_the_three_decimal_key():
    class Sim:
        uploads = 0

        def set_logs(self, a):
            Sim.uploads += 1
    cap = D.Capture(Sim(), None, np.zeros(8 * 16, np.float32))
    keys = set()
    for j in range(int(8.9 * 240), int(9.3 * 240)):
        o = D.log_offsets_at(cap.step_t(j))
        cap.apply_logs(o)
        keys.add(f"{o['index']}:{D._fixed(o['s'], 3)}" if o else "")
    assert Sim.uploads == len(keys) and 25 <= Sim.uploads <= 32     # ~29 eased steps plus the start


@needs_node
def test_node_and_python_cue_sheets_agree():
    r = subprocess.run([NODE, RUN, "-"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    py = [dict(t=e["t"], type=e["type"], strength=e["strength"], pos=list(e["pos"])) for e in D.EVENTS]
    js = [dict(t=e["t"], type=e["type"], strength=e["strength"], pos=list(e["pos"])) for e in d["events"]]
    assert js == py
    rows = np.array(d["rows"])
    for row in rows:
        t = row[0]
        p = D.sim_at(t)
        o = D.log_offsets_at(t)
        mine = [p["intensity"], p["burstAmt"], p["burstX"], p["burstY"], p["burstZ"], p["windX"], p["windZ"],
                o["s"] if o else -1]
        assert np.allclose(row[1:], mine, rtol=1e-14, atol=1e-15), (t, row[1:], mine)   # libm sin/exp: last bits
    logs, _ = F.build_campfire(7)
    assert np.array_equal(np.array(d["logs"], np.float32), D.apply_log_offsets(logs, D.log_offsets_at(9.06)))


@pytest.mark.gpu
@needs_node
def test_node_capture_loop_matches_python_bit_for_bit(tmp_path):
    """The same compressed cue sheet (pop, log shift, gust, crackle) through both capture loops:
    identical fields, geometry and sparks."""
    print("capture loop node vs python 32x48x32", flush=True)
    r = subprocess.run([NODE, RUN, str(tmp_path), "6"], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    meta = json.loads((tmp_path / "meta.json").read_text())
    events = sorted([dict(t=0.05, type="pop", strength=1.0, pos=(0.06, 0.14, 0.12), aim=(0.5, 0, 0.86)),
                     dict(t=0.1, type="log-shift", strength=1.0, pos=(0.05, 0.12, 0.0), log=3),
                     dict(t=0.0, type="gust", strength=1.0, pos=(0, 0.3, 0), dur=0.2),
                     dict(t=0.15, type="crackle-burst", strength=0.5, pos=(-0.10, 0.08, 0.06)),
                     dict(t=0.2, type="pop", strength=0.4, pos=(0.02, 0.1, -0.05))], key=lambda e: e["t"])
    logs, _ = F.build_campfire(7)
    sim = F.FireSim(res=(32, 48, 32), size=(1.2, 1.8, 1.2), logs=logs, noise_n=32, use_graph=True)
    sparks = F.Sparks(sim, 1024)
    cap = D.Capture(sim, sparks, logs, sim_dt=1 / 240, events=events)
    cap.begin(preroll=0.1)
    for i in range(6):
        cap.frame(i)
    assert meta["n"] == cap.n == 8 and meta["logUploads"] == cap.log_uploads and meta["frame"] == sim.frame
    for f, fid in (("vel", F.VEL), ("S", F.S), ("S2", F.S2), ("geo", F.GEO)):
        a = np.fromfile(tmp_path / f"{f}.bin", np.uint16)
        assert np.array_equal(a, sim.read(fid, "native").ravel()), f
    a = np.fromfile(tmp_path / "sparks.bin", np.float32).reshape(1024, 16)
    b = sparks.read()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert (b[:, 14] > 0.5).sum() > 20
