"""Pins of the spark oracle (src/sparks.js) on the CPU.

The reference has no spark tests; the restatement is pinned by what the update's physics
fixes in closed form and by distribution checks that a plausible mistake breaks:
  * still air (zero velocity and curl textures): the horizontal velocity relaxes as
    v0 exp(-n dt / tau) with tau = 0.4 d^2 s (Stokes response, src/sparks.js:115);
  * gravity: v_y settles at the terminal value -g' dt c / k, g' = 9.81 (1 - 0.004),
    c = 0.2 for d < 0.3 mm, else 1, k = 1 - exp(-dt / tau) (src/sparks.js:121);
  * a burning spark in still air relaxes to 1180 + 60 (seed - 0.5) K with a 0.3 s time
    constant; after burn-out it cools to 293 K with 0.06 + 0.08 d s (src/sparks.js:125-131);
  * kill rules: T < 760 K, below the floor, above 4.5 m, beyond 3 m, older than 12 s;
  * spawn: base sparks on the 0.2 m bed disc between 5 and 17 cm, sizes in [0.15, 2.5] mm with
    median 0.5 mm (log-normal, sigma 0.55); burst sparks inside the burst box; the Box-Muller
    normals have mean 0 and variance 1; one burst emits per step, round(n min(t/dur, 1)) in all;
  * host logic: base spawns follow the carried fraction of rate dt per step (within 1 of rate T).
"""
import math

import numpy as np
import pytest

import fsoracle as O


def _still_sim():
    orc = O.OracleSim(res=(16, 24, 16), size=(0.6, 0.9, 0.6), precision=O.FP32, noise_n=16)
    (nx, ny, nz), ch = orc.dims(O.VEL)
    orc.write(O.VEL, np.zeros((nz, ny, nx, 4)))
    n = 16
    orc.write(O.CURL, np.zeros((n, n, n, 4)))
    return orc


def _one_spark(d_mm, v, T=1200.0, tb=10.0, seed=0.5, pos=(0.0, 0.3, 0.0)):
    a = np.zeros((4, 16), np.float32)      # one alive spark; the other rows are dead slots
    a[0, 0:3] = pos
    a[0, 4:7] = v
    a[0, 7] = tb
    a[0, 8:11] = pos
    a[0, 11] = T
    a[0, 12], a[0, 13], a[0, 14] = d_mm, seed, 1.0
    return a


def test_stokes_relaxation_and_terminal_velocity():
    orc = _still_sim()
    sp = O.OracleSparks(orc, 4)
    sp.set(rate=0.0)
    d, dt, n = 1.0, 1.0 / 240.0, 60
    sp.write(_one_spark(d, (1.0, 0.0, -0.5)))
    for _ in range(n):
        sp.step(dt)
    a = sp.read()[0]
    tau = 0.4 * d * d
    assert a[4] == pytest.approx(1.0 * math.exp(-n * dt / tau), rel=2e-5)
    assert a[6] == pytest.approx(-0.5 * math.exp(-n * dt / tau), rel=2e-5)
    # v_y: geometric series towards -g' dt / k
    k = 1.0 - math.exp(-dt / tau)
    g = 9.81 * 0.996
    vy = sum(-g * dt * (1 - k) ** j for j in range(n))
    assert a[5] == pytest.approx(vy, rel=2e-5)
    # light sparks (d < 0.3 mm) feel a fifth of it
    sp.write(_one_spark(0.25, (0.0, 0.0, 0.0)))
    sp.step(dt)
    tau = max(0.4 * 0.25 * 0.25, 0.01)
    k = 1.0 - math.exp(-dt / tau)
    assert sp.read()[0][5] == pytest.approx(-g * dt * 0.2, rel=1e-5)


def test_burning_and_cooling_temperatures():
    orc = _still_sim()
    sp = O.OracleSparks(orc, 4)
    sp.set(rate=0.0)
    dt = 1.0 / 120.0
    sp.write(_one_spark(1.0, (0.0, 0.0, 0.0), T=1300.0, tb=5.0, seed=0.75))
    n = 24
    for _ in range(n):
        sp.step(dt)
    a = sp.read()[0]
    # still air: rel = |u - v| is the (small) fall speed; the target is 1180 + 60 rel + 60 (0.75 - 0.5)
    vy = abs(a[5])
    target = 1180.0 + 60.0 * 0.25
    assert a[11] == pytest.approx(target + (1300.0 - target) * math.exp(-n * dt / 0.3), abs=60.0 * vy + 0.05)
    # burned out: exponential cooling towards 293 K
    sp.write(_one_spark(1.0, (0.0, 0.0, 0.0), T=1200.0, tb=0.0))
    sp.step(dt)
    a = sp.read()[0]
    assert a[11] == pytest.approx(293.0 + 907.0 * math.exp(-dt / (0.06 + 0.08)), rel=1e-6)
    assert a[14] == 1.0


@pytest.mark.parametrize("case", ["cold", "floor", "high", "far", "old"])
def test_kill_rules(case):
    orc = _still_sim()
    sp = O.OracleSparks(orc, 4)
    sp.set(rate=0.0)
    a = _one_spark(1.0, (0.0, 0.0, 0.0))
    if case == "cold":
        a[0, 11] = 700.0
        a[0, 7] = 0.0
    elif case == "floor":
        a[0, 1] = 0.001
        a[0, 5] = -5.0
    elif case == "high":
        a[0, 1] = 4.6
    elif case == "far":
        a[0, 0] = 3.1
    else:
        a[0, 3] = 12.5
    sp.write(a)
    sp.step(1.0 / 120.0)
    assert sp.read()[0][14] == 0.0


def test_base_spawn_layout_sizes_and_count():
    orc = _still_sim()
    sp = O.OracleSparks(orc, 20000)
    sp.set(rate=40000.0)              # 1/120 s -> 333 base sparks per step
    dt = 1.0 / 120.0
    for _ in range(6):
        sp.step(dt)
    a = sp.read()
    born = a[:, 14] > 0.5
    carry, total = 0.0, 0          # the host's carry (src/sparks.js:242-243), in double like JS
    for _ in range(6):
        carry += 40000.0 * dt
        k = math.floor(carry)
        carry -= k
        total += k
    assert born.sum() == total and abs(total - 40000.0 * 6 * dt) <= 1
    # first-step sparks have moved one step; check the spawn rows of the last step only via the q (prev) of age 0
    fresh = born & (a[:, 3] == 0.0)
    r = np.hypot(a[fresh, 0], a[fresh, 2])
    assert r.max() <= 0.2 + 1e-6
    assert a[fresh, 1].min() >= 0.05 - 1e-6 and a[fresh, 1].max() <= 0.17 + 1e-6
    d = a[born, 12]
    assert d.min() >= 0.15 and d.max() <= 2.5
    assert np.median(d) == pytest.approx(0.5, rel=0.05)
    # log-normal width: sigma of log(d) over the unclamped ones is 0.55
    ln = np.log(d[(d > 0.16) & (d < 2.4)] / 0.5)
    assert np.std(ln) == pytest.approx(0.55, rel=0.08)
    # initial temperatures T0 = 1230 + 70 N(0, 1) (base sparks): mean 1230, sd 70
    T0 = a[fresh, 11]
    assert np.mean(T0) == pytest.approx(1230.0, abs=15.0)
    assert np.std(T0) == pytest.approx(70.0, rel=0.15)


def test_burst_schedule_and_box():
    orc = _still_sim()
    sp = O.OracleSparks(orc, 4096)
    sp.set(rate=0.0)
    dt = 1.0 / 120.0
    sp.burst(n=140, pos=(0.1, 0.2, -0.1), speed=3.2, radius=0.12, dur=0.5)
    sp.burst(n=30, pos=(0.0, 0.1, 0.0), speed=3.0, radius=0.02, dur=0.05)
    counts = []
    for s in range(1, 80):
        sp.step(dt)
        a = sp.read()
        counts.append(int(((a[:, 14] > 0.5) & (a[:, 3] == 0.0)).sum()))   # spawned this step: age 0
        if s == 1:
            fresh = a[:, 14] > 0.5
            p = a[fresh, 8:11]       # previous position = spawn position
            assert np.all(np.abs(p - np.array([0.1, 0.2, -0.1], np.float32)) <= 0.12 + 1e-6)
    # the first burst emits round(140 t / 0.5) as t advances, the second waits until the first is done
    expect_first = [min(140, math.floor(140 * min(k * dt / 0.5, 1.0) + 0.5)) for k in range(1, 80)]
    emitted = np.cumsum(counts)
    assert list(emitted[:60]) == expect_first[:60]  # one burst emits per step: the second waits
    assert emitted[59] == 140                      # 0.5 s: the first burst is complete
    assert emitted[-1] == 170                      # and then the 30 of the second


def test_rand2_uniform_and_box_muller_moments():
    orc = _still_sim()
    sp = O.OracleSparks(orc, 50000)
    sp.set(rate=0.0)
    sp.burst(n=50000, pos=(0.0, 1.0, 0.0), speed=1.0, radius=0.5, dur=1e-3, fan=0.0)
    sp.step(1.0 / 120.0)
    a = sp.read()
    born = a[:, 14] > 0.5
    assert born.sum() == 50000
    # positions in the burst box are pos + (U - 0.5) 2 radius: uniform on [-0.5, 0.5] about pos
    u = (a[born, 8] - 0.0) / 1.0 + 0.5
    assert np.mean(u) == pytest.approx(0.5, abs=0.01)
    assert np.var(u) == pytest.approx(1.0 / 12.0, rel=0.03)
    # seeds r(19) uniform
    assert np.mean(a[born, 13]) == pytest.approx(0.5, abs=0.01)
    # burst T0 = 1310 + 70 N(0, 1)
    T0 = a[born, 11]
    assert np.mean(T0) == pytest.approx(1310.0, abs=1.5)
    assert np.std(T0) == pytest.approx(70.0, rel=0.02)
