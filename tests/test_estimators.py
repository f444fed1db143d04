"""CPU tests of the flame-statistics estimators on series with known answers."""
import numpy as np
import pytest

from firesim import estimators as E


def test_puffing_frequency_recovers_a_known_tone():
    dt = 1e-3
    t = np.arange(12000) * dt          # 12 s: Welch segments of 4 s
    rng = np.random.default_rng(0)
    for f0 in (1.5 / np.sqrt(0.3), 2.1, 3.7):
        q = 55e3 + 8e3 * np.sin(2 * np.pi * f0 * t + 0.3) + 2e3 * rng.standard_normal(t.size) + 3e3 * t
        f = E.puffing_frequency(q, dt)
        assert abs(f - f0) <= 0.03 * f0, (f, f0)


def test_puffing_correlation_matches_reference_examples():
    # physics.md:361: D = 0.3 m -> 2.7 Hz, D = 0.5 m -> 2.1 Hz, D = 1 m -> 1.5 Hz
    assert E.puffing_correlation(0.3) == pytest.approx(2.74, abs=0.01)
    assert E.puffing_correlation(0.5) == pytest.approx(2.12, abs=0.01)
    assert E.puffing_correlation(1.0) == 1.5


def test_flame_height_is_the_half_intermittency_point():
    # flame on the axis up to a height h(t) uniform in [0.2, 0.6] m: I(z) = (0.6 - z) / 0.4 on
    # [0.2, 0.6], so I = 0.5 at z = 0.4 m
    dx = 0.005
    ny = 192
    z = (np.arange(ny) + 0.5) * dx
    rng = np.random.default_rng(1)
    h = rng.uniform(0.2, 0.6, 20000)
    axisT = np.where(z[None, :] < h[:, None], 1300.0, 400.0)
    assert E.flame_height(axisT, dx) == pytest.approx(0.4, abs=2 * dx)
    assert E.tip_height(h) == pytest.approx(0.4, abs=0.01)
    I = E.intermittency(axisT)
    assert I[0] == 1.0 and I[-1] == 0.0


def test_flame_height_edge_cases():
    dx = 0.01
    assert E.flame_height(np.full((10, 50), 300.0), dx) == 0.0          # never burning
    assert E.flame_height(np.full((10, 50), 1500.0), dx) == 50 * dx     # burning to the top


def test_summary_fields():
    dt, dx = 1e-3, 0.005
    n, ny = 3000, 64
    t = np.arange(n) * dt
    s = {"time": t, "Q": 5e4 + 1e4 * np.sin(2 * np.pi * 2.5 * t), "Tmax": np.full(n, 1800.0),
         "tip": np.full(n, 0.3), "axisT": np.where(np.arange(ny)[None, :] < 40, 1200.0, 300.0).repeat(n, 0)}
    out = E.summary(s, dx, dt, skip=100)
    assert out["records"] == n - 100
    assert out["puffing_Hz_Q"] == pytest.approx(2.5, rel=0.03)
    assert out["flame_height_m"] == pytest.approx(40 * dx, abs=dx)
