"""Device diagnostics against host reductions of the same fields: FIRE.stats and FIRE.profile
(src/main.js:268-355) restated in numpy, and the per-step record against the state read back
after each step."""
import numpy as np
import pytest

import firesim as F
from firesim import scenes

pytestmark = pytest.mark.gpu


def _stats_np(a):
    """FIRE.stats on a (nz, ny, nx, 4) array: min/max/mean over finite values, bad count, first argmax."""
    nz, ny, nx, _ = a.shape
    out = []
    for c in range(4):
        v = a[..., c].astype(np.float64).ravel()
        fin = np.isfinite(v)
        vf = np.where(fin, v, -np.inf)
        arg = int(np.argmax(vf)) if fin.any() else -1
        out.append({"min": v[fin].min() if fin.any() else np.inf, "max": v[fin].max() if fin.any() else -np.inf,
                    "sum": v[fin].sum(), "bad": int((~fin).sum()),
                    "argmax": None if arg < 0 else (arg % nx, (arg // nx) % ny, arg // (nx * ny))})
    return out


def _profile_np(S, dx):
    """FIRE.profile: |values| as the reference's sign-less decode, every 4th y-plane."""
    nz, ny, nx, _ = S.shape
    a = np.abs(S.astype(np.float64))
    T = 293 + 1000 * a[..., 0]
    q = a[..., 3] * 1e4
    Q = np.sum(np.where(q > 0, 354800 * q / np.where(q > 0, T, 1.0), 0.0)) * dx ** 3
    k, i = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    centre = (np.abs(i - nx / 2) < 3) & (np.abs(k - nz / 2) < 3)
    rows = []
    for j in range(0, ny, 4):
        Tj = T[:, j, :]
        rows.append({"h": (j + 0.5) * dx, "Tc": Tj[centre].mean(), "Tmax": Tj.max(), "soot": a[:, j, :, 2].max(),
                     "fuel": a[:, j, :, 1].max(), "hotCells": int((Tj > 1000).sum())})
    return Q, rows


@pytest.mark.parametrize("prec", [F.PREC_FP16, F.PREC_FP32], ids=["fp16", "fp32"])
def test_stats_and_profile_match_host_reductions(prec):
    print(f"diag stats/profile spots4_64 prec {prec}", flush=True)
    sim = scenes.make_scene("spots4_64", precision=prec, noise_n=32)
    sim.step(0.002, 40)
    sim.sync()
    for name, fid in (("S", F.S), ("vel", F.VEL), ("S2", F.S2)):
        a = sim.read(fid)
        got = sim.stats(name)
        want = _stats_np(a)
        n = a.shape[0] * a.shape[1] * a.shape[2]
        for c in range(4):
            assert got[c]["min"] == want[c]["min"] and got[c]["max"] == want[c]["max"], (name, c)
            assert got[c]["argmax"] == want[c]["argmax"], (name, c)
            assert got[c]["bad"] == want[c]["bad"]
            assert abs(got[c]["mean"] - want[c]["sum"] / n) <= 1e-12 * max(abs(want[c]["sum"] / n), 1e-30)
    # non-finite values are counted, not folded into min/max/mean
    S = sim.read(F.S)
    S[3, 5, 7, 1] = np.nan
    S[4, 6, 8, 1] = np.inf
    sim.write(F.S, S)
    got = sim.stats("S")
    assert got[1]["bad"] == 2 and np.isfinite(got[1]["max"])
    # profile
    S = sim.read(F.S)
    S[3, 5, 7, 1] = 0.0
    S[4, 6, 8, 1] = 0.0
    sim.write(F.S, S)
    S = sim.read(F.S)
    p = sim.profile()
    Q, rows = _profile_np(S, float(np.float32(sim.dx)))
    assert Q > 0 and abs(p["Q_kW"] * 1e3 - Q) <= 1e-12 * Q
    assert len(p["rows"]) == len(rows)
    for g, w in zip(p["rows"], rows):
        assert g["Tmax"] == w["Tmax"] and g["soot"] == w["soot"] and g["fuel"] == w["fuel"]
        assert g["hotCells"] == w["hotCells"]
        assert abs(g["Tc"] - w["Tc"]) <= 1e-12 * w["Tc"]
        assert abs(g["h"] - w["h"]) <= 1e-9


@pytest.mark.parametrize("use_graph", [False, True], ids=["eager", "graph"])
def test_per_step_record_matches_the_state_after_each_step(use_graph):
    print(f"diag record spots4_64 graph={use_graph}", flush=True)
    sim = scenes.make_scene("spots4_64", precision=F.PREC_FP16, noise_n=32, use_graph=use_graph)
    sim.step(0.002, 30)
    sim.diag_enable(4, t_flame=900.0)          # a ring smaller than the run: keeps the newest
    dx = float(np.float32(sim.dx))
    want = []
    for _ in range(6):
        sim.step(0.002, 1)
        sim.sync()
        S = sim.read(F.S)
        Q, _ = _profile_np(S, dx)
        T = 293 + 1000 * np.abs(S[..., 0].astype(np.float64))
        planeT = T.max(axis=(0, 2))
        hot = np.nonzero(planeT > 900.0)[0]
        nz, ny, nx, _ = S.shape
        k, i = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
        centre = (np.abs(i - nx / 2) < 3) & (np.abs(k - nz / 2) < 3)
        axis = np.array([T[:, j, :][centre].max() for j in range(ny)])
        want.append((sim.time, Q, T.max(), (hot[-1] + 1) * dx if hot.size else 0.0, axis))
    s = sim.diag_series()
    assert s["time"].size == 4
    for r, w in enumerate(want[-4:]):
        assert s["time"][r] == pytest.approx(w[0], abs=1e-12)
        assert s["Q"][r] == pytest.approx(w[1], rel=1e-12)
        assert s["Tmax"][r] == w[2]
        assert s["tip"][r] == pytest.approx(w[3], abs=1e-12)
        assert np.array_equal(s["axisT"][r], w[4])
    assert s["Q"].min() > 0
    sim.diag_enable(0)
    assert sim.diag_series()["time"].size == 0
