"""The adjoint and tangent with hot-gas viscosity on (src/sim.js:339-346), against FP64 differences.

The viscous term adds kv(T) lap(v) to the velocity, with lap the 7-point Laplacian of the velocity over
clamped texel loads and kv = min(visc 1.5e-5 (max(T, 293)/293)^1.7 dt/dx^2, 0.15).  Its derivative has a
local part (through T and the cell's own velocity) and a stencil part (the neighbours' velocities), which
the adjoint spreads with a transposed Laplacian (k_visc_adj) and the tangent forms from the neighbours'
tangents.

Workload: the adjoint scene of test_gpu_adjoint.py (8 pyrolysis spots, floor fuel flux, 40x48x40 at
1 cm, dt 2 ms, FP32 storage, pressure solve refined to 1e-9, 30-step warm-up) with visc = 2 and
lamAge = 0.2, the viscosity and laminar gate of the reference's production preset
(tools/render_final.mjs:21).  References: central differences of the FP64 oracle from the identical
state (bar: relative error <= 1e-3), and the tangent (dot-product test over 50 steps, bar 1e-4).
Each derivative is also compared with the oracle's differences at visc = 0 from the same state, to show
that the viscous term moves it by more than the bar (so the checks see it).
"""
import numpy as np
import pytest

import firesim as F
import fsoracle as O
from firesim import scenes
from test_gpu_adjoint import DT, NAME, RES, TOL, WARM, _loss, _oracle_load, _random_seeds, _smooth, _weighted

pytestmark = pytest.mark.gpu

VISC = dict(visc=2.0, lamAge=0.2)


def _sim():
    sim = scenes.make_scene(NAME, res=RES, precision=F.PREC_FP32, solve=F.SOLVE_REFINE64, div_tol=TOL, max_vcycles=40,
                            noise_n=32)
    sim.set_face_flux(scenes.scene_face_flux(NAME, RES))
    sim.set_params(**VISC)
    return sim


def _oracle(gpu, visc=True):
    sc = scenes.SCENES[NAME]
    size = tuple(float(np.float32(n * sc["dx"])) for n in RES)
    orc = O.OracleSim(res=RES, size=size, logs=np.zeros(8, np.float32), n_logs=0, precision=O.FP64,
                      vcycles=sc["vcycles"], max_levels=sc["max_levels"], solve=O.SOLVE_REFINE64, div_tol=1e-11,
                      max_vcycles=40, noise_n=32)
    prm = dict(scenes.scene_params(NAME))
    prm.update(VISC if visc else dict(visc=0.0, lamAge=VISC["lamAge"]))
    orc.set_params(**prm)
    for f in (F.GEO, F.NOISE, F.CURL):
        orc.write(f, gpu.read(f, np.float64))
    return orc


@pytest.fixture(scope="module")
def vcase():
    print(f"viscous adjoint case {NAME} {RES} warm {WARM} {VISC}", flush=True)
    sim = _sim()
    sim.step(DT, WARM)
    sim.sync()
    rows = sim.spots()[:, :6].astype(np.float32)
    face = scenes.scene_face_flux(NAME, RES)
    base = sim.save_state()
    rng = np.random.default_rng(2025)
    n = 6
    pr = rows.copy()
    pr[:, 4] *= (1.0 + rng.normal(0.0, 0.2, len(rows))).astype(np.float32)
    sim.set_spots(pr)
    sim.step(DT, n)
    S = sim.read("S")
    target = (S[..., 0].astype(np.float64), S[..., 2].astype(np.float64))
    sim.set_spots(rows)
    sim.load_state(base)
    return dict(sim=sim, rows=rows, face=face, base=base, target=target, n=n)


def test_viscous_adjoint_is_accepted_and_the_preset_is_on(vcase):
    print("viscous adjoint: parameters", flush=True)
    p = vcase["sim"].params
    assert p["visc"] == np.float32(2.0) and p["lamAge"] == np.float32(0.2)


def test_viscous_spot_flux_gradient_against_fp64_differences(vcase):
    print("viscous adjoint: spot-flux gradient vs FP64 FD, 6 steps", flush=True)
    sim, rows, target, n = vcase["sim"], vcase["rows"], vcase["target"], vcase["n"]
    sim.load_state(vcase["base"])
    adj = F.Adjoint(sim, n, 4)
    adj.forward(DT, n)
    adj.seed_mismatch(*[np.ascontiguousarray(a, np.float32) for a in target])
    adj.backward(solve_tol=TOL, max_cycles=40)
    g_adj, _ = adj.grads(len(rows), face=True)

    def fd(visc):
        orc = _oracle(sim, visc)
        orc.set_face_flux(vcase["face"])
        orc.set_spots(rows.astype(np.float64))
        g = np.zeros(len(rows))
        for q in range(len(rows)):
            f0 = float(rows[q, 4])
            vals = {}
            for sgn in (1, -1):
                r = rows[:, 4].astype(np.float64)
                r[q] = f0 * (1 + sgn * 1e-6)
                orc.set_flux64(spot=r)
                _oracle_load(orc, vcase["base"])
                orc.step(DT, n)
                vals[sgn] = _loss(orc.read(O.S), target)
            g[q] = (vals[1] - vals[-1]) / (2e-6 * f0)
        return g

    g64 = fd(True)
    g64_0 = fd(False)
    e64 = np.linalg.norm(g_adj - g64) / np.linalg.norm(g64)
    moved = np.linalg.norm(g64_0 - g64) / np.linalg.norm(g64)
    print("adjoint", g_adj, "\nfp64 FD", g64, "\nfp64 FD at visc 0", g64_0,
          f"\nrel err {e64:.3e}; the viscous term moves the FP64 gradient by {moved:.3e}", flush=True)
    assert e64 <= 1e-3
    assert moved > 3e-3


def test_viscous_initial_velocity_gradient_one_step(vcase):
    """The stencil part of the viscous term's derivative, over one step: L = w . vel with w random on the
    hot interior cells (theta > 0.5, at least 4 cells from every face) and zero elsewhere, along a velocity
    direction on the same cells that alternates every two cells (cos(pi/2 (x + y + z + c))), which the
    7-point Laplacian amplifies about sixfold; FP64 differences with and without viscosity."""
    print("viscous adjoint: initial-velocity gradient on the hot cells, 1 step, vs FP64 FD", flush=True)
    sim = vcase["sim"]
    vel, S, S2, p, t, fr = vcase["base"]
    nz, ny, nx = vel.shape[:3]
    hot = np.zeros((nz, ny, nx), bool)
    hot[4:-4, 4:-4, 4:-4] = True
    hot &= S[..., 0] > 0.5
    print(f"hot interior cells: {int(hot.sum())}", flush=True)
    assert hot.sum() >= 50
    rng = np.random.default_rng(19)
    w = {"vel": np.zeros(vel.shape, np.float32), "S": np.zeros(S.shape, np.float32), "S2": np.zeros(S2.shape, np.float32)}
    w["vel"][hot, :3] = rng.normal(0.0, 1.0, (int(hot.sum()), 3))
    sim.load_state(vcase["base"])
    adj = F.Adjoint(sim, 1, 1)
    adj.forward(DT, 1)
    for f in ("vel", "S", "S2"):
        adj.seed(f, w[f])
    adj.backward(solve_tol=TOL, max_cycles=40)
    g = adj.read("vel").astype(np.float64)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    d = np.zeros(vel.shape, np.float64)
    for c in range(3):
        d[..., c] = np.where(hot, 1e-2 * np.cos(0.5 * np.pi * (x + y + z + c)), 0.0)
    pred = float(np.sum(g * d))
    res = {}
    for visc in (True, False):
        orc = _oracle(sim, visc)
        orc.set_spots(vcase["rows"].astype(np.float64))
        orc.set_face_flux(vcase["face"])
        h = 1e-4
        vals = {}
        for sgn in (1, -1):
            _oracle_load(orc, (vel.astype(np.float64) + sgn * h * d, S.astype(np.float64), S2.astype(np.float64), p, t, fr))
            orc.step(DT, 1)
            vals[sgn] = _weighted(orc, w)
        res[visc] = (vals[1] - vals[-1]) / (2 * h)
    err = abs(pred - res[True]) / abs(res[True])
    moved = abs(res[False] - res[True]) / abs(res[True])
    print(f"adjoint {pred:.6e} FP64 FD {res[True]:.6e} rel err {err:.3e}; at visc 0 {res[False]:.6e} "
          f"(moved {moved:.3e})", flush=True)
    assert err <= 1e-3
    assert moved > 3e-3


def test_viscous_adjoint_is_the_transpose_of_the_tangent_over_50_steps(vcase):
    print("viscous adjoint/tangent dot product, 50 steps", flush=True)
    sim, rows, face = vcase["sim"], vcase["rows"], vcase["face"]
    n = 50
    rng = np.random.default_rng(21)
    w = _random_seeds(23)
    vel, S, S2 = vcase["base"][:3]
    d_spot = rng.normal(0.0, 1.0, len(rows)) * 1e-3
    d_face = (np.where(face > 0, rng.normal(0.0, 1.0, face.shape), 0.0) * 1e-3).astype(np.float32)
    d_vel = np.zeros_like(vel)
    d_vel[4:-4, 4:-4, 4:-4, :3] = 1e-2 * rng.normal(0.0, 1.0, d_vel[4:-4, 4:-4, 4:-4, :3].shape)
    d_S = np.where(S > 1e-3, 1e-3 * rng.normal(0.0, 1.0, S.shape), 0.0).astype(np.float32)
    d_S2 = np.where(S2 > 1e-3, 1e-3 * rng.normal(0.0, 1.0, S2.shape), 0.0).astype(np.float32)
    sim.load_state(vcase["base"])
    adj = F.Adjoint(sim, n, 10)
    t = adj.tangent(DT, n, spot_dir=d_spot, face_dir=d_face, state_dir=(d_vel.astype(np.float32), d_S, d_S2),
                    solve_tol=TOL, max_cycles=40)
    lhs = sum(float(np.sum(t[f].astype(np.float64) * w[f])) for f in ("vel", "S", "S2"))
    for f in ("vel", "S", "S2"):
        adj.seed(f, w[f])
    adj.backward(solve_tol=TOL, max_cycles=40)
    gs, gf = adj.grads(len(rows), face=True)
    g = {f: adj.read(f).astype(np.float64) for f in ("vel", "S", "S2")}
    parts = [float(np.sum(gs * d_spot)), float(np.sum(gf * d_face)), float(np.sum(g["vel"] * d_vel)),
             float(np.sum(g["S"] * d_S)), float(np.sum(g["S2"] * d_S2))]
    rhs = sum(parts)
    err = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"<w, J d> {lhs:.9e}  <J^T w, d> {rhs:.9e}  rel {err:.2e}  parts {parts}", flush=True)
    assert err <= 1e-4
