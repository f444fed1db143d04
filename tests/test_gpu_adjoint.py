"""Gradients of the GPU adjoint (fs_adjoint_*) against finite differences and the tangent.

Workload: the adjoint scene (8 pyrolysis spots of U(1.5, 3) g/s on a 0.3 m disc plus a
floor fuel-flux texture under the disc, 1 cm cells, dt 2 ms, FP32 storage, pressure solve
refined to 1e-9) on a 40x48x40 grid after a 30-step warm-up; the loss is the
temperature/soot mismatch L = sum (theta - theta*)^2 + (soot - soot*)^2 after N steps
against a target made by the same N steps with every flux scaled by (1 + N(0, 0.2)), or a
random weighting L = w . (vel, S, S2).

References:
  * FP64: central differences of the FP64 oracle (float64 storage and arithmetic, solve
    converged to 1e-11) from the identical state, the fluxes perturbed in double precision:
    the FP64 reference derivative.  Bar: relative error <= 1e-3.
  * FD32: central differences of the GPU's own FP32 steps, h = 1e-3 relative.
  * Tangent: the forward-mode derivative along the same FP32 trajectory; over 200 steps the
    adjoint must be its transpose (dot-product test).  Beyond ~50 steps differences stop
    being a reference: the FP64 derivative moves by tens of percent under a 1e-7 change of
    the initial state, while the FP32 and FP64 trajectories drift apart by ~1e-2.
"""
import numpy as np
import pytest

import firesim as F
import fsoracle as O
from firesim import scenes

pytestmark = pytest.mark.gpu

NAME = "adjoint8_64"
RES = (40, 48, 40)
DT = 0.002
WARM = 30
TOL = 1e-9


def _sim(precision=F.PREC_FP32, solve=F.SOLVE_REFINE64):
    sim = scenes.make_scene(NAME, res=RES, precision=precision, solve=solve, div_tol=TOL, max_vcycles=40, noise_n=32)
    sim.set_face_flux(scenes.scene_face_flux(NAME, RES))
    return sim


def _oracle(gpu):
    sc = scenes.SCENES[NAME]
    size = tuple(float(np.float32(n * sc["dx"])) for n in RES)
    orc = O.OracleSim(res=RES, size=size, logs=np.zeros(8, np.float32), n_logs=0, precision=O.FP64,
                      vcycles=sc["vcycles"], max_levels=sc["max_levels"], solve=O.SOLVE_REFINE64, div_tol=1e-11,
                      max_vcycles=40, noise_n=32)
    orc.set_params(**scenes.scene_params(NAME))
    for f in (F.GEO, F.NOISE, F.CURL):
        orc.write(f, gpu.read(f, np.float64))
    return orc


def _oracle_load(orc, st):
    vel, S, S2, p, t, fr = st
    orc.write(O.VEL, vel.astype(np.float64))
    orc.write(O.S, S.astype(np.float64))
    orc.write(O.S2, S2.astype(np.float64))
    orc.write(O.P, p)
    orc.set_clock(t, fr)


def _loss(S, target):
    S = np.asarray(S, np.float64)
    return float(np.sum((S[..., 0] - target[0]) ** 2) + np.sum((S[..., 2] - target[1]) ** 2))


@pytest.fixture(scope="module")
def case():
    """The warmed-up sim, its base state, spot rows, face flux and the mismatch target (N = 6)."""
    print(f"adjoint case {NAME} {RES} warm {WARM}", flush=True)
    sim = _sim()
    sim.step(DT, WARM)
    sim.sync()
    rows = sim.spots()[:, :6].astype(np.float32)
    face = scenes.scene_face_flux(NAME, RES)
    base = sim.save_state()
    rng = np.random.default_rng(2024)
    n = 6
    pr = rows.copy()
    pr[:, 4] *= (1.0 + rng.normal(0.0, 0.2, len(rows))).astype(np.float32)
    pf = (face * (1.0 + rng.normal(0.0, 0.2, face.shape))).astype(np.float32)
    sim.set_spots(pr)
    sim.set_face_flux(pf)
    sim.step(DT, n)
    S = sim.read("S")
    target = (S[..., 0].astype(np.float64), S[..., 2].astype(np.float64))
    sim.set_spots(rows)
    sim.set_face_flux(face)
    sim.load_state(base)
    return dict(sim=sim, rows=rows, face=face, base=base, target=target, n=n)


def _adjoint_grads(c, every=4):
    sim = c["sim"]
    sim.load_state(c["base"])
    adj = F.Adjoint(sim, c["n"], every)
    adj.forward(DT, c["n"])
    t32 = [np.ascontiguousarray(a, np.float32) for a in c["target"]]
    loss = adj.seed_mismatch(*t32)
    adj.backward(solve_tol=TOL, max_cycles=40)
    gs, gf = adj.grads(len(c["rows"]), face=True)
    return adj, loss, gs, gf


def test_mismatch_loss_and_final_state(case):
    print("adjoint loss value and final state", flush=True)
    sim = case["sim"]
    adj, loss, gs, gf = _adjoint_grads(case)
    S_after = sim.read("S")
    t32 = [np.asarray(a, np.float32) for a in case["target"]]
    # the kernel forms theta - theta* in float32 before squaring in double
    assert np.isclose(loss, _loss(S_after, t32), rtol=1e-6, atol=0)
    # the same steps through fs_step land on the same state (the solve's stopping test is
    # checked from the host there, on the device here: agreement to the solve tolerance)
    sim.load_state(case["base"])
    sim.step(DT, case["n"])
    S_step = sim.read("S")
    err = np.abs(S_step - S_after).max()
    assert err < 1e-5, err
    assert np.all(np.isfinite(gs)) and np.all(np.isfinite(gf))
    assert np.abs(gs).max() > 0
    # fuel from the floor has not reached the flame after 6 steps: no theta/soot sensitivity
    assert np.abs(gf).max() == 0


def test_spot_flux_gradient_against_fp64_and_fp32_differences(case):
    print("adjoint spot-flux gradient vs FD (FP64 oracle, FP32 GPU)", flush=True)
    sim, rows, target, n = case["sim"], case["rows"], case["target"], case["n"]
    _, _, g_adj, _ = _adjoint_grads(case)
    orc = _oracle(sim)
    orc.set_face_flux(case["face"])
    g64 = np.zeros(len(rows))
    g32 = np.zeros(len(rows))
    for q in range(len(rows)):
        f0 = float(rows[q, 4])
        vals = {}
        orc.set_spots(rows.astype(np.float64))
        for sgn in (1, -1):
            r = rows[:, 4].astype(np.float64)
            r[q] = f0 * (1 + sgn * 1e-6)
            orc.set_flux64(spot=r)
            _oracle_load(orc, case["base"])
            orc.step(DT, n)
            vals[sgn] = _loss(orc.read(O.S), target)
        g64[q] = (vals[1] - vals[-1]) / (2e-6 * f0)
        for sgn in (1, -1):
            r = rows.copy()
            r[q, 4] = np.float32(f0 * (1 + sgn * 1e-3))
            sim.set_spots(r)
            sim.load_state(case["base"])
            sim.step(DT, n)
            vals[sgn] = _loss(sim.read("S"), target)
        g32[q] = (vals[1] - vals[-1]) / (float(np.float32(f0 * 1.001)) - float(np.float32(f0 * 0.999)))
    sim.set_spots(rows)
    e64 = np.linalg.norm(g_adj - g64) / np.linalg.norm(g64)
    e32 = np.linalg.norm(g_adj - g32) / np.linalg.norm(g64)
    print("adjoint", g_adj, "\nfp64 FD", g64, "\nfp32 FD", g32, f"\nrel err vs FP64 {e64:.3e}, vs FD32 {e32:.3e}",
          flush=True)
    assert e64 <= 1e-3
    assert e32 <= 2e-2


def _random_seeds(seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = RES
    w = {f: rng.normal(0.0, 1.0, (nz, ny, nx, 4)).astype(np.float32) for f in ("vel", "S", "S2")}
    w["vel"][..., 3] = 0.0
    return w


def _weighted(orc, w):
    """L = sum of w . (vel, S, S2) of the oracle's state."""
    return sum(float(np.sum(orc.read(fid) * w[f])) for f, fid in (("vel", O.VEL), ("S", O.S), ("S2", O.S2)))


def _smooth(shape, k):
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return (np.sin(np.pi * (x + 0.5) / nx + k) * np.sin(np.pi * (y + 0.5) / ny + 0.5 * k) *
            np.sin(2 * np.pi * (z + 0.5) / nz + 0.3 * k))


def test_face_flux_gradient_against_fp64_differences(case):
    """L = w . (vel, S, S2) after 6 steps (random weights): the fuel the floor flux adds is in L."""
    print("adjoint floor-flux gradient vs FD (FP64 oracle)", flush=True)
    sim, n, face = case["sim"], case["n"], case["face"]
    w = _random_seeds(7)
    sim.load_state(case["base"])
    adj = F.Adjoint(sim, n, 4)
    adj.forward(DT, n)
    for f in ("vel", "S", "S2"):
        adj.seed(f, w[f])
    adj.backward(solve_tol=TOL, max_cycles=40)
    _, gf = adj.grads(len(case["rows"]), face=True)
    orc = _oracle(sim)
    orc.set_spots(case["rows"].astype(np.float64))
    rng = np.random.default_rng(5)
    for label, d in (("random", np.where(face > 0, rng.normal(0.0, 1e-3, face.shape), 0.0)),
                     ("uniform", np.where(face > 0, 1e-3, 0.0))):
        pred = float(np.sum(gf * d))
        vals = {}
        orc.set_face_flux(face)
        for sgn in (1, -1):
            orc.set_flux64(face=face.astype(np.float64) + sgn * 1e-3 * d)
            _oracle_load(orc, case["base"])
            orc.step(DT, n)
            vals[sgn] = _weighted(orc, w)
        orc.set_flux64()
        fd = (vals[1] - vals[-1]) / 2e-3
        err = abs(pred - fd) / abs(fd)
        print(f"face direction {label}: adjoint {pred:.6e} FP64 FD {fd:.6e} rel err {err:.3e}", flush=True)
        assert err <= 1e-3


def test_initial_state_gradient_one_step(case):
    """Directional derivatives of L = w . (vel, S, S2) after one step along smooth directions
    of the initial state.  Scalar directions stay inside cells where the channel is positive:
    at zero every advected scalar sits on its clamp max(., 0), where the step has a one-sided
    derivative only."""
    print("adjoint initial-state gradient, 1 step, vs FP64 FD", flush=True)
    sim = case["sim"]
    w = _random_seeds(9)
    sim.load_state(case["base"])
    adj = F.Adjoint(sim, 1, 1)
    adj.forward(DT, 1)
    for f in ("vel", "S", "S2"):
        adj.seed(f, w[f])
    adj.backward(solve_tol=TOL, max_cycles=40)
    g = {f: adj.read(f).astype(np.float64) for f in ("vel", "S", "S2")}
    orc = _oracle(sim)
    orc.set_spots(case["rows"].astype(np.float64))
    orc.set_face_flux(case["face"])
    vel, S, S2, p, t, fr = case["base"]
    st0 = {"vel": vel.astype(np.float64), "S": S.astype(np.float64), "S2": S2.astype(np.float64)}
    dirs = []
    # velocity: at least 4 cells from every face.  On the floor the projection leaves v.y = 0
    # exactly (sim.js:545), so the backtrace lands on a texel centre and the limiter's texel
    # box (the floor of the backtrace position, sim.js:180-185) switches under any change of
    # v.y: a jump.  Near the faces the floor clamp, the outside fade and the revert test
    # (sim.js:137-140, 190) put kinks within reach of the differences.
    d = np.zeros_like(st0["vel"])
    for c in range(3):
        d[..., c] = 1e-2 * _smooth(vel.shape[:3], c)
    inner = np.zeros(vel.shape[:3], bool)
    inner[4:-4, 4:-4, 4:-4] = True
    d[~inner] = 0.0
    dirs.append(("vel", d))
    for f, chans in (("S", (0, 1, 2)), ("S2", (0, 1))):
        for c in chans:
            d = np.zeros_like(st0[f])
            d[..., c] = 1e-3 * _smooth(vel.shape[:3], c + 3) * (st0[f][..., c] > 1e-3)
            dirs.append((f"{f}[{c}]", d))
    for label, d in dirs:
        f = label.split("[")[0]
        pred = float(np.sum(g[f] * d))
        for h in (1e-4, 1e-5):
            vals = {}
            for sgn in (1, -1):
                st = dict(st0)
                st[f] = st0[f] + sgn * h * d
                _oracle_load(orc, (st["vel"], st["S"], st["S2"], p, t, fr))
                orc.step(DT, 1)
                vals[sgn] = _weighted(orc, w)
            fd = (vals[1] - vals[-1]) / (2 * h)
            err = abs(pred - fd) / abs(fd)
            print(f"direction {label}: adjoint {pred:.6e} FP64 FD(h={h:g}) {fd:.6e} rel err {err:.3e}", flush=True)
            assert err <= 1e-3


def _directions(case, seed):
    """A direction in every input: spot fluxes, floor flux, initial velocity (4 cells in from
    every face) and initial scalars (where the channel is positive)."""
    rng = np.random.default_rng(seed)
    rows, face = case["rows"], case["face"]
    vel, S, S2 = case["base"][:3]
    d_spot = rng.normal(0.0, 1.0, len(rows)) * 1e-3
    d_face = np.where(face > 0, rng.normal(0.0, 1.0, face.shape), 0.0) * 1e-3
    d_vel = np.zeros_like(vel)
    d_vel[4:-4, 4:-4, 4:-4, :3] = 1e-2 * rng.normal(0.0, 1.0, d_vel[4:-4, 4:-4, 4:-4, :3].shape)
    d_S = np.where(S > 1e-3, 1e-3 * rng.normal(0.0, 1.0, S.shape), 0.0)
    d_S2 = np.where(S2 > 1e-3, 1e-3 * rng.normal(0.0, 1.0, S2.shape), 0.0)
    return d_spot, d_face.astype(np.float32), d_vel.astype(np.float32), d_S.astype(np.float32), d_S2.astype(np.float32)


def test_tangent_against_fp64_differences(case):
    """Forward mode over 6 steps along a spot-flux + floor-flux direction vs FP64 differences.
    (Per-cell random directions in the scalars are left to the dot-product test: they move
    cells off exact ties, e.g. products saturated at 1 in neighbouring texels, where the step
    has one-sided derivatives only.)"""
    print("tangent vs FP64 FD, 6 steps", flush=True)
    sim, n, rows, face = case["sim"], case["n"], case["rows"], case["face"]
    w = _random_seeds(11)
    d_spot, d_face = _directions(case, 3)[:2]
    d_S = d_S2 = np.zeros(1, np.float32)
    sim.load_state(case["base"])
    adj = F.Adjoint(sim, n, 4)
    t = adj.tangent(DT, n, spot_dir=d_spot, face_dir=d_face, solve_tol=TOL, max_cycles=40)
    pred = sum(float(np.sum(t[f].astype(np.float64) * w[f])) for f in ("vel", "S", "S2"))
    orc = _oracle(sim)
    orc.set_spots(rows.astype(np.float64))
    orc.set_face_flux(face)
    vel, S, S2, p, t0, fr = case["base"]
    h = 1e-5
    vals = {}
    for sgn in (1, -1):
        orc.set_flux64(spot=rows[:, 4].astype(np.float64) + sgn * h * d_spot,
                       face=face.astype(np.float64) + sgn * h * d_face)
        _oracle_load(orc, (vel, S + sgn * h * d_S.astype(np.float64), S2 + sgn * h * d_S2.astype(np.float64), p, t0, fr))
        orc.step(DT, n)
        vals[sgn] = _weighted(orc, w)
    fd = (vals[1] - vals[-1]) / (2 * h)
    err = abs(pred - fd) / abs(fd)
    print(f"tangent {pred:.6e} FP64 FD {fd:.6e} rel err {err:.3e}", flush=True)
    assert err <= 1e-3


def test_adjoint_is_the_transpose_of_the_tangent_over_200_steps(case):
    """Dot-product test along one recorded 200-step FP32 trajectory (checkpoints every 20):
    <w, tangent(d)> = <adjoint(w), d> over spot fluxes, floor flux and the initial state."""
    print("adjoint/tangent dot product, 200 steps", flush=True)
    sim, rows = case["sim"], case["rows"]
    n = 200
    w = _random_seeds(13)
    d_spot, d_face, d_vel, d_S, d_S2 = _directions(case, 4)
    sim.load_state(case["base"])
    adj = F.Adjoint(sim, n, 20)
    t = adj.tangent(DT, n, spot_dir=d_spot, face_dir=d_face, state_dir=(d_vel, d_S, d_S2), solve_tol=TOL,
                    max_cycles=40)
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


def test_autograd_function_matches_the_c_abi(case):
    print("adjoint torch.autograd.Function", flush=True)
    torch = pytest.importorskip("torch")
    from firesim.autograd import mismatch, rollout
    sim, rows, face, n = case["sim"], case["rows"], case["face"], case["n"]
    _, loss_c, gs_c, gf_c = _adjoint_grads(case)
    sim.load_state(case["base"])
    adj = F.Adjoint(sim, n, 4)
    flux = torch.tensor(rows[:, 4].astype(np.float64), requires_grad=True)
    face_t = torch.tensor(face, requires_grad=True)
    vel, S, S2 = rollout(adj, DT, n, flux, face_t, solve_tol=TOL)
    tT = torch.tensor(case["target"][0])
    tS = torch.tensor(case["target"][1])
    loss = mismatch(S.double(), tT, tS)
    loss.backward()
    assert abs(loss.item() - loss_c) <= 1e-6 * loss_c
    gs = flux.grad.numpy()
    gf = face_t.grad.double().numpy()
    assert np.linalg.norm(gs - gs_c) <= 1e-4 * np.linalg.norm(gs_c)
    assert np.linalg.norm(gf - gf_c) <= 1e-4 * np.linalg.norm(gf_c)


def test_adjoint_requirements_are_checked():
    print("adjoint requirement errors", flush=True)
    sim16 = _sim(precision=F.PREC_FP16)
    with pytest.raises(F.FsError, match="FP32"):
        F.Adjoint(sim16, 4, 2)
    simv = _sim(solve=F.SOLVE_VCYCLES)
    with pytest.raises(F.FsError, match="REFINE64"):
        F.Adjoint(simv, 4, 2)
    sim = _sim()
    adj = F.Adjoint(sim, 4, 2)
    with pytest.raises(F.FsError, match="forward"):
        adj.backward()
    with pytest.raises(F.FsError):
        adj.forward(DT, 5)
    adj.forward(DT, 2)
    with pytest.raises(F.FsError, match="backward"):
        adj.read("S")
