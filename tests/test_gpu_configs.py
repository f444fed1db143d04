"""The measured workloads at their full sizes, on the GPU, against the oracle and closed forms.

  * 128x192x128 fire step (8 spots of U(1.5, 3) g/s, 5 mm, dt 1 ms, vortEps 0.3): the plain
    FP32 step bit-identical to the oracle after 1 and 10 steps; the mixed step (fp16 storage,
    FP32 V-cycles in FP64 defect correction to 1e-5) within temperature relative L2 <= 1e-3 of
    the FP64 oracle after 1 and 10 steps, residual <= 1e-5 every step.
  * 256x384x256 isolated Poisson solve (64 Gaussian blobs + white noise, 6 levels, V(2,2)
    Gauss-Seidel in FP64 defect correction to 1e-8): the residual recomputed in float64 from the
    returned iterate, and the error against the exact solution by DST-II / DCT-IV transforms.
  * 64x96x64 adjoint scene (8 spots + a 64x64 floor flux, checkpoints every 20 steps): the
    50-step gradient against central differences of the FP64 oracle along the gradient
    (<= 1e-3), and the 200-step adjoint as the transpose of the tangent (dot-product test).
  * 192x288x192 fine-grid scene (12 spots on 3 segments, 5.4 mm, dt 0.75 ms, 6 levels): the
    mixed step keeps every step's residual <= 1e-5 and finite fields, and its heat release and
    burned fuel (FP64 budget) agree with the plain FP32 step's.
"""
import numpy as np
import pytest

import firesim as F
import fsoracle as O
from firesim import scenes
from helpers import copy_state, rel_l2

pytestmark = pytest.mark.gpu


def _T(S):
    return 293.0 + 1000.0 * np.asarray(S, np.float64)[..., 0]


# ------------------------------------------------------------------ 128x192x128
def _c2_oracle(precision, **kw):
    name = "spots8_128"
    sc = scenes.SCENES[name]
    orc = O.OracleSim(res=sc["res"], size=scenes.scene_size(name), logs=np.zeros(8, np.float32), n_logs=0,
                      precision=precision, vcycles=sc["vcycles"], noise_n=128, **kw)
    orc.set_params(**scenes.scene_params(name))
    orc.set_spots(scenes.scene_spots(name, 7))
    return orc


def test_c2_plain_fp32_step_bit_identical_at_full_size():
    name = "spots8_128"
    sc = scenes.SCENES[name]
    print(f"C2 plain FP32 {sc['res']} vs oracle", flush=True)
    gpu = scenes.make_scene(name, precision=F.PREC_FP32, solve=F.SOLVE_VCYCLES, noise_n=128)
    gpu.step(sc["dt"], 20)
    gpu.sync()
    orc = _c2_oracle(O.FP32)
    copy_state(gpu, orc)
    for n in (1, 9):
        gpu.step(sc["dt"], n)
        orc.step(sc["dt"], n)
        gpu.sync()
        for f in (F.VEL, F.S, F.S2, F.P):
            g, o = gpu.read(f, np.float64), orc.read(f)
            print(f"  after {gpu.frame - 20} steps field {f}: differing {int(np.count_nonzero(g != o))}", flush=True)
            assert np.array_equal(g, o), (n, f, rel_l2(g, o))
    assert _T(gpu.read(F.S, np.float64)).max() > 1000.0


def test_c2_mixed_step_error_against_fp64_oracle_at_full_size():
    name = "spots8_128"
    sc = scenes.SCENES[name]
    print(f"C2 mixed {sc['res']} vs FP64 oracle", flush=True)
    gpu = scenes.make_scene(name, precision=F.PREC_FP16, solve=F.SOLVE_REFINE64, div_tol=1e-5, max_vcycles=12,
                            noise_n=128)
    gpu.step(sc["dt"], 20)
    gpu.sync()
    orc = _c2_oracle(O.FP64, solve=O.SOLVE_REFINE64, div_tol=1e-10, max_vcycles=40)
    copy_state(gpu, orc)
    errs = {}
    for k in range(1, 11):
        gpu.step(sc["dt"], 1)
        orc.step(sc["dt"], 1)
        rel, used = gpu.residual()
        assert rel <= 1e-5, (k, rel, used)
        if k in (1, 10):
            errs[k] = {n: rel_l2(gpu.read(f, np.float64)[..., :3], orc.read(f)[..., :3])
                       for n, f in (("vel", F.VEL), ("S", F.S), ("S2", F.S2))}
            errs[k]["T"] = rel_l2(_T(gpu.read(F.S, np.float64)), _T(orc.read(O.S)))
            print(f"  step {k}: residual {rel:.2e} ({used} cycles), errors " +
                  ", ".join(f"{n} {v:.2e}" for n, v in errs[k].items()), flush=True)
    assert errs[1]["T"] <= 1e-3 and errs[10]["T"] <= 1e-3, errs
    assert _T(orc.read(O.S)).max() > 1000.0


# ------------------------------------------------------------------ 256x384x256 Poisson
def _apply_A(p, h):
    """The step's pressure operator: ghosts -p on the x / z faces and the top, +p below the floor."""
    nz, ny, nx = p.shape
    lap = -6.0 * p
    for axis, lo_sign, hi_sign in ((2, -1.0, -1.0), (0, -1.0, -1.0), (1, 1.0, -1.0)):
        P = np.moveaxis(p, axis, 0)
        nb = np.empty_like(P)
        nb[1:] = P[:-1]
        nb[0] = lo_sign * P[0]
        lap += np.moveaxis(nb, 0, axis)
        nb[:-1] = P[1:]
        nb[-1] = hi_sign * P[-1]
        lap += np.moveaxis(nb, 0, axis)
    return lap / (h * h)


def _spectral_solve(b, h):
    """Exact solution of A p = b: DST-II in x and z, DCT-IV in y (the operator's eigenbases)."""
    import scipy.fft as sfft
    nz, ny, nx = b.shape
    B = sfft.dst(b, type=2, axis=2, norm="ortho")
    B = sfft.dst(B, type=2, axis=0, norm="ortho")
    B = sfft.dct(B, type=4, axis=1, norm="ortho")
    lx = -(2 - 2 * np.cos(np.pi * (np.arange(nx) + 1) / nx))
    lz = -(2 - 2 * np.cos(np.pi * (np.arange(nz) + 1) / nz))
    ly = -(2 - 2 * np.cos(np.pi * (np.arange(ny) + 0.5) / ny))
    B /= (lz[:, None, None] + ly[None, :, None] + lx[None, None, :]) / (h * h)
    p = sfft.idct(B, type=4, axis=1, norm="ortho")
    p = sfft.idst(p, type=2, axis=0, norm="ortho")
    return sfft.idst(p, type=2, axis=2, norm="ortho")


def test_c3_isolated_poisson_at_full_size():
    dims = (256, 384, 256)
    h = 1.0 / 256
    print(f"C3 Poisson {dims}", flush=True)
    b = scenes.blob_rhs(dims)
    ps = F.Poisson(dims, h, max_levels=6)
    levels, _ = ps.info()
    assert len(levels) == 6 and levels[-1] == (8, 12, 8)
    ps.set_rhs(b)
    ps.set_x(np.zeros_like(b))
    hist = ps.solve(1e-8, 40)
    x = ps.read_x(True)
    ps.close()
    print(f"  {len(hist) - 1} cycles, history {np.array2string(hist, precision=2)}", flush=True)
    assert hist[-1] <= 1e-8 and np.all(np.diff(hist) < 0)
    # the solve's right-hand side is FP32 (the workload's): the device holds b rounded to f32
    b = b.astype(np.float32).astype(np.float64)
    nb = np.linalg.norm(b)
    rel = np.linalg.norm(b - _apply_A(x, h)) / nb
    print(f"  float64 residual of the returned iterate {rel:.3e} (device {hist[-1]:.3e})", flush=True)
    assert rel <= 1e-8 and abs(rel / hist[-1] - 1) < 1e-3
    exact = _spectral_solve(b, h)
    err = np.linalg.norm(x - exact) / np.linalg.norm(exact)
    print(f"  error against the exact solution {err:.3e}", flush=True)
    assert err <= 1e-7


# ------------------------------------------------------------------ 64x96x64 adjoint
C4 = "adjoint8_64"


@pytest.fixture(scope="module")
def c4():
    sc = scenes.SCENES[C4]
    print(f"C4 adjoint scene {sc['res']}", flush=True)
    sim = scenes.make_scene(C4, precision=F.PREC_FP32, solve=F.SOLVE_REFINE64, div_tol=1e-9, max_vcycles=40,
                            noise_n=128)
    rows = sim.spots()[:, :6].astype(np.float32)
    face = scenes.scene_face_flux(C4)
    base = sim.save_state()
    rng = np.random.default_rng(2024)
    pr = rows.copy()
    pr[:, 4] *= (1.0 + rng.normal(0.0, 0.2, len(rows))).astype(np.float32)
    pf = (face * (1.0 + rng.normal(0.0, 0.2, face.shape))).astype(np.float32)

    def target(n):
        sim.set_spots(pr)
        sim.set_face_flux(pf)
        sim.load_state(base)
        sim.step(sc["dt"], n)
        S = sim.read("S")
        sim.set_spots(rows)
        sim.set_face_flux(face)
        return np.ascontiguousarray(S[..., 0]), np.ascontiguousarray(S[..., 2])
    return dict(sim=sim, rows=rows, face=face, base=base, target=target, dt=sc["dt"], rng=rng)


def _loss(S, tgt):
    S = np.asarray(S, np.float64)
    return float(np.sum((S[..., 0] - tgt[0]) ** 2) + np.sum((S[..., 2] - tgt[1]) ** 2))


def test_c4_adjoint_50_steps_against_fp64_differences(c4):
    sim, rows, face, base, dt = c4["sim"], c4["rows"], c4["face"], c4["base"], c4["dt"]
    n = 50
    tgt = c4["target"](n)
    sim.load_state(base)
    adj = F.Adjoint(sim, n, 20)
    adj.forward(dt, n)
    adj.seed_mismatch(*tgt)
    adj.backward(solve_tol=1e-9, max_cycles=40)
    g_spot, g_face = adj.grads(len(rows), face=True)
    adj.close()
    sc = scenes.SCENES[C4]
    orc = O.OracleSim(res=sc["res"], size=scenes.scene_size(C4), logs=np.zeros(8, np.float32), n_logs=0,
                      precision=O.FP64, vcycles=sc["vcycles"], solve=O.SOLVE_REFINE64, div_tol=1e-11, max_vcycles=40,
                      noise_n=128)
    orc.set_params(**scenes.scene_params(C4))
    for f in (F.GEO, F.NOISE, F.CURL):
        orc.write(f, sim.read(f, np.float64))
    orc.set_spots(rows.astype(np.float64))
    orc.set_face_flux(face)
    vel0, S0, S20, p0, t0, fr0 = base

    def fd64(d_spot, d_face, h):
        v = {}
        for sgn in (1, -1):
            orc.set_flux64(spot=rows[:, 4].astype(np.float64) + sgn * h * d_spot,
                           face=face.astype(np.float64) + sgn * h * d_face)
            orc.write(O.VEL, vel0.astype(np.float64))
            orc.write(O.S, S0.astype(np.float64))
            orc.write(O.S2, S20.astype(np.float64))
            orc.write(O.P, p0)
            orc.set_clock(t0, fr0)
            orc.step(dt, n)
            v[sgn] = _loss(orc.read(O.S), tgt)
        return (v[1] - v[-1]) / (2 * h)

    # Along the spot gradient (scaled to the fluxes' size): relative error.  Along the floor-flux
    # gradient (scaled to the floor flux's size): after 50 steps the floor's fuel has risen only to
    # the spots' lower edge, where it meets the flame only in cells at the oxygen clamp
    # o = clamp(1 - Y - X, 0, 1) (src/sim.js:302), so that derivative is near zero and not smooth;
    # its error is measured against the spot derivative, the size of the whole gradient.
    ds = g_spot / np.linalg.norm(g_spot) * np.linalg.norm(rows[:, 4])
    pred_s = float(np.sum(g_spot * ds))
    fd_s = fd64(ds, np.zeros_like(face, np.float64), 1e-6)
    err_s = abs(pred_s - fd_s) / abs(fd_s)
    print(f"  C4 50 steps along the spot gradient: adjoint {pred_s:.9e}, FP64 differences {fd_s:.9e}, "
          f"rel err {err_s:.2e}", flush=True)
    assert err_s <= 1e-3, (pred_s, fd_s)
    if np.linalg.norm(g_face) > 0:
        df = g_face / np.linalg.norm(g_face) * np.linalg.norm(face)
        pred_f = float(np.sum(g_face * df))
        fd_f = fd64(np.zeros(len(rows)), df, 1e-6)
        err_f = abs(pred_f - fd_f) / abs(fd_s)
        print(f"  C4 50 steps along the floor gradient: adjoint {pred_f:.9e}, FP64 differences {fd_f:.9e}, "
              f"error / spot derivative {err_f:.2e}", flush=True)
        assert err_f <= 1e-3, (pred_f, fd_f, fd_s)


def test_c4_adjoint_is_the_transpose_of_the_tangent_over_200_steps(c4):
    sim, rows, face, base, dt, rng = c4["sim"], c4["rows"], c4["face"], c4["base"], c4["dt"], c4["rng"]
    n = 200
    tgt = c4["target"](n)
    sim.load_state(base)
    adj = F.Adjoint(sim, n, 20)
    adj.forward(dt, n)
    adj.seed_mismatch(*tgt)
    adj.backward(solve_tol=1e-9, max_cycles=40)
    g_spot, g_face = adj.grads(len(rows), face=True)
    d_spot = rng.normal(0.0, 1.0, len(rows)) * 1e-3
    d_face = (np.where(face > 0, rng.normal(0.0, 1.0, face.shape), 0.0) * 1e-3).astype(np.float32)
    sim.load_state(base)
    tan = adj.tangent(dt, n, spot_dir=d_spot, face_dir=d_face, solve_tol=1e-9, max_cycles=40)
    S = sim.read("S").astype(np.float64)
    jd = float(np.sum(2.0 * (S[..., 0] - tgt[0]) * tan["S"][..., 0]) + np.sum(2.0 * (S[..., 2] - tgt[1]) * tan["S"][..., 2]))
    jtw = float(np.sum(g_spot * d_spot) + np.sum(g_face * d_face))
    adj.close()
    rel = abs(jd - jtw) / max(abs(jd), abs(jtw))
    print(f"  C4 200 steps: <seed, J d> {jd:.9e}, <J^T seed, d> {jtw:.9e}, rel {rel:.2e}", flush=True)
    assert rel <= 1e-5 and abs(jd) > 0 and np.linalg.norm(g_face) > 0


# ------------------------------------------------------------------ 192x288x192
def test_c5_fine_grid_mixed_and_fp32_steps():
    name = "segments12_192"
    sc = scenes.SCENES[name]
    steps = 120
    rec = {}
    for form, prec, solve in (("mixed", F.PREC_FP16, F.SOLVE_REFINE64), ("fp32", F.PREC_FP32, F.SOLVE_VCYCLES)):
        print(f"C5 {form} {sc['res']} {steps} steps", flush=True)
        sim = scenes.make_scene(name, precision=prec, solve=solve, div_tol=1e-5, max_vcycles=12, noise_n=128)
        levels, _ = sim.mg_info()
        assert len(levels) == 6
        sim.budget_enable()
        for k in range(steps):
            sim.step(sc["dt"], 1)
            if form == "mixed":
                rel, used = sim.residual()
                assert rel <= 1e-5, (k, rel, used)
        S = sim.read(F.S, np.float64)
        assert all(np.all(np.isfinite(sim.read(f, np.float64))) for f in (F.VEL, F.S, F.S2))
        rec[form] = sim.budget()
        rec[form]["Tmax"] = _T(S).max()
        sim.close()
        print(f"  {form}: heat released {rec[form]['heat_released']:.6e} J, fuel burned "
              f"{rec[form]['fuel_burned']:.6e} kg, max T {rec[form]['Tmax']:.0f} K", flush=True)
    assert rec["mixed"]["Tmax"] > 1200
    for k in ("heat_released", "fuel_burned", "fuel_injected"):
        d = abs(rec["mixed"][k] / rec["fp32"][k] - 1)
        assert d <= 0.02, (k, rec["mixed"][k], rec["fp32"][k])
