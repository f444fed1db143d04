"""GPU parity: every kernel of the step, and whole steps, against the CPU oracle.

The HIP library is driven through its C ABI (libfiresim.so); the oracle is the C++
restatement of src/sim.js.  Both start from the identical state (the GPU state after
a few steps of the campfire from rest, copied into the oracle), run the same stage
or step, and are compared field by field.

Every stage and whole steps (1 and 10) are compared bit for bit: both sides run the same f32
operations in the same order (no FMA contraction on the GPU), the same exp/log/pow
constructions and the same round-to-nearest-even stores, so the storage values agree exactly.
The relative L2 is printed beside each comparison.
"""
import numpy as np
import pytest

import firesim as F
import fsoracle as O
from helpers import DT, TEST_RES, TEST_SIZE, copy_state, make_pair, rel_l2

pytestmark = pytest.mark.gpu

PRECS = [F.PREC_FP32, F.PREC_FP16, F.PREC_BF16]
PREC_IDS = ["fp32", "fp16", "bf16"]


def _f16_ulp_close(a, b, ulps=2):
    """|a - b| within `ulps` binary16 ulps of |b| (or 2^-24 absolute near zero)."""
    b = np.asarray(b, np.float64)
    ulp = np.maximum(np.abs(b), 2.0 ** -14) * 2.0 ** -10
    return np.abs(np.asarray(a, np.float64) - b) <= ulps * ulp + 2.0 ** -24


@pytest.mark.parametrize("n", [32, 128])
def test_noise_textures_match_oracle(n):
    """n^3 noise + curl bake (src/noise.js:8-80) on the GPU vs the oracle's bake; 128 is the
    reference app's size (src/main.js:36)."""
    print(f"noise bake n={n} seed=1", flush=True)
    gpu = F.FireSim(res=(8, 12, 8), size=(0.1, 0.15, 0.1), noise_n=n, noise_seed=1)
    noise_o, curl_o = O.bake_noise(n, 1)
    noise_g = gpu.read(F.NOISE)
    curl_g = gpu.read(F.CURL)
    assert noise_g.shape == noise_o.shape
    same = np.mean(noise_g == noise_o)
    assert same > 0.999, same
    assert np.all(_f16_ulp_close(noise_g, noise_o, 1))
    # curl amplifies noise differences by N/2 (a 1-ulp input difference is 16 output units)
    assert np.mean(curl_g == curl_o) > 0.99
    assert rel_l2(curl_g, curl_o) < 1e-3


def test_geometry_matches_oracle():
    """Emitter geometry kernel (src/sim.js:498-523) for the seed-7 campfire."""
    print("geo 64x96x64", flush=True)
    gpu, orc = make_pair(F.PREC_FP32, warm_steps=0)
    gpu.run_stage(F.ST_GEO)
    orc.stage(O.ST_GEO)
    g = gpu.read(F.GEO, np.float64)
    o = orc.read(O.GEO)
    assert np.max(np.abs(g - o)) < 1e-5
    assert rel_l2(g, o) < 1e-6
    assert g[..., 0].max() > 0.5          # pyrolysis band present
    assert g[..., 2].sum() > 100          # solid wood cells present


STAGES = [(F.ST_ADV1, (F.VELX, F.SX, F.S2X)), (F.ST_ADV2, (F.VELY, F.SY, F.S2Y)), (F.ST_CURL, (F.VORT,)),
          (F.ST_FORCES, (F.VELX, F.S, F.S2)), (F.ST_DIV, (F.B,)), (F.ST_MG, (F.P,)),
          (F.ST_PROJECT, (F.VEL,))]
STAGE_IDS = ["adv1", "adv2", "curl", "forces", "div", "mg", "project"]


@pytest.mark.parametrize("prec", PRECS, ids=PREC_IDS)
@pytest.mark.parametrize("stage,outs", STAGES, ids=STAGE_IDS)
def test_stage_parity(prec, stage, outs):
    print(f"stage {stage} prec {prec}", flush=True)
    gpu, orc = make_pair(prec, warm_steps=30)
    # make the intermediates consistent with the state: one step's adv1/adv2/curl/forces/div
    for st in (F.ST_ADV1, F.ST_ADV2, F.ST_CURL, F.ST_FORCES, F.ST_DIV):
        if st == stage:
            break
        gpu.run_stage(st, DT)
    gpu.sync()
    copy_state(gpu, orc)
    gpu.run_stage(stage, DT)
    orc.stage(stage, DT)
    gpu.sync()
    for f in outs:
        g = gpu.read(f, np.float64)
        o = orc.read(f)
        err = rel_l2(g, o)
        print(f"  field {f}: rel L2 {err:.3e}, differing {int(np.count_nonzero(g != o))}", flush=True)
        assert np.all(np.isfinite(g))
        assert np.array_equal(g, o), (f, err)


@pytest.mark.parametrize("prec", PRECS, ids=PREC_IDS)
def test_step_parity(prec):
    """Whole steps (src/sim.js:627-648) after the fire has developed for 30 steps."""
    print(f"step parity prec {prec}", flush=True)
    gpu, orc = make_pair(prec, warm_steps=30)
    for nsteps in (1, 9):
        gpu.step(DT, nsteps)
        orc.step(DT, nsteps)
        gpu.sync()
        assert abs(gpu.time - orc.time) < 1e-12 and gpu.frame == orc.frame
        for f in (F.VEL, F.S, F.S2, F.P):
            g, o = gpu.read(f, np.float64), orc.read(f)
            err = rel_l2(g, o)
            print(f"  after {gpu.frame - 30} steps field {f}: rel L2 {err:.3e}", flush=True)
            assert np.array_equal(g, o), (nsteps, f, err)


def test_refine_reaches_tolerance():
    """FP32 V-cycles inside FP64 defect correction reach ||b - A p||/||b|| <= 1e-5 each step."""
    print("refine64", flush=True)
    gpu, orc = make_pair(F.PREC_FP16, warm_steps=20, solve=F.SOLVE_REFINE64, div_tol=1e-5, max_vcycles=12)
    for _ in range(3):
        gpu.step(DT, 1)
        orc.step(DT, 1)
        rel, used = gpu.residual()
        orel, oused = orc.residual()
        print(f"  gpu rel {rel:.3e} used {used}; oracle rel {orel:.3e} used {oused}", flush=True)
        assert rel <= 1e-5 and used >= 1
        assert abs(used - oused) <= 1
    for f in (F.VEL, F.S, F.P):
        assert rel_l2(gpu.read(f, np.float64), orc.read(f)) < 1e-3


def test_graph_replay_matches_eager():
    """A step replayed from the captured HIP graph equals the eagerly launched step, bit for bit."""
    print("graph vs eager", flush=True)
    a = F.FireSim(res=TEST_RES, size=TEST_SIZE, noise_n=32, use_graph=False)
    b = F.FireSim(res=TEST_RES, size=TEST_SIZE, noise_n=32, use_graph=True)
    for _ in range(4):   # several replays with host work and synchronisation between them
        a.step(DT, 5)
        b.step(DT, 5)
        a.sync()
        b.sync()
        for f in (F.VEL, F.S, F.S2, F.P):
            assert np.array_equal(a.read(f, "native"), b.read(f, "native")), f
    assert a.time == b.time and a.frame == b.frame


def test_deterministic_rerun():
    """Two runs from reset give identical bits (the reference's determinism check, results/result-new-determinism.json)."""
    sims = []
    for _ in range(2):
        s = F.FireSim(res=TEST_RES, size=TEST_SIZE, noise_n=32)
        s.step(DT, 20)
        sims.append(s)
    for f in (F.VEL, F.S, F.S2, F.P):
        assert np.array_equal(sims[0].read(f, "native"), sims[1].read(f, "native"))
    sims[0].reset()
    sims[0].step(DT, 20)
    for f in (F.VEL, F.S, F.S2):
        assert np.array_equal(sims[0].read(f, "native"), sims[1].read(f, "native"))


@pytest.mark.parametrize("direct", [0, -1], ids=["exact-coarse", "full-V"])
@pytest.mark.parametrize("sweeps", [(0, 0), (1, 2)], ids=["V22", "V12"])
def test_tol32_solve_parity(sweeps, direct):
    """FP32 cycles until the FP64 residual reaches the tolerance: same cycle counts and bits as the oracle,
    with the coarse level solved exactly by separable transforms (the default) and with the V-cycle
    down to the coarsest level."""
    print(f"tol32 sweeps {sweeps} direct {direct}", flush=True)
    gpu, orc = make_pair(F.PREC_FP32, warm_steps=20, solve=F.SOLVE_TOL32, div_tol=1e-4, max_vcycles=8, sweeps=sweeps,
                         direct_level=direct)
    for _ in range(3):
        gpu.step(DT, 1)
        orc.step(DT, 1)
        rel, used = gpu.residual()
        orel, oused = orc.residual()
        print(f"  gpu rel {rel:.3e} used {used}; oracle rel {orel:.3e} used {oused}", flush=True)
        assert used == oused and abs(rel - orel) <= 1e-9 * max(orel, 1e-30) + 1e-15
    for f in (F.VEL, F.S, F.S2, F.P):
        assert rel_l2(gpu.read(f, np.float64), orc.read(f)) == 0.0, f


@pytest.mark.parametrize("solve,first,tol", [(F.SOLVE_REFINE64, 0, 1e-5), (F.SOLVE_REFINE64, 1, 1e-7),
                                             (F.SOLVE_REFINE64, 3, 1e-8), (F.SOLVE_TOL32, 0, 1e-5)],
                         ids=["refine64", "refine64-first1", "refine64-tol1e-8", "tol32"])
def test_adaptive_and_async_solves_agree(solve, first, tol):
    """Host-adaptive iteration counts (one sync per step) and the fully asynchronous form
    (max_vcycles iterations, early exit on the device flag) give identical bits, graph or not.  The
    graph-replayed adaptive refinement closes its last correction inside the projection kernel; the
    tighter tolerances and one-cycle first corrections make steps take several corrections, so the
    extra-correction parts (right-hand side formed first, residual left to the projection) run too.  Steps
    go three to a call, so the next step's first advection pass runs ahead of the stopping test."""
    print(f"adaptive vs async solve {solve} first {first} tol {tol}", flush=True)
    sims = [F.FireSim(res=TEST_RES, size=TEST_SIZE, noise_n=32, precision=F.PREC_FP16, solve=solve, div_tol=tol,
                      max_vcycles=10, use_graph=g, adaptive=a, first_cycles=first)
            for a, g in ((True, True), (False, True), (True, False))]
    used = []
    for _ in range(5):
        for s in sims:
            s.step(DT, 3)
        r = [s.residual() for s in sims]
        assert r[0] == r[1] == r[2], r
        used.append(r[0][1])
    print(f"  V-cycles of each call's last step {used}", flush=True)
    if first == 1:
        assert min(used) >= 2                    # several one-cycle corrections per step
    for f in (F.VEL, F.S, F.S2, F.P):
        a = sims[0].read(f, "native")
        for s in sims[1:]:
            assert np.array_equal(a, s.read(f, "native")), f


def test_refine_steps_run_ahead_equal_one_step_per_call():
    """The refinement's steps of one call run one ahead of the host's stopping test (a step whose predecessor
    needs another correction skips its stores and is enqueued again); one step per call runs unpipelined.  With
    a tolerance that makes steps take more corrections than the last step did, twelve steps as four calls of three
    equal twelve calls of one, bit for bit, with graphs and without."""
    print("refine run-ahead vs one step per call", flush=True)
    for g in (True, False):
        sims = [F.FireSim(res=TEST_RES, size=TEST_SIZE, noise_n=32, precision=F.PREC_FP16, solve=F.SOLVE_REFINE64,
                          div_tol=1e-8, max_vcycles=10, use_graph=g) for _ in range(2)]
        for _ in range(4):
            sims[0].step(DT, 3)
            for _ in range(3):
                sims[1].step(DT, 1)
            assert sims[0].residual() == sims[1].residual()
        for f in (F.VEL, F.S, F.S2, F.P):
            assert np.array_equal(sims[0].read(f, "native"), sims[1].read(f, "native")), (g, f)
        for s in sims:
            s.close()


@pytest.mark.parametrize("res", [(64, 72, 80), (96, 88, 48)], ids=["32x36x40", "48x44x24"])
def test_tol32_exact_coarse_partial_tiles(res):
    """The exact coarse solve on levels whose y and z are not multiples of 16 (32x36x40, 48x44x24): the
    matrix-core products then run padded 16 x 16 tiles and 8-column inverse blocks.  Same cycle counts and
    bits as the oracle's FMA chains.  direct_level=1 is refused unless the kernels take the level."""
    print(f"tol32 exact coarse level on {res}", flush=True)
    dx = TEST_SIZE[0] / TEST_RES[0]
    size = tuple(n * dx for n in res)
    gpu, orc = make_pair(F.PREC_FP32, res=res, size=size, warm_steps=10, solve=F.SOLVE_TOL32, div_tol=1e-4,
                         max_vcycles=8, sweeps=(1, 2), direct_level=1)
    for _ in range(2):
        gpu.step(DT, 1)
        orc.step(DT, 1)
        rel, used = gpu.residual()
        orel, oused = orc.residual()
        print(f"  gpu rel {rel:.3e} used {used}; oracle rel {orel:.3e} used {oused}", flush=True)
        assert used == oused and abs(rel - orel) <= 1e-9 * max(orel, 1e-30) + 1e-15
    for f in (F.VEL, F.S, F.S2, F.P):
        assert rel_l2(gpu.read(f, np.float64), orc.read(f)) == 0.0, f


def test_direct_level_choice_and_refusal():
    """Automatic choice: level 1 of 64x96x64 (32x48x32); an explicit level that the kernels do not take
    is refused at create; the plain V-cycle and the exact coarse level reach the same tolerance."""
    print("direct level choice", flush=True)
    with pytest.raises(F.FsError):
        F.FireSim(res=TEST_RES, size=TEST_SIZE, noise_n=32, solve=F.SOLVE_REFINE64, direct_level=4)
    sims = [F.FireSim(res=TEST_RES, size=TEST_SIZE, noise_n=32, precision=F.PREC_FP16, solve=F.SOLVE_REFINE64,
                      div_tol=1e-6, max_vcycles=12, direct_level=d) for d in (0, 1, -1)]
    for s in sims:
        s.step(DT, 15)
    r = [s.residual() for s in sims]
    assert r[0] == r[1] and all(rel <= 1e-6 for rel, _ in r), r
    p = [s.read(F.P, np.float64) for s in sims]
    assert np.array_equal(p[0], p[1]) and rel_l2(p[2], p[0]) < 1e-3
