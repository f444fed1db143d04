"""GPU parity on the reference's non-default parameter branches, against the CPU oracle.

The default parameter block leaves several branches of the source kernel and of the
semi-Lagrangian pass switched off.  These tests turn them on and compare the HIP library
with the oracle, bit for bit, stage by stage and over whole steps:

  * FINAL: the production preset of tools/render_final.mjs:21 (visc 2 -> the hot-gas
    viscous Laplacian, src/sim.js:339-346; lamAge 0.2 -> the laminar gate, :350;
    srcFloor -0.3 -> the pyrolysis cut-off, :253-254; anchorFreq / anchorGain / fuelSrc;
    vortEps 4; turbAmp / turbFreq / turbEvolve);
  * burst: a live burst puff centred in the flame (src/sim.js:288-296);
  * wind: ambient wind and a gust (the inflow of adv1, :153-154; the relaxation, :372-375;
    the side sponge toward the wind, :382-385);
  * all: the three together.

Every case runs with the 128^3 noise and curl textures of the reference app
(src/main.js:36), so the source kernel's repeat-addressed samples wrap a 128-texel period.
The noise textures are copied from the GPU into the oracle (the texture bake itself is
compared in test_gpu_parity.py), then both sides run from the identical state.
"""
import numpy as np
import pytest

import firesim as F
import fsoracle as O
from helpers import DT, copy_state, make_pair, rel_l2

pytestmark = pytest.mark.gpu

FINAL = dict(visc=2.0, lamAge=0.2, anchorFreq=1.0, srcFloor=-0.3, anchorGain=2.7, fuelSrc=0.15, vortEps=4.0,
             turbAmp=0.4, turbFreq=0.7, turbEvolve=0.28)
BURST = dict(burstAmt=1.0, burstX=0.02, burstY=0.16, burstZ=-0.01)
WIND = dict(windX=1.6, windZ=-0.9, gust=0.7)
CASES = {"final": FINAL, "burst": BURST, "wind": WIND, "all": {**FINAL, **BURST, **WIND}}

PRECS = [F.PREC_FP32, F.PREC_FP16, F.PREC_BF16]
PREC_IDS = ["fp32", "fp16", "bf16"]
NOISE_N = 128


def _same(g, o):
    """Bit-identical values (both sides hold exactly representable storage values)."""
    g = np.asarray(g, np.float64)
    o = np.asarray(o, np.float64)
    return g.shape == o.shape and np.array_equal(g, o)


def _describe(g, o):
    d = np.abs(np.asarray(g, np.float64) - np.asarray(o, np.float64))
    return f"differing values {int(np.count_nonzero(d))}, max |diff| {d.max():.3e}, rel L2 {rel_l2(g, o):.3e}"


STAGES = [(F.ST_ADV1, (F.VELX, F.SX, F.S2X)), (F.ST_ADV2, (F.VELY, F.SY, F.S2Y)),
          (F.ST_FORCES, (F.VELX, F.S, F.S2)), (F.ST_PROJECT, (F.VEL,))]
STAGE_IDS = ["adv1", "adv2", "forces", "project"]


@pytest.mark.parametrize("prec", PRECS, ids=PREC_IDS)
@pytest.mark.parametrize("case", list(CASES), ids=list(CASES))
def test_stage_parity_params(case, prec):
    """adv1, adv2, forces and project under the case's parameters: bit-identical to the oracle."""
    params = CASES[case]
    print(f"stage parity case {case} prec {prec} noise {NOISE_N}", flush=True)
    gpu, orc = make_pair(prec, warm_steps=30, noise_n=NOISE_N, params=params)
    for stage, outs in STAGES:
        copy_state(gpu, orc)
        gpu.run_stage(stage, DT)
        orc.stage(stage, DT)
        gpu.sync()
        for f in outs:
            g, o = gpu.read(f, np.float64), orc.read(f)
            assert np.all(np.isfinite(g))
            print(f"  stage {stage} field {f}: {_describe(g, o)}", flush=True)
            assert _same(g, o), (stage, f, _describe(g, o))
        # the next stage starts from the GPU's outputs of this one (copied again above)
    # the branches really ran: the burst heats its puff, the wind moved the side cells
    s = gpu.read(F.S, np.float64)
    assert s[..., 0].max() > 0.3


@pytest.mark.parametrize("prec", PRECS, ids=PREC_IDS)
@pytest.mark.parametrize("case", list(CASES), ids=list(CASES))
def test_step_parity_params(case, prec):
    """Whole steps under the case's parameters: bit-identical after 1 and after 10 steps."""
    params = CASES[case]
    print(f"step parity case {case} prec {prec} noise {NOISE_N}", flush=True)
    gpu, orc = make_pair(prec, warm_steps=30, noise_n=NOISE_N, params=params)
    for nsteps in (1, 9):
        gpu.step(DT, nsteps)
        orc.step(DT, nsteps)
        gpu.sync()
        assert abs(gpu.time - orc.time) < 1e-12 and gpu.frame == orc.frame
        for f in (F.VEL, F.S, F.S2, F.P):
            g, o = gpu.read(f, np.float64), orc.read(f)
            print(f"  after {gpu.frame - 30} steps field {f}: {_describe(g, o)}", flush=True)
            assert _same(g, o), (nsteps, f, _describe(g, o))


def test_params_change_the_step():
    """Each case changes the result of a step from the default block (the branches are live)."""
    print("params change the step", flush=True)
    base, _ = make_pair(F.PREC_FP32, warm_steps=30, noise_n=NOISE_N)
    base.step(DT, 1)
    ref = {f: base.read(f, np.float64) for f in (F.VEL, F.S)}
    for case, params in CASES.items():
        gpu, _ = make_pair(F.PREC_FP32, warm_steps=30, noise_n=NOISE_N, params=params)
        gpu.step(DT, 1)
        d = max(rel_l2(gpu.read(f, np.float64), ref[f]) for f in ref)
        print(f"  {case}: rel L2 vs default {d:.3e}", flush=True)
        assert d > 1e-3, case


class _OracleHost:
    """The capture loop's view of a sim (set_params(dict), set_logs, reset, step) over the oracle."""

    def __init__(self, orc):
        self.orc = orc
        self.names = set(O.param_names())

    def set_params(self, p):
        self.orc.set_params(**{k: float(v) for k, v in p.items() if k in self.names})

    def set_logs(self, logs):
        self.orc.set_logs(logs)

    def reset(self):
        self.orc.reset()

    def step(self, dt, n=1):
        self.orc.step(dt, n)


def test_capture_loop_matches_oracle():
    """The deterministic capture loop (src/main.js:153-182) with the FINAL preset, pops, a log shift
    (geometry rebuilt) and a gust, driving the GPU and the oracle with the same per-step parameters:
    bit-identical fields and geometry after every frame."""
    import firesim.driver as D
    print("capture loop gpu vs oracle 32x48x32 noise 128", flush=True)
    events = sorted([dict(t=0.05, type="pop", strength=1.0, pos=(0.06, 0.14, 0.12), aim=(0.5, 0, 0.86)),
                     dict(t=0.1, type="log-shift", strength=1.0, pos=(0.05, 0.12, 0.0), log=3),
                     dict(t=0.0, type="gust", strength=1.0, pos=(0, 0.3, 0), dur=0.2),
                     dict(t=0.15, type="crackle-burst", strength=0.5, pos=(-0.10, 0.08, 0.06))],
                    key=lambda e: e["t"])
    res, size = (32, 48, 32), (1.2, 1.8, 1.2)
    logs, _ = F.build_campfire(7)
    gpu = F.FireSim(res=res, size=size, logs=logs, precision=F.PREC_FP16, noise_n=NOISE_N, use_graph=True)
    gpu.set_params(FINAL)
    orc = O.OracleSim(res=res, size=size, logs=logs, n_logs=16, precision=O.FP16, noise_n=NOISE_N)
    orc.set_params(**FINAL)
    orc.write(O.NOISE, gpu.read(F.NOISE, np.float64))
    orc.write(O.CURL, gpu.read(F.CURL, np.float64))
    cap_g = D.Capture(gpu, None, logs, sim_dt=1 / 240, events=events)
    cap_o = D.Capture(_OracleHost(orc), None, logs, sim_dt=1 / 240, events=events)
    cap_g.begin(preroll=0.1)
    cap_o.begin(preroll=0.1)
    burst_seen = False
    for i in range(6):
        cap_g.frame(i)
        cap_o.frame(i)
        gpu.sync()
        burst_seen |= gpu.params["burstAmt"] > 0
        assert gpu.frame == orc.frame and abs(gpu.time - orc.time) < 1e-12
        for f in (F.VEL, F.S, F.S2, F.GEO):
            g, o = gpu.read(f, np.float64), orc.read(f)
            assert _same(g, o), (i, f, _describe(g, o))
        print(f"  frame {i}: steps {gpu.frame}, identical", flush=True)
    assert cap_g.log_uploads == cap_o.log_uploads >= 2
    assert burst_seen
