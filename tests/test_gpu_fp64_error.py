"""Error of the GPU step against the FP64 oracle.

The FP64 oracle runs the same step with float64 storage, float64 arithmetic and its
pressure solve converged to ||b - A p|| / ||b|| <= 1e-10; it starts from the GPU's own
state (the spot-fed 1 cm scene after a warm-up), so the comparison isolates what the
GPU's precision choices do over 1 and 10 steps:
  * mixed: fp16 storage, f32 math, FP32 V-cycles in FP64 defect correction to 1e-5;
  * mixed_spectral: the same with the exact separable inner solve (FS_SOLVE_SPECTRAL64);
  * plain FP32: float4 storage, the reference's 2 x V(2,2) per step.
Bar for the mixed step: temperature relative L2 <= 1e-3 after 1 and after 10 steps,
and the divergence residual <= 1e-5 every step.
"""
import numpy as np
import pytest

import firesim as F
import fsoracle as O
from firesim import scenes
from helpers import copy_state, rel_l2

pytestmark = pytest.mark.gpu

NAME = "spots4_64"


def _temperature(S):
    return 293.0 + 1000.0 * np.asarray(S, np.float64)[..., 0]


def _errors(gpu, orc):
    out = {}
    for name, f in (("vel", F.VEL), ("S", F.S), ("S2", F.S2)):
        out[name] = rel_l2(gpu.read(f, np.float64)[..., :3], orc.read(f)[..., :3])
    out["T"] = rel_l2(_temperature(gpu.read(F.S, np.float64)), _temperature(orc.read(F.S)))
    out["theta"] = rel_l2(gpu.read(F.S, np.float64)[..., 0], orc.read(F.S)[..., 0])      # T - 293 K
    return out


def _fp64_oracle(sc, res):
    size = scenes.scene_size(NAME)
    orc = O.OracleSim(res=res, size=size, logs=np.zeros(8, np.float32), n_logs=0, precision=O.FP64,
                      vcycles=sc["vcycles"], max_levels=sc["max_levels"], solve=O.SOLVE_REFINE64, div_tol=1e-10,
                      max_vcycles=40, noise_n=32)
    orc.set_params(**scenes.scene_params(NAME))
    orc.set_spots(scenes.scene_spots(NAME, 7))
    return orc


@pytest.mark.parametrize("form", ["mixed", "mixed_spectral", "fp32"])
def test_step_error_against_fp64_oracle(form):
    print(f"fp64 error {NAME} {form}", flush=True)
    sc = scenes.SCENES[NAME]
    res = sc["res"]
    if form in ("mixed", "mixed_spectral"):
        solve = F.SOLVE_REFINE64 if form == "mixed" else F.SOLVE_SPECTRAL64
        gpu = scenes.make_scene(NAME, precision=F.PREC_FP16, solve=solve, div_tol=1e-5, max_vcycles=12, noise_n=32)
    else:
        gpu = scenes.make_scene(NAME, precision=F.PREC_FP32, solve=F.SOLVE_VCYCLES, noise_n=32)
    gpu.step(sc["dt"], 40)
    gpu.sync()
    orc = _fp64_oracle(sc, res)
    copy_state(gpu, orc)
    errs = {}
    for n_done, n in ((1, 1), (10, 9)):
        for _ in range(n):
            gpu.step(sc["dt"], 1)
            orc.step(sc["dt"], 1)
            if form != "fp32":
                rel, used = gpu.residual()
                assert rel <= 1e-5, (rel, used)
        gpu.sync()
        errs[n_done] = _errors(gpu, orc)
    print(form, {k: {f: f"{v:.2e}" for f, v in e.items()} for k, e in errs.items()}, flush=True)
    assert _temperature(orc.read(O.S)).max() > 600.0          # a burning state
    if form != "fp32":
        assert errs[1]["T"] <= 1e-3 and errs[10]["T"] <= 1e-3, errs
    else:
        assert errs[1]["T"] <= 1e-2, errs                       # reported; the reference's solve is not converged
