"""Shared helpers of the parity tests: GPU/oracle pairs with identical state, error norms."""
import numpy as np

import firesim as F
import fsoracle as O

# the reference app's physical domain (src/main.js:38) on a coarse grid: dx = 1.875 cm
TEST_RES = (64, 96, 64)
TEST_SIZE = (1.2, 1.8, 1.2)
DT = 1.0 / 90.0              # src/main.js:48 (test mode simDt)

STATE_FIELDS = (F.VEL, F.S, F.S2, F.VORT, F.GEO, F.VELX, F.SX, F.S2X, F.VELY, F.SY, F.S2Y)
PREC_TO_ORACLE = {F.PREC_FP32: O.FP32, F.PREC_FP16: O.FP16, F.PREC_BF16: O.BF16}


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b.ravel())
    d = np.linalg.norm((a - b).ravel())
    return d / nb if nb > 0 else d


def copy_state(gpu, orc):
    """Copies every field, the pressure, the noise textures and the clock GPU -> oracle."""
    for f in STATE_FIELDS + (F.NOISE, F.CURL):
        orc.write(f, gpu.read(f, np.float64))
    orc.write(O.P, gpu.read(F.P, np.float64))
    orc.write(O.B, gpu.read(F.B, np.float64))
    orc.set_clock(gpu.time, gpu.frame)


def make_pair(precision=F.PREC_FP16, res=TEST_RES, size=TEST_SIZE, warm_steps=30, dt=DT, solve=F.SOLVE_VCYCLES,
              vcycles=2, div_tol=1e-5, max_vcycles=8, noise_n=32, params=None, use_graph=False, sweeps=(0, 0),
              direct_level=0):
    """A GPU sim stepped `warm_steps` from rest and an oracle holding the identical state."""
    logs, _ = F.build_campfire(7)
    gpu = F.FireSim(res=res, size=size, logs=logs, vcycles=vcycles, precision=precision, solve=solve,
                    div_tol=div_tol, max_vcycles=max_vcycles, noise_n=noise_n, use_graph=use_graph,
                    pre_sweeps=sweeps[0], post_sweeps=sweeps[1], direct_level=direct_level)
    if params:
        gpu.set_params(params)
    if warm_steps:
        gpu.step(dt, warm_steps)
    gpu.sync()
    orc = O.OracleSim(res=res, size=size, logs=logs, n_logs=16, precision=PREC_TO_ORACLE[precision], vcycles=vcycles,
                      solve=solve, div_tol=div_tol, max_vcycles=max_vcycles, noise_n=noise_n, pre_sweeps=sweeps[0],
                      post_sweeps=sweeps[1], direct_level=direct_level)
    if params:
        orc.set_params(**params)
    copy_state(gpu, orc)
    return gpu, orc
