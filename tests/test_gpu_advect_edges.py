"""Advection on small and odd grids with velocities that carry samples across every edge, bit for bit.

The advection kernels read their fields through buffer resources, take x-neighbour texel pairs in one 16-byte
load for 8-byte texels (clamped edges selected on the raw words), and start the backtrace from the cell's own
texel.  Random velocities of up to ~3 cells per step on grids down to nx = 2 put trilinear samples and limiter
corners before the first and past the last texel on every axis (ix = -1 .. n), which the developed-fire scenes
reach only at a few cells.  adv1 and adv2 must equal the oracle's restatement exactly, in all three storages.
"""
import numpy as np
import pytest

import firesim as F
import fsoracle as O
from helpers import copy_state, make_pair

pytestmark = pytest.mark.gpu

PRECS = [F.PREC_FP32, F.PREC_FP16, F.PREC_BF16]
PREC_IDS = ["fp32", "fp16", "bf16"]
GRIDS = [(2, 6, 4), (7, 9, 5), (33, 20, 17)]


@pytest.mark.parametrize("prec", PRECS, ids=PREC_IDS)
@pytest.mark.parametrize("res", GRIDS, ids=["x".join(map(str, r)) for r in GRIDS])
def test_advection_across_edges_is_bit_exact(prec, res):
    print(f"advection edges res {res} prec {prec}", flush=True)
    dx, dt = 0.01, 0.004
    size = tuple(dx * n for n in res)
    gpu, orc = make_pair(prec, res=res, size=size, warm_steps=0, noise_n=16)
    rng = np.random.default_rng(sum(res) * 10 + prec)
    nx, ny, nz = res
    vmax = 3.0 * dx / dt                                  # 3 cells per step
    vel = rng.uniform(-vmax, vmax, (nz, ny, nx, 4)).astype(np.float32)
    vel[..., 3] = 0.0
    gpu.write(F.VEL, vel)
    gpu.write(F.S, rng.uniform(0.0, 1.0, (nz, ny, nx, 4)).astype(np.float32))
    gpu.write(F.S2, rng.uniform(0.0, 1.0, (nz, ny, nx, 4)).astype(np.float32))
    gpu.sync()
    copy_state(gpu, orc)
    # first-order backtrace positions: samples leave the domain on both sides of every axis
    v = gpu.read(F.VEL, np.float64)
    for ax, n in ((0, nx), (1, ny), (2, nz)):
        c = np.arange(n) + 0.5
        shape = [1, 1, 1]
        shape[2 - ax] = n
        pos = c.reshape(shape) - v[..., ax] * dt / dx
        assert (pos < 0.5).any() and (pos > n - 0.5).any(), ax
    for stage, outs in ((F.ST_ADV1, (F.VELX, F.SX, F.S2X)), (F.ST_ADV2, (F.VELY, F.SY, F.S2Y))):
        gpu.run_stage(stage, dt)
        orc.stage(stage, dt)
        gpu.sync()
        for f in outs:
            g, o = gpu.read(f, np.float64), orc.read(f)
            assert np.array_equal(g, o), (stage, f, int(np.sum(g != o)))
