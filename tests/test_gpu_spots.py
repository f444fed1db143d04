"""GPU parity of the spot and floor-flux sources (fs_set_spots / fs_set_face_flux).

Spot-fed scenes (no logs) run on the GPU through the C ABI and, from the identical state,
in the oracle; the source stage and whole steps are compared bit for bit, the spot mass
normalisation must be the oracle's, and the injected fuel mass must equal flux*dt.
"""
import numpy as np
import pytest

import firesim as F
import fsoracle as O
from firesim import scenes
from helpers import PREC_TO_ORACLE, copy_state, rel_l2

pytestmark = pytest.mark.gpu

PRECS = [F.PREC_FP32, F.PREC_FP16, F.PREC_BF16]
PREC_IDS = ["fp32", "fp16", "bf16"]
RHO0 = 1.204


def scene_pair(name, precision, res=(64, 96, 64), warm_steps=20, params=None, seed=7, use_graph=False):
    """A GPU sim of a spot scene stepped `warm_steps`, and an oracle holding the identical state."""
    sc = scenes.SCENES[name]
    gpu = scenes.make_scene(name, seed=seed, res=res, precision=precision, noise_n=32, use_graph=use_graph)
    if params:
        gpu.set_params(params)
    if warm_steps:
        gpu.step(sc["dt"], warm_steps)
    gpu.sync()
    size = tuple(float(np.float32(n * sc["dx"])) for n in res)
    orc = O.OracleSim(res=res, size=size, logs=np.zeros(8, np.float32), n_logs=0,
                      precision=PREC_TO_ORACLE[precision], vcycles=sc["vcycles"], max_levels=sc["max_levels"],
                      noise_n=32)
    p = scenes.scene_params(name)
    p.update(params or {})
    orc.set_params(**p)
    orc.set_spots(scenes.scene_spots(name, seed))
    face = scenes.scene_face_flux(name)
    if face is not None and tuple(res) == tuple(sc["res"]):
        orc.set_face_flux(face)
    copy_state(gpu, orc)
    return gpu, orc


def test_spot_normalisation_matches_oracle():
    print("spot norms spots8_128 @64x96x64", flush=True)
    gpu, orc = scene_pair("spots8_128", F.PREC_FP32, warm_steps=0)
    g = gpu.spots()
    o = orc.set_spots(scenes.scene_spots("spots8_128", 7))
    assert g.shape == (8, 7)
    assert np.array_equal(g[:, 6].astype(np.float32), o.astype(np.float32))


@pytest.mark.parametrize("prec", PRECS, ids=PREC_IDS)
def test_source_stage_with_spots_and_face_flux_bit_exact(prec):
    print(f"forces adjoint8_64 prec {prec}", flush=True)
    gpu, orc = scene_pair("adjoint8_64", prec, warm_steps=25)
    for st in (F.ST_ADV1, F.ST_ADV2, F.ST_CURL):
        gpu.run_stage(st, 0.002)
    gpu.sync()
    copy_state(gpu, orc)
    gpu.run_stage(F.ST_FORCES, 0.002)
    orc.stage(O.ST_FORCES, 0.002)
    for f in (F.VELX, F.S, F.S2):
        assert np.array_equal(gpu.read(f, np.float64), orc.read(f)), f
    assert gpu.read(F.S)[..., 1].max() > 0.01          # fuel present


@pytest.mark.parametrize("prec", [F.PREC_FP32, F.PREC_FP16], ids=["fp32", "fp16"])
@pytest.mark.parametrize("name", ["spots4_64", "adjoint8_64"])
def test_steps_with_spots_match_oracle(prec, name):
    print(f"steps {name} prec {prec}", flush=True)
    sc = scenes.SCENES[name]
    gpu, orc = scene_pair(name, prec, warm_steps=30)
    gpu.step(sc["dt"], 1)
    orc.step(sc["dt"], 1)
    gpu.sync()
    for f in (F.VEL, F.S, F.S2):
        assert np.array_equal(gpu.read(f, np.float64), orc.read(f)), f
    gpu.step(sc["dt"], 9)
    orc.step(sc["dt"], 9)
    gpu.sync()
    for f in (F.VEL, F.S, F.S2):
        assert rel_l2(gpu.read(f, np.float64), orc.read(f)) <= 1e-4, f
    assert gpu.read(F.S)[..., 0].max() > 0.3            # burning: T > 593 K somewhere


def test_injected_fuel_mass_is_flux_times_dt_on_gpu():
    print("mass spots4_64 fp32", flush=True)
    sc = scenes.SCENES["spots4_64"]
    gpu = scenes.make_scene("spots4_64", res=(64, 96, 64), precision=F.PREC_FP32, noise_n=32)
    gpu.set_params(kMix=0.0, fuelDissip=0.0, turbAmp=0.0)
    gpu.run_stage(F.ST_GEO)
    gpu.run_stage(F.ST_FORCES, sc["dt"])
    Y = gpu.read(F.S, np.float64)[..., 1]
    got = Y.sum() * RHO0 * sc["dx"] ** 3
    want = float(scenes.scene_spots("spots4_64")[:, 4].astype(np.float64).sum()) * sc["dt"]
    assert abs(got - want) <= 2e-5 * want, (got, want)


def test_graph_replay_with_spots_matches_eager():
    print("graph vs eager spots8_128 @64x96x64", flush=True)
    a = scenes.make_scene("spots8_128", res=(64, 96, 64), precision=F.PREC_FP16, noise_n=32, use_graph=True)
    b = scenes.make_scene("spots8_128", res=(64, 96, 64), precision=F.PREC_FP16, noise_n=32, use_graph=False)
    for _ in range(3):
        a.step(0.001, 5)
        b.step(0.001, 5)
        a.sync()
        b.sync()
        for f in (F.VEL, F.S, F.S2):
            assert np.array_equal(a.read(f, np.float64), b.read(f, np.float64))
    # changing the spots between replays takes effect without a re-capture
    a.set_spots(scenes.scene_spots("spots8_128", 3))
    b.set_spots(scenes.scene_spots("spots8_128", 3))
    a.step(0.001, 5)
    b.step(0.001, 5)
    a.sync()
    b.sync()
    assert np.array_equal(a.read(F.S, np.float64), b.read(F.S, np.float64))
