"""CPU tests of the spot and floor-flux sources (build extension of the source kernel).

A spot injects fuel mass flux*dt per step into the cells of its sphere with the weight
w = 1 - smoothstep(r/2, r, d), normalised so the injected mass is exact; the floor flux
adds flux*dt/(rho0*dx) to the fuel fraction of each floor cell.  With burning and fuel
dissipation switched off (kMix = 0, fuelDissip = 0) the fuel mass added by one source
pass is therefore known in closed form: sum(flux_k)*dt for the spots and
sum(face)*dx^2*dt for the floor.  These pin the oracle's restatement of both terms.
The scene layouts are pinned against the reference's generator (src/scene.js:13-16).
"""
import numpy as np
import pytest

import fsoracle as O
from firesim import scenes

RHO0 = 1.204


def _scene_oracle(name, precision=O.FP64, res=(32, 48, 32), **params):
    dx = scenes.SCENES[name]["dx"]
    size = tuple(float(np.float32(n * dx)) for n in res)
    orc = O.OracleSim(res=res, size=size, logs=np.zeros(8, np.float32), n_logs=0, precision=precision, noise_n=32)
    p = scenes.scene_params(name)
    p.update(params)
    orc.set_params(**p)
    return orc, dx


def _fuel_mass(orc, dx):
    return orc.read(O.S)[..., 1].sum() * RHO0 * dx ** 3


@pytest.mark.parametrize("precision,tol", [(O.FP64, 1e-6), (O.FP32, 2e-5)])
def test_spot_injection_mass_is_flux_times_dt(precision, tol):
    orc, dx = _scene_oracle("spots4_64", precision, kMix=0.0, fuelDissip=0.0, turbAmp=0.0)
    spots = np.array([[0.0, 0.1, 0.0, 0.02, 2e-3, 0.75],
                      [0.07, 0.12, -0.03, 0.025, 1e-3, 0.75],
                      [-0.05, 0.08, 0.06, 0.013, 3e-3, 0.75]], np.float32)
    norms = orc.set_spots(spots)
    assert np.all(norms > 0)
    dt = 0.002
    orc.stage(O.ST_GEO)
    orc.stage(O.ST_FORCES, dt)
    want = float(spots[:, 4].astype(np.float64).sum()) * dt
    got = _fuel_mass(orc, dx)
    assert abs(got - want) <= tol * want, (got, want)


def test_spot_weights_follow_the_smoothstep_profile():
    """Inside r/2 the weight is 1: a spot's centre cell gets exactly flux*dt*norm; cells beyond r get nothing."""
    orc, dx = _scene_oracle("spots4_64", O.FP64, kMix=0.0, fuelDissip=0.0, turbAmp=0.0)
    # centre at a cell centre: cell (16, 10, 16) of a 32x48x32 grid at 1 cm with origin (-0.16, 0, -0.16)
    c = np.array([-0.16 + 16.5 * dx, 10.5 * dx, -0.16 + 16.5 * dx])
    norm = orc.set_spots([[*c, 0.03, 2e-3, 0.75]])[0]
    dt = 0.002
    orc.stage(O.ST_GEO)
    orc.stage(O.ST_FORCES, dt)
    Y = orc.read(O.S)[..., 1]
    # the kernel reads dt as the f32 uniform of src/sim.js:620-624
    assert Y[16, 10, 16] == pytest.approx(float(np.float32(2e-3)) * float(np.float32(dt)) * float(np.float32(norm)), rel=1e-12)
    z, y, x = np.nonzero(Y)
    d = np.sqrt(((x - 16) ** 2 + (y - 10) ** 2 + (z - 16) ** 2)) * dx
    assert d.max() <= 0.03 + 1e-12
    # the weight sum: norm = 1 / (rho0 * sum(w) * dx^3), with w computed independently here
    zz, yy, xx = np.meshgrid(np.arange(32), np.arange(48), np.arange(32), indexing="ij")
    dd = np.sqrt(((xx - 16) ** 2 + (yy - 10) ** 2 + (zz - 16) ** 2)) * dx
    t = np.clip((dd - 0.015) / 0.015, 0, 1)
    w = np.where(dd < 0.03, 1 - t * t * (3 - 2 * t), 0.0)
    assert norm == pytest.approx(1.0 / (RHO0 * w.sum() * dx ** 3), rel=1e-6)


def test_face_flux_mass_is_flux_times_area_times_dt():
    orc, dx = _scene_oracle("adjoint8_64", O.FP64, kMix=0.0, fuelDissip=0.0, turbAmp=0.0)
    rng = np.random.default_rng(3)
    face = (rng.random((32, 32)) * 4e-3).astype(np.float32)
    orc.set_face_flux(face)
    dt = 0.002
    orc.stage(O.ST_GEO)
    orc.stage(O.ST_FORCES, dt)
    S = orc.read(O.S)
    assert np.all(S[:, 1:, :, 1] == 0)          # only the floor row receives it
    want = float(face.astype(np.float64).sum()) * dx * dx * dt
    got = _fuel_mass(orc, dx)
    assert abs(got - want) <= 1e-6 * want
    orc.set_face_flux(None)
    orc.reset()
    orc.stage(O.ST_GEO)
    orc.stage(O.ST_FORCES, dt)
    assert _fuel_mass(orc, dx) == 0.0


def test_spot_mass_is_linear_in_flux_and_additive():
    """Two spots far apart inject the sum of what each injects alone; doubling a flux doubles its mass."""
    dt = 0.002
    a = [-0.08, 0.1, 0.0, 0.02, 1e-3, 0.75]
    b = [0.08, 0.1, 0.0, 0.02, 2e-3, 0.75]
    masses = []
    for sp in ([a], [b], [a, b]):
        orc, dx = _scene_oracle("spots4_64", O.FP64, kMix=0.0, fuelDissip=0.0, turbAmp=0.0)
        orc.set_spots(sp)
        orc.stage(O.ST_GEO)
        orc.stage(O.ST_FORCES, dt)
        masses.append(_fuel_mass(orc, dx))
    assert masses[2] == pytest.approx(masses[0] + masses[1], rel=1e-12)
    assert masses[1] == pytest.approx(2.0 * masses[0], rel=1e-6)


def test_scene_layouts_use_the_reference_lcg():
    # the generator of src/scene.js:13-16, evaluated by hand for seed 7
    R = scenes.lcg(7)
    s = 7
    for _ in range(5):
        s = (s * 1664525 + 1013904223) % 2 ** 32
        assert R() == s / 2 ** 32
    sp = scenes.scene_spots("spots4_64", 7)
    assert sp.shape == (4, 6)
    assert np.all(np.hypot(sp[:, 0], sp[:, 2]) <= 0.15 + 1e-6)
    assert np.all(sp[:, 1] == np.float32(0.1)) and np.all(sp[:, 3] == np.float32(0.02))
    assert np.all(sp[:, 4] == np.float32(2e-3))
    sp8 = scenes.scene_spots("spots8_128", 11)
    assert sp8.shape == (8, 6) and np.all((sp8[:, 4] >= 1.5e-3) & (sp8[:, 4] <= 3e-3))
    sp12 = scenes.scene_spots("segments12_192", 7)
    assert sp12.shape == (12, 6) and np.all((sp12[:, 4] >= 1e-3) & (sp12[:, 4] <= 3e-3))
    assert len(set(np.round(sp12[:, 1], 4))) == 3          # three segments at three heights
    assert scenes.scene_size("segments12_192") == pytest.approx((1.0368, 1.5552, 1.0368), rel=1e-6)
    assert scenes.scene_face_flux("adjoint8_64").shape == (64, 64)
