"""CPU tests of the oracle's FP64 mass and energy budget (the checker of fs_budget_*).

* The fuel and heat records telescope: the totals after the steps equal the totals at enable
  plus the recorded terms, to FP64 summation rounding.
* The fuel injected by the spots and the floor flux equals sum(flux) dt + sum(F) dx^2 dt (the
  sources' mass normalisation), to f32 rounding, where no cell saturates.
* Heat released = cp heatOut x fuel burned where no cell reaches the 2100 K cap
  (src/sim.js:306: dT = min(heatOut b, 2100 - T0)), never more.
* The totals at enable are the plain sums of the state.
"""
import numpy as np

import firesim as F
import fsoracle as O
from firesim import scenes

K = {k: i for i, k in enumerate(F.BUDGET_KEYS)}
CP = 1005.0
RHO0 = float(np.float32(F.RHO0))     # FS_RHO0 is an f32 constant of the C ABI


def _scene_oracle(precision=O.FP32, res=(32, 48, 32), face=True, params=None):
    name = "adjoint8_64"
    sc = scenes.SCENES[name]
    size = tuple(float(np.float32(n * sc["dx"])) for n in res)
    o = O.OracleSim(res=res, size=size, logs=np.zeros(8, np.float32), n_logs=0, precision=precision, noise_n=32)
    o.set_params(**{**scenes.scene_params(name), **(params or {})})
    o.set_spots(scenes.scene_spots(name, 7))
    if face:
        o.set_face_flux(scenes.scene_face_flux(name, res=res))
    return o, sc["dt"]


def _check_identities(b, rtol=1e-12):
    fuel = (b[K["fuel0"]] + b[K["fuel_transport"]] + b[K["fuel_sources"]] + b[K["fuel_injected"]]
            - b[K["fuel_burned"]] - b[K["fuel_dissipated"]] - b[K["fuel_removed"]] + b[K["fuel_rounding"]])
    heat = (b[K["heat0"]] + b[K["heat_transport"]] + b[K["heat_sources"]] + b[K["heat_released"]]
            - b[K["heat_radiated"]] - b[K["heat_entrained"]] - b[K["heat_removed"]] + b[K["heat_rounding"]])
    scale_f = max(abs(b[K["fuel"]]), b[K["fuel_injected"]], 1e-30)
    scale_h = max(abs(b[K["heat"]]), b[K["heat_released"]], 1e-30)
    assert abs(fuel - b[K["fuel"]]) <= rtol * scale_f, (fuel, b[K["fuel"]])
    assert abs(heat - b[K["heat"]]) <= rtol * scale_h, (heat, b[K["heat"]])


def test_budget_terms_telescope_and_match_the_state():
    o, dt = _scene_oracle()
    o.step(dt, 15)                     # a burning state
    o.budget_enable()
    b0 = o.budget()
    S = o.read(O.S)
    V = float(np.float32(0.01)) ** 3
    assert np.isclose(b0[K["fuel0"]], RHO0 * V * S[..., 1].sum(), rtol=1e-13, atol=0)
    assert np.isclose(b0[K["heat0"]], RHO0 * CP * V * 1000 * S[..., 0].sum(), rtol=1e-13, atol=0)
    o.step(dt, 10)
    b = o.budget()
    assert b[K["steps"]] == 10 and b[K["bad_cells"]] == 0
    _check_identities(b)
    S = o.read(O.S)
    assert np.isclose(b[K["fuel"]], RHO0 * V * S[..., 1].sum(), rtol=1e-13, atol=0)
    assert b[K["fuel_burned"]] > 0 and b[K["heat_released"]] > 0 and b[K["heat_radiated"]] >= 0
    # the spot centres saturate at a fuel fraction of 1 within a few steps: injected < expected
    assert b[K["fuel_injected"]] < b[K["fuel_expected"]]
    # heat released per kg burned: cp heatOut, less where the 2100 K cap binds
    ratio = b[K["heat_released"]] / (CP * 7700.0 * b[K["fuel_burned"]])
    assert 0.5 < ratio <= 1 + 1e-6, ratio


def test_injected_fuel_is_the_sources_mass_flux():
    """From rest, before any cell saturates: the injected mass is the sources' normalisation,
    flux dt per spot and F dx^2 dt per floor cell, to f32 rounding."""
    o, dt = _scene_oracle()
    o.budget_enable()
    o.step(dt, 2)
    b = o.budget()
    _check_identities(b)
    assert o.read(O.S)[..., 1].max() < 0.9
    assert abs(b[K["fuel_injected"]] / b[K["fuel_expected"]] - 1) < 1e-5, (b[K["fuel_injected"]], b[K["fuel_expected"]])


def test_heat_per_burned_fuel_is_cp_heatout_without_the_cap():
    """A weak, cool scene (one spot, fuel fraction far from stoichiometric): no cell reaches 2100 K."""
    o, dt = _scene_oracle(precision=O.FP64, face=False, params={"heatOut": 600.0})
    o.step(dt, 8)
    o.budget_enable()
    o.step(dt, 4)
    b = o.budget()
    _check_identities(b)
    assert b[K["fuel_burned"]] > 0
    assert abs(b[K["heat_released"]] / (CP * 600.0 * b[K["fuel_burned"]]) - 1) < 1e-9


def test_fp16_storage_rounding_is_recorded():
    o, dt = _scene_oracle(precision=O.FP16)
    o.step(dt, 10)
    o.budget_enable()
    o.step(dt, 5)
    b = o.budget()
    _check_identities(b)
    assert b[K["fuel_rounding"]] != 0 and abs(b[K["fuel_rounding"]]) < 1e-2 * b[K["fuel"]]
