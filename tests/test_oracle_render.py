"""Pins of the render-volume oracle (oracle/render_oracle.cpp) and of the blackbody LUT.

The oracle is checked against what the physics and geometry fix, not against its own formulas:
  * blackbody LUT: the library's LUT against an independent numpy restatement, luminance
    normalised to a 1500 K blackbody, white balance turning a blackbody at Twb neutral, Planck's
    law against its Wien and Rayleigh-Jeans limits, the CIE fit's peak;
  * volume march: an empty field leaves (L, Tr) = (0, 1); for a slab of constant temperature and
    soot the discrete march telescopes to the emission-absorption solution
    L = (e / sigma)(1 - Tr), and -ln Tr matches sigma times the slab's thickness (Beer-Lambert);
    single scattering of a distant light gives L = I p(c) / r^2 (sigma_s / sigma_t)(1 - Tr) with
    the Henyey-Greenstein phase p at the known angle; the surface depth cuts the march;
  * light extraction: a uniform hot field puts every band's light at the band's centre with
    the sample lattice's variance as its spread and the band's emitted power as its intensity;
    the temporal filter relaxes with its published time constants; a cold field gives no light;
  * prefilter: linear fields pass unchanged, a parabola loses k/3 (the discrete Laplacian),
    peaks are never raised.
"""
import math

import numpy as np
import pytest

import fsoracle as O
import firesim as F
from render_common import RP, axis_camera, bb_interp, make_uniforms

LUT = O.blackbody_lut()


# ---- blackbody LUT (src/blackbody.js) --------------------------------------------------
def test_library_lut_matches_restatement():
    a = F.blackbody_lut()
    np.testing.assert_allclose(a, LUT, rtol=2e-7, atol=0)


def test_lut_normalisation_and_white_balance():
    # Y (column 3) is the luminance relative to a 1500 K blackbody: exactly 1 there
    lut = F.blackbody_lut(n=2, tmin=1500.0, tmax=2500.0, twb=5000.0)
    assert lut[0, 0, 3] == pytest.approx(1.0, rel=1e-7)
    # a blackbody at the white-balance temperature comes out neutral (sRGB maps D65 to R=G=B)
    lut = F.blackbody_lut(n=4, tmin=2000.0, tmax=5000.0, twb=5000.0)
    r, g, b = lut[0, 3, :3]
    assert r == pytest.approx(g, rel=2e-4) and b == pytest.approx(g, rel=2e-4)
    # ... and a cooler one is red-shifted
    r, g, b = lut[0, 0, :3]
    assert r > g > b
    # without white balance the 5000 K blackbody is warm (its chromaticity is not D65's)
    lut0 = F.blackbody_lut(n=4, tmin=2000.0, tmax=5000.0, twb=0.0)
    assert lut0[0, 3, 0] > 1.02 * lut0[0, 3, 2]
    # soot (emissivity ~ lambda^-1.39) is bluer than the blackbody at the same temperature
    assert lut[1, 1, 2] / lut[1, 1, 0] > lut[0, 1, 2] / lut[0, 1, 0]
    # luminance rises with temperature
    assert np.all(np.diff(LUT[0, :, 3]) > 0)


def test_planck_limits_and_cie_peak():
    h, c, k = 6.62607015e-34, 2.99792458e8, 1.380649e-23
    lnm, T = 500.0, 1000.0        # h c / (lambda k T) = 28.8: Wien limit
    l = lnm * 1e-9
    wien = 2 * h * c * c / l ** 5 * math.exp(-h * c / (l * k * T))
    assert O.planck(lnm, T) == pytest.approx(wien, rel=1e-12)
    lnm, T = 1e6, 5000.0           # 1 mm at 5000 K: Rayleigh-Jeans limit 2 c k T / lambda^4
    l = lnm * 1e-9
    assert O.planck(lnm, T) == pytest.approx(2 * c * k * T / l ** 4, rel=2e-3)
    lam = np.arange(400.0, 700.0, 0.5)
    y = O.cie_fit(lam)[:, 1]
    assert 550 <= lam[y.argmax()] <= 560 and y.max() == pytest.approx(1.0, abs=0.03)


# ---- volume march -------------------------------------------------------------------------
RES, SIZE = (48, 64, 48), (0.48, 0.64, 0.48)
SLAB = (slice(18, 30), slice(5, 21), slice(18, 30))     # x, y, z cells; fade is 1 around it


def slab_fields(temp=0.0, soot=0.0, smoke=0.0, heat=0.0):
    nx, ny, nz = RES
    S = np.zeros((nz, ny, nx, 4), np.float32)
    S2 = np.zeros_like(S)
    S[..., 0] = temp
    xs, ys, zs = SLAB
    S[zs, ys, xs, 2] = soot
    S[zs, ys, xs, 3] = heat
    S2[zs, ys, xs, 2] = smoke
    return S, S2


CURL = np.zeros((8, 8, 8, 4), np.float32)
NO_LIGHTS = np.zeros((8, 8), np.float32)
CAM = axis_camera(0.0, 0.12, -2.0, fov_deg=0.5)       # through g = (24, 12, *), along +z


def march(S, S2, params, lights=NO_LIGHTS, depth=None, cam=CAM, size=32):
    p = dict(detailAmt=0.0, stepMul=0.1)
    p.update(params)
    U = make_uniforms(cam, RES, SIZE, p, width=size, height=size, vol_scale=0.5)
    return O.render_volume(U, S, S2, LUT, lights, CURL, depth)


def test_empty_field_is_transparent():
    S, S2 = slab_fields()
    for cubic in (0, 1):
        vol, aux = march(S, S2, dict(cubic=cubic))
        assert np.all(vol[..., :3] == 0) and np.all(vol[..., 3] == 1)
        assert np.all(aux[..., 0] == 0) and np.all(aux[..., 1] == np.float32(1e6))
    # rays that miss the box march nothing
    vol, aux = march(S, S2, {}, cam=axis_camera(5.0, 0.12, -2.0))
    assert np.all(vol[..., 3] == 1) and np.all(vol[..., :3] == 0)


@pytest.mark.parametrize("cubic", [0, 1])
def test_emission_absorption_slab(cubic):
    soot, temp = 5.0, 1.5
    S, S2 = slab_fields(temp=temp, soot=soot)
    prm = dict(cubic=cubic, sootAbs=4.0, sootEmit=10.0, emitK=4.0)
    vol, aux = march(S, S2, prm)
    c = vol[7:9, 7:9].reshape(-1, 4).astype(np.float64)      # rays within a cell of the axis
    Tr = c[:, 3]
    tau = prm["sootAbs"] * soot * 12 * 0.01                    # 12 cells of 1 cm
    np.testing.assert_allclose(-np.log(Tr), tau, rtol=1e-2)    # Beer-Lambert
    ratio = bb_interp(LUT, 293.0 + temp * 1000.0, 1) * prm["sootEmit"] * prm["emitK"] / prm["sootAbs"]
    for k in range(3):                                         # L = (e / sigma)(1 - Tr), binary16 output
        np.testing.assert_allclose(c[:, k], ratio[k] * (1 - Tr), rtol=2e-3)
    # a colder slab (below emitTempLo) absorbs the same and emits nothing
    S, S2 = slab_fields(temp=0.5, soot=soot)
    vol2, _ = march(S, S2, prm)
    assert np.all(vol2[..., :3] == 0)
    np.testing.assert_array_equal(vol2[..., 3], vol[..., 3])


@pytest.mark.parametrize("g", [0.0, 0.5])
def test_single_scattering_of_a_distant_light(g):
    S, S2 = slab_fields(smoke=1.0)
    lights = np.zeros((8, 8), np.float32)
    lights[0, :3] = (0.0, 0.12, -1002.0)                      # behind the camera, on the axis
    lights[0, 4:7] = (1e6, 2e6, 3e6)
    vol, _ = march(S, S2, dict(smokeScat=10.0, smokeG=g, lightK=0.6), lights=lights)
    c = vol[7:9, 7:9].reshape(-1, 4).astype(np.float64)
    Tr = c[:, 3]
    np.testing.assert_allclose(-np.log(Tr), 10.0 * 1.1 * 0.12, rtol=1e-2)
    phase = (1 - g * g) / (4 * math.pi * (1 + g * g + 2 * g) ** 1.5)   # cos = -1: back-scattering
    for k in range(3):
        want = lights[0, 4 + k] * phase / 1002.0 ** 2 * 0.6 / 1.1 * (1 - Tr)
        np.testing.assert_allclose(c[:, k], want, rtol=3e-3)


def test_surface_depth_cuts_the_march():
    S, S2 = slab_fields(soot=5.0)
    full, _ = march(S, S2, {})
    depth = np.array([[1e6, 2.0]], np.float32)                # right half: a surface at the slab's middle
    vol, aux = march(S, S2, {}, depth=depth)
    tau = 4.0 * 5.0 * 0.12
    left, right = vol[7:9, 7], vol[7:9, 8]
    np.testing.assert_allclose(-np.log(left[:, 3].astype(np.float64)), tau, rtol=1e-2)
    np.testing.assert_allclose(-np.log(right[:, 3].astype(np.float64)), tau / 2, rtol=2e-2)
    assert np.all(aux[:, :8, 1] == np.float32(1e6)) and np.all(aux[:, 8:, 1] == np.float32(2.0))
    np.testing.assert_array_equal(vol[:, :8], full[:, :8])


# ---- light extraction ------------------------------------------------------------------------
LRES, LSIZE = (32, 48, 32), (0.32, 0.48, 0.32)


def light_uniforms(light_dt=-1.0, **params):
    return make_uniforms(axis_camera(0.0, 0.2, -2.0), LRES, LSIZE, params, light_dt=light_dt)


def uniform_field(temp, soot):
    nx, ny, nz = LRES
    S = np.zeros((nz, ny, nx, 4), np.float32)
    S[..., 0] = temp
    S[..., 2] = soot
    return S


def test_lights_of_a_uniform_fire():
    U = light_uniforms()
    lstate = np.zeros((3, 8, 4), np.float32)
    L = O.render_lights(U, uniform_field(1.5, 2.0), LUT, lstate)
    dx = 0.01
    e = 2.0 * 10.0 * bb_interp(LUT, 1793.0, 1) * 4.0
    lum = np.array([0.2126, 0.7152, 0.0722])
    J = e * (32 * 6 * 32) * dx ** 3                               # the band's emitted power
    spread = dx * dx * (4 * (16 ** 2 - 1) / 12 * 2 + 4 * (3 ** 2 - 1) / 12)   # lattice variance, 3 axes
    for k in range(8):
        np.testing.assert_allclose(L[k, :3], [0.0, (6 * k + 3) * dx, 0.0], atol=2e-7)
        np.testing.assert_allclose(L[k, 4:7], J, rtol=1e-5)
        assert L[k, 3] == pytest.approx(J @ lum, rel=1e-5)
        assert L[k, 7] == pytest.approx(spread, rel=1e-4)


def test_lights_temporal_filter():
    lstate = np.zeros((3, 8, 4), np.float32)
    O.render_lights(light_uniforms(), uniform_field(1.5, 2.0), LUT, lstate)
    L = O.render_lights(light_uniforms(light_dt=0.05, flickerK=0.35), uniform_field(1.5, 4.0), LUT, lstate)
    J1 = 2.0 * 10.0 * bb_interp(LUT, 1793.0, 1) * 4.0 * (32 * 6 * 32) * 1e-6
    a, am = 1 - math.exp(-0.05 / 0.08), 1 - math.exp(-0.05 / 1.5)   # output 80 ms, running mean 1.5 s
    sc, sm = J1 * (1 + a), J1 * (1 + am)
    Jc = sc * ((1 + a) / (1 + am)) ** (0.35 - 1.0)
    np.testing.assert_allclose(L[0, 4:7], Jc, rtol=2e-5)
    np.testing.assert_allclose(lstate[2, 0, :3], sm, rtol=2e-5)
    assert lstate[2, 0, 3] == 1.0


def test_cold_field_gives_no_light():
    lstate = np.zeros((3, 8, 4), np.float32)
    L = O.render_lights(light_uniforms(), uniform_field(0.3, 5.0), LUT, lstate)   # 593 K < 700 K
    for k in range(8):
        y = np.float32(6 * k + 3) * np.float32(0.01)             # (y0 + bandH / 2) dx in f32
        np.testing.assert_array_equal(L[k], np.float32([0.0, y, 0.0, 0.0, 0.0, 0.0, 0.0, 0.04]))


# ---- prefilter ------------------------------------------------------------------------------
def test_sharpen_linear_and_parabola():
    nx, ny, nz = 16, 12, 10
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    S = np.stack([x + 2 * y + 3 * z, 5 * x - y, z + 0 * x, 7 + 0 * x], -1).astype(np.float32)
    out = O.render_sharpen(1.0, S)
    np.testing.assert_array_equal(out[1:-1, 1:-1, 1:-1], S[1:-1, 1:-1, 1:-1])   # linear: unchanged
    P = np.zeros_like(S)
    P[..., 0] = (x.astype(np.float32) / 4) ** 2
    for k in (0.5, 1.0):
        out = O.render_sharpen(k, P)
        want = (x / 4.0) ** 2 - k * (2 / 16) / 6                 # s + k (s - mean) = s - k h^2 s'' / 6
        np.testing.assert_allclose(out[1:-1, 1:-1, 2:-1, 0], want[1:-1, 1:-1, 2:-1], rtol=2.0 ** -11, atol=1e-7)   # binary16


def test_sharpen_never_raises_peaks_or_leaves_the_local_range():
    rng = np.random.default_rng(3)
    S = rng.random((10, 12, 14, 4)).astype(np.float32)
    out = O.render_sharpen(1.0, S)
    pad = np.pad(S, ((1, 1), (1, 1), (1, 1), (0, 0)), mode="edge")
    nb = [pad[1:-1, 1:-1, 2:], pad[1:-1, 1:-1, :-2], pad[1:-1, 2:, 1:-1], pad[1:-1, :-2, 1:-1], pad[2:, 1:-1, 1:-1],
          pad[:-2, 1:-1, 1:-1]]
    lo = np.minimum.reduce([S] + nb)
    hi = np.maximum.reduce([S] + nb)
    tol = 1e-3
    assert np.all(out >= lo - tol) and np.all(out <= hi + tol)
    peak = np.zeros((5, 5, 5, 4), np.float32)
    peak[2, 2, 2] = 10.0
    out = O.render_sharpen(1.0, peak)
    np.testing.assert_array_equal(out, peak)
    np.testing.assert_array_equal(O.render_sharpen(0.0, S), S.astype(np.float16).astype(np.float32))   # k = 0


def test_uniform_block_layout():
    assert len(RP) == 88 and RP["sootEdge"] == 87 and RP["lightDt"] == 83
    U = make_uniforms(CAM, RES, SIZE, dict(cubic=1), width=32, height=32, vol_scale=0.5)
    assert U[RP["vw"]] == 16 and U[RP["vh"]] == 16 and U[RP["dx"]] == np.float32(0.01)
    assert U[RP["ox"]] == np.float32(-0.24) and U[RP["cubic"]] == 1
