"""CPU tests that pin the oracle's step kernels against closed forms, invariants and brute force.

Each kernel of src/sim.js that the oracle restates is run on analytic inputs whose result the
mathematics fixes, so that a plausible restatement error (a dropped term, a wrong sign or index,
an Euler step where the reference takes a midpoint, a missing edge rule) fails a test:

  * adv1 (src/sim.js:123-157): translation of a linear field by a constant velocity is exact
    under trilinear sampling, clamped to the edge texel; the RK2 midpoint backtrace on a
    velocity that varies linearly; outsideFade returns ambient (zero scalars, the wind profile)
    for traces leaving the box;
  * adv2 (src/sim.js:159-197): the MacCormack correction vanishes on linear fields; the limiter
    keeps every value inside the 8 texels around the backtrace; a whole 4x4x4 MacCormack step on
    random fields against an independent numpy evaluation;
  * curl, div, project (src/sim.js:200-211, 399-413, 526-550): exact central differences of
    linear and trigonometric fields, with the clamped edge rows, the floor and the Dirichlet
    ghosts worked by hand; the solid, floor, speed and NaN rules of the projection;
  * geo (src/sim.js:498-523, src/shaders/lib.js:58-62): the campfire's capsule distances, pyrolysis
    band, coal bed and solid mask against a numpy evaluation;
  * forces (src/sim.js:214-396): one-cell closed forms of each sub-model with the others gated off
    through the parameters (burning, soot formation and oxidation, radiative and entrainment
    cooling, buoyancy, viscosity, the laminar gate, confinement on a known vortex, wind
    relaxation, velocity dissipation, the speed limit);
  * the oracle's f32 exp / log / pow / sin / cos against libm within stated ulp bounds.

The FP64 oracle (double storage and math) is compared at 1e-12; the FP32 oracle at f32 rounding.
"""
import numpy as np
import pytest

import fsoracle as O

NOLOGS = np.zeros(8, np.float32)


def f32(v):
    """A value as the f32 uniform block holds it (dt, dx and every parameter are f32 there)."""
    return float(np.float32(v))


def _sim(res, dx=0.01, precision=O.FP64, **params):
    size = tuple(n * dx for n in res)
    o = O.OracleSim(res=res, size=size, logs=NOLOGS, n_logs=0, precision=precision, noise_n=16)
    if params:
        o.set_params(**params)
    return o


def _grid(res):
    nx, ny, nz = res
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return i.astype(np.float64), j.astype(np.float64), k.astype(np.float64)


def _smoothstep(e0, e1, x):
    t = np.clip((x - e0) / (e1 - e0), 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t)


def _fade(bx, by, bz, res):
    """outsideFade (src/sim.js:137-140), positions in cell units."""
    nx, ny, nz = res
    lo = np.minimum(bx, bz)
    hi = np.maximum(bx - nx, np.maximum(by - ny, bz - nz))
    return np.clip(1.0 - np.maximum(-lo, hi), 0.0, 1.0)


def _trilinear(field, pos):
    """Independent clamp-to-edge trilinear sampler: field (nz, ny, nx, C), pos (..., 3) in cell units
    (texel centres at i + 0.5)."""
    nz, ny, nx = field.shape[:3]
    t = pos - 0.5
    i0 = np.floor(t).astype(int)
    fr = t - i0
    out = 0.0
    for dz in (0, 1):
        for dy in (0, 1):
            for dxx in (0, 1):
                w = ((fr[..., 0] if dxx else 1 - fr[..., 0]) * (fr[..., 1] if dy else 1 - fr[..., 1])
                     * (fr[..., 2] if dz else 1 - fr[..., 2]))
                ii = np.clip(i0[..., 0] + dxx, 0, nx - 1)
                jj = np.clip(i0[..., 1] + dy, 0, ny - 1)
                kk = np.clip(i0[..., 2] + dz, 0, nz - 1)
                out = out + w[..., None] * field[kk, jj, ii]
    return out


def _centres(res):
    i, j, k = _grid(res)
    return np.stack([i + 0.5, j + 0.5, k + 0.5], -1)


# ============================================================ adv1
@pytest.mark.parametrize("precision,tol", [(O.FP64, 1e-12), (O.FP32, 2e-6)], ids=["fp64", "fp32"])
def test_adv1_translates_linear_fields_exactly_with_edge_clamp_fade_and_wind(precision, tol):
    res = (16, 12, 10)
    dx, dt = 0.01, 0.012
    u, w = 1.3, -1.6                         # shifts of 1.56 cells in +x and 1.92 cells in -z
    wind = (0.7, -0.2)
    o = _sim(res, dx, precision, windX=wind[0], windZ=wind[1])
    nx, ny, nz = res
    i, j, k = _grid(res)
    vel = np.zeros((nz, ny, nx, 4))
    vel[..., 0], vel[..., 2] = u, w
    cS = np.array([[0.3, 0.02, -0.01, 0.005], [0.1, -0.004, 0.03, 0.002], [0.7, 0.01, 0.02, -0.03],
                   [0.05, 0.003, 0.001, 0.004]])           # per channel: a, kx, ky, kz
    cS2 = cS[::-1] * 0.5
    lin = lambda c, x, y, z: c[0] + c[1] * x + c[2] * y + c[3] * z
    S = np.stack([lin(c, i, j, k) for c in cS], -1)
    S2 = np.stack([lin(c, i, j, k) for c in cS2], -1)
    o.write(O.VEL, vel)
    o.write(O.S, S)
    o.write(O.S2, S2)
    o.stage(O.ST_ADV1, dt)
    dt, dx, wind = f32(dt), f32(dx), (f32(wind[0]), f32(wind[1]))
    sx, sz = u * dt / dx, w * dt / dx
    tx = np.clip(i - sx, 0, nx - 1)          # texel coordinate of the trace, clamped to the edge texel
    tz = np.clip(k - sz, 0, nz - 1)
    bx, by, bz = i + 0.5 - sx, j + 0.5, k + 0.5 - sz
    f = _fade(bx, by, bz, res)
    assert f.min() < 0.5 and f.max() == 1.0 and np.any((f > 0) & (f < 1))
    expS = np.stack([lin(c, tx, j, tz) for c in cS], -1) * f[..., None]
    expS2 = np.stack([lin(c, tx, j, tz) for c in cS2], -1) * f[..., None]
    wprof = _smoothstep(0.03, 0.4, by * dx)
    expV = np.zeros_like(vel)
    expV[..., 0] = wind[0] * wprof * (1 - f) + u * f
    expV[..., 2] = wind[1] * wprof * (1 - f) + w * f
    for fid, exp in ((O.SX, expS), (O.S2X, expS2), (O.VELX, expV)):
        got = o.read(fid)
        np.testing.assert_allclose(got, exp, rtol=tol, atol=tol, err_msg=str(fid))
    # a trace leaving the box by a full cell is ambient: zero scalars, pure wind
    out = f == 0
    assert np.any(out) and np.all(o.read(O.SX)[out] == 0)


@pytest.mark.parametrize("precision,tol", [(O.FP64, 1e-12), (O.FP32, 2e-6)], ids=["fp64", "fp32"])
def test_adv1_backtrace_is_the_rk2_midpoint(precision, tol):
    """u(x) = alpha + beta x (texel units): the midpoint rule c - dt v(c - dt v(c)/2dx)/dx, exact
    under trilinear sampling of linear fields (an Euler trace differs by beta dt^2 v / 2dx^2)."""
    res = (20, 8, 6)
    dx, dt = 0.01, 0.005
    alpha, beta = 0.5, 0.08
    o = _sim(res, dx, precision)
    nx, ny, nz = res
    i, j, k = _grid(res)
    vel = np.zeros((nz, ny, nx, 4))
    vel[..., 0] = alpha + beta * i
    S = np.zeros((nz, ny, nx, 4))
    S[..., 1] = 2.0 + 0.25 * i
    o.write(O.VEL, vel)
    o.write(O.S, S)
    o.write(O.S2, np.zeros_like(S))
    o.stage(O.ST_ADV1, dt)
    r = f32(dt) / f32(dx)
    tm = i - 0.5 * r * (alpha + beta * i)            # midpoint, texel coordinate
    tb = i - r * (alpha + beta * tm)                 # trace, texel coordinate
    inner = (i >= 2) & (i <= nx - 2)
    got = o.read(O.SX)[..., 1]
    np.testing.assert_allclose(got[inner], (2.0 + 0.25 * tb)[inner], rtol=tol, atol=tol)
    np.testing.assert_allclose(o.read(O.VELX)[..., 0][inner], (alpha + beta * tb)[inner], rtol=tol, atol=tol)
    euler = 2.0 + 0.25 * (i - r * (alpha + beta * i))
    assert np.max(np.abs(got[inner] - euler[inner])) > 1e-4


# ============================================================ adv2
def _adv_numpy(vel, S, dt, dx, res, wind=(0.0, 0.0)):
    """Both MacCormack passes (src/sim.js:123-197), written independently in float64 numpy."""
    c = _centres(res)
    v = vel[..., :3]

    def sample_v(p):
        return _trilinear(v, p)

    v0 = sample_v(c)
    mid = c - 0.5 * dt * v0 / dx
    b = c - dt * sample_v(mid) / dx
    f = _fade(b[..., 0], b[..., 1], b[..., 2], res)
    wprof = _smoothstep(0.03, 0.4, b[..., 1] * dx)
    windv = np.stack([wind[0] * wprof, 0 * wprof, wind[1] * wprof], -1)
    hat_v = windv * (1 - f[..., None]) + _trilinear(v, b) * f[..., None]
    hat_s = _trilinear(S, b) * f[..., None]
    fw = c + dt * v / dx
    vN = hat_v + 0.5 * (v - _trilinear(hat_v, fw))
    sN = hat_s + 0.5 * (S - _trilinear(hat_s, fw))
    bb = b - 0.5
    b0 = np.floor(bb).astype(int)
    nx, ny, nz = res
    vals_v, vals_s = [], []
    for kk in range(8):
        q = b0 + np.array([kk & 1, (kk >> 1) & 1, (kk >> 2) & 1])
        qi = np.clip(q[..., 0], 0, nx - 1)
        qj = np.clip(q[..., 1], 0, ny - 1)
        qk = np.clip(q[..., 2], 0, nz - 1)
        vals_v.append(v[qk, qj, qi])
        vals_s.append(S[qk, qj, qi])
    vmn, vmx = np.min(vals_v, 0), np.max(vals_v, 0)
    smn, smx = np.min(vals_s, 0), np.max(vals_s, 0)
    revert = _fade(b[..., 0], b[..., 1], b[..., 2], res) < 1.0
    vN = np.where(revert[..., None], hat_v, vN)
    sN = np.where(revert[..., None], hat_s, sN)
    vN = np.clip(vN, np.minimum(vmn, hat_v), np.maximum(vmx, hat_v))
    sN = np.clip(sN, np.minimum(smn, hat_s), np.maximum(smx, hat_s))
    return hat_v, hat_s, vN, np.maximum(sN, 0.0), (smn, smx), revert


def test_adv2_correction_vanishes_on_linear_fields():
    res = (14, 10, 12)
    dx, dt = 0.01, 0.003
    o = _sim(res, dx)
    nx, ny, nz = res
    i, j, k = _grid(res)
    vel = np.zeros((nz, ny, nx, 4))
    vel[..., 0], vel[..., 1], vel[..., 2] = 0.9, 0.6, -0.4
    S = np.stack([1.0 + 0.05 * i + 0.02 * j - 0.03 * k, 0.5 + 0.01 * k, 2 + 0 * i, 0.3 + 0.01 * j], -1)
    o.write(O.VEL, vel)
    o.write(O.S, S)
    o.write(O.S2, S * 0.5)
    o.stage(O.ST_ADV1, dt)
    o.stage(O.ST_ADV2, dt)
    hat, got = o.read(O.SX), o.read(O.SY)
    # interior cells whose forward probe and trace stay between texel centres
    inner = (i >= 2) & (i <= nx - 3) & (j >= 2) & (j <= ny - 3) & (k >= 2) & (k <= nz - 3)
    np.testing.assert_allclose(got[inner], hat[inner], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(o.read(O.VELY)[inner][:, :3], vel[inner][:, :3], rtol=1e-12)
    sh = np.array([0.9, 0.6, -0.4]) * f32(dt) / f32(dx)
    exact = 1.0 + 0.05 * (i - sh[0]) + 0.02 * (j - sh[1]) - 0.03 * (k - sh[2])
    np.testing.assert_allclose(got[..., 0][inner], exact[inner], rtol=1e-12)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_adv2_full_maccormack_step_on_4x4x4_matches_numpy(seed):
    """Random fields, random velocities carrying traces across every face: both passes of the
    FP64 oracle equal the independent numpy evaluation in every cell."""
    rng = np.random.default_rng(seed)
    res = (4, 4, 4)
    dx, dt = 0.01, 0.01
    wind = (0.3, -0.5)
    o = _sim(res, dx, windX=wind[0], windZ=wind[1])
    vel = np.zeros((4, 4, 4, 4))
    vel[..., :3] = rng.uniform(-0.9, 0.9, (4, 4, 4, 3))
    S = rng.uniform(0, 1, (4, 4, 4, 4))
    S2 = rng.uniform(0, 1, (4, 4, 4, 4))
    o.write(O.VEL, vel)
    o.write(O.S, S)
    o.write(O.S2, S2)
    o.stage(O.ST_ADV1, dt)
    o.stage(O.ST_ADV2, dt)
    dt, dx, wind = f32(dt), f32(dx), (f32(wind[0]), f32(wind[1]))
    hv, hs, vN, sN, _, revert = _adv_numpy(vel, S, dt, dx, res, wind)
    _, hs2, _, s2N, _, _ = _adv_numpy(vel, S2, dt, dx, res, wind)
    assert 0 < revert.sum() < revert.size
    np.testing.assert_allclose(o.read(O.VELX)[..., :3], hv, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(o.read(O.SX), hs, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(o.read(O.S2X), hs2, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(o.read(O.VELY)[..., :3], vN, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(o.read(O.SY), sN, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(o.read(O.S2Y), s2N, rtol=1e-12, atol=1e-12)
    assert np.all(o.read(O.VELY)[..., 3] == 0)


def test_adv2_limiter_keeps_values_inside_the_trace_box():
    """FP32 oracle, rough random scalars on 24x20x16 under a constant velocity: every corrected
    value lies within min/max of the 8 source texels at floor(trace - 1/2) (Selle et al. 2008)."""
    rng = np.random.default_rng(5)
    res = (24, 20, 16)
    dx, dt = 0.01, 0.004
    o = _sim(res, dx, O.FP32)
    nz, ny, nx = res[2], res[1], res[0]
    vel = np.zeros((nz, ny, nx, 4))
    vel[..., 0], vel[..., 1], vel[..., 2] = 1.1, -0.7, 0.45
    S = rng.uniform(0, 1, (nz, ny, nx, 4)).astype(np.float32).astype(np.float64)
    o.write(O.VEL, vel)
    o.write(O.S, S)
    o.write(O.S2, S)
    o.stage(O.ST_ADV1, dt)
    o.stage(O.ST_ADV2, dt)
    got = o.read(O.SY)
    _, _, _, _, (smn, smx), _ = _adv_numpy(vel, S, f32(dt), f32(dx), res)
    assert np.all(got >= smn - 1e-6) and np.all(got <= smx + 1e-6)
    # the correction is active: the result is not the plain semi-Lagrangian value
    assert np.mean(np.abs(got - o.read(O.SX))) > 1e-3


# ============================================================ curl, div
def _one_sided_span(n, idx):
    """Index span of the clamped central difference: (min(i+1, n-1) - max(i-1, 0))."""
    return np.minimum(idx + 1, n - 1) - np.maximum(idx - 1, 0)


def test_curl_of_linear_field_with_clamped_edges():
    res = (9, 11, 7)
    o = _sim(res)
    nx, ny, nz = res
    i, j, k = _grid(res)
    M = np.array([[0.1, -0.3, 0.25], [0.4, 0.05, -0.2], [-0.15, 0.35, 0.02]])
    pos = np.stack([i, j, k], -1)
    vel = np.zeros((nz, ny, nx, 4))
    vel[..., :3] = pos @ M.T + np.array([0.5, -0.2, 0.1])
    o.write(O.VELY, vel)
    o.stage(O.ST_CURL)
    got = o.read(O.VORT)
    sx, sy, sz = _one_sided_span(nx, i), _one_sided_span(ny, j), _one_sided_span(nz, k)
    # dv_a/dx_b over the clamped span: M[a, b] * span_b
    d = lambda a, b, s: M[a, b] * s
    w = 0.5 * np.stack([d(2, 1, sy) - d(1, 2, sz), d(0, 2, sz) - d(2, 0, sx), d(1, 0, sx) - d(0, 1, sy)], -1)
    np.testing.assert_allclose(got[..., :3], w, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(got[..., 3], np.linalg.norm(w, axis=-1), rtol=1e-12)
    # interior: the exact curl (times dx), (M21 - M12, M02 - M20, M10 - M01)
    np.testing.assert_allclose(got[3, 4, 4, :3], [M[2, 1] - M[1, 2], M[0, 2] - M[2, 0], M[1, 0] - M[0, 1]], rtol=1e-12)


def test_curl_of_a_shear_wave():
    """v = (sin(a j), 0, 0): w.z = -(sin(a (j+1)) - sin(a (j-1)))/2 = -sin(a) cos(a j) inside."""
    res = (6, 24, 5)
    a = 0.37
    o = _sim(res)
    i, j, k = _grid(res)
    vel = np.zeros((5, 24, 6, 4))
    vel[..., 0] = np.sin(a * j)
    o.write(O.VELY, vel)
    o.stage(O.ST_CURL)
    got = o.read(O.VORT)
    inner = (j >= 1) & (j <= 22)
    np.testing.assert_allclose(got[..., 2][inner], (-np.sin(a) * np.cos(a * j))[inner], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(got[..., 0], 0, atol=0)
    np.testing.assert_allclose(got[..., 2][j == 0], -0.5 * np.sin(a), rtol=1e-12)   # one-sided at the floor


def test_divergence_of_linear_field_with_floor_and_edges():
    res = (8, 10, 6)
    dx = 0.02
    o = _sim(res, dx)
    nx, ny, nz = res
    i, j, k = _grid(res)
    M = np.array([[0.3, 0.1, -0.2], [0.05, -0.45, 0.15], [0.2, 0.1, 0.6]])
    c0 = np.array([0.1, 0.7, -0.3])
    vel = np.zeros((nz, ny, nx, 4))
    vel[..., :3] = np.stack([i, j, k], -1) @ M.T + c0
    vel[..., 3] = 0.25 + 0.01 * i                   # phi, the expansion target
    o.write(O.VELX, vel)
    o.stage(O.ST_DIV)
    got = o.read(O.B)
    sx, sz = _one_sided_span(nx, i), _one_sided_span(nz, k)
    dy = np.where(j == 0, vel[..., 1] + np.roll(vel[..., 1], -1, axis=1),       # floor: B = -v.y
                  M[1, 1] * _one_sided_span(ny, j))
    dx = f32(dx)
    exp = (M[0, 0] * sx + dy + M[2, 2] * sz) / (2 * dx) - vel[..., 3]
    np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-12)
    # interior: trace(M)/dx - phi
    assert abs(got[2, 4, 3] - ((M[0, 0] + M[1, 1] + M[2, 2]) / dx - vel[2, 4, 3, 3])) < 1e-12


# ============================================================ project
def test_projection_of_linear_pressure_with_ghost_rules():
    res = (7, 9, 6)
    dx = 0.01
    o = _sim(res, dx)
    nx, ny, nz = res
    i, j, k = _grid(res)
    g = np.array([0.004, -0.003, 0.002])
    p = 0.01 + g[0] * i + g[1] * j + g[2] * k
    vel = np.zeros((nz, ny, nx, 4))
    vel[..., 0], vel[..., 1], vel[..., 2], vel[..., 3] = 0.2, -0.5, 0.1, 7.0
    o.write(O.VELX, vel)
    o.write(O.P, p)
    geo = np.zeros((nz, ny, nx, 4))
    geo[3, 4, 2, 2] = 1.0                            # one solid cell
    o.write(O.GEO, geo)
    o.stage(O.ST_PROJECT)
    got = o.read(O.VEL)

    def grad(axis_idx, n, plus, minus, me):
        # pr(c +- e): -me outside x/z faces and the top; +me below the floor
        return plus - minus

    # x: ghosts -p at both faces
    pxp = np.where(i == nx - 1, -p, 0.01 + g[0] * (i + 1) + g[1] * j + g[2] * k)
    pxm = np.where(i == 0, -p, 0.01 + g[0] * (i - 1) + g[1] * j + g[2] * k)
    pyp = np.where(j == ny - 1, -p, 0.01 + g[0] * i + g[1] * (j + 1) + g[2] * k)
    pym = np.where(j == 0, p, 0.01 + g[0] * i + g[1] * (j - 1) + g[2] * k)
    pzp = np.where(k == nz - 1, -p, 0.01 + g[0] * i + g[1] * j + g[2] * (k + 1))
    pzm = np.where(k == 0, -p, 0.01 + g[0] * i + g[1] * j + g[2] * (k - 1))
    dx = f32(dx)
    G = np.stack([pxp - pxm, pyp - pym, pzp - pzm], -1) / (2 * dx)
    exp = vel[..., :3] - G
    exp[..., 1] = np.where(j == 0, np.maximum(exp[..., 1], 0.0), exp[..., 1])
    exp[3, 4, 2] = 0.0
    np.testing.assert_allclose(got[..., :3], exp, rtol=1e-12, atol=1e-14)
    assert np.all(got[..., 3] == 0)
    # interior cells: v - g/dx exactly
    np.testing.assert_allclose(got[2, 3, 3, :3], vel[2, 3, 3, :3] - g / dx, rtol=1e-12)


def test_projection_speed_limit_and_nan_reset():
    res = (6, 6, 6)
    o = _sim(res)
    vel = np.zeros((6, 6, 6, 4))
    vel[..., 0] = 30.0
    vel[..., 1] = 40.0
    vel[2, 2, 2, 0] = np.nan
    o.write(O.VELX, vel)
    o.write(O.P, np.zeros((6, 6, 6)))
    o.write(O.GEO, np.zeros((6, 6, 6, 4)))
    o.stage(O.ST_PROJECT)
    got = o.read(O.VEL)
    np.testing.assert_allclose(got[3, 3, 3, :3], [12.0, 16.0, 0.0], rtol=1e-12)   # |v| = 50 -> 20
    assert np.all(got[2, 2, 2] == 0)


# ============================================================ geo
def _sd_capsule(p, a, b, r):
    pa, ba = p - a, b - a
    h = np.clip((pa @ ba) / (ba @ ba), 0, 1)
    return np.linalg.norm(pa - ba * h[..., None], axis=-1) - r


def test_geometry_of_the_campfire_matches_numpy():
    res = (40, 60, 40)
    size = (1.2, 1.8, 1.2)
    dx = size[0] / res[0]
    logs, n = O.build_campfire(7)
    o = O.OracleSim(res=res, size=size, logs=logs, n_logs=n, precision=O.FP64, noise_n=16)
    prm = dict(zip(O.param_names(), o.get_params().astype(np.float64)))
    o.stage(O.ST_GEO)
    got = o.read(O.GEO)
    i, j, k = _grid(res)
    p = np.stack([-size[0] / 2 + (i + 0.5) * dx, (j + 0.5) * dx, -size[2] / 2 + (k + 0.5) * dx], -1)
    L = logs.reshape(-1, 8).astype(np.float64)
    rxz = np.hypot(p[..., 0], p[..., 2])
    dmin = np.full(p.shape[:3], 1e9)
    band = np.zeros(p.shape[:3])
    for q in range(n):
        A, B = L[q, :4], L[q, 4:]
        dd = _sd_capsule(p, A[:3], B[:3], A[3]) + A[3] * 0.3 * (1 - _smoothstep(0.0, 0.16, rxz)) * B[3]
        dmin = np.minimum(dmin, dd)
        bnd = 1 - _smoothstep(0.0, 1.6 * dx, np.abs(dd - 0.6 * dx))
        hot = B[3] * (1 - _smoothstep(prm["hotR0"], prm["hotR1"], rxz)) * (1 - _smoothstep(0.08, prm["hotY1"], p[..., 1]))
        band = np.maximum(band, bnd * hot)
    bedH = 0.05 * (1 - _smoothstep(0.0, prm["bedRadius"], rxz)) + 0.02
    bed = (1 - _smoothstep(0.0, 2 * dx, p[..., 1] - bedH)) * (1 - _smoothstep(prm["bedRadius"] * 0.6, prm["bedRadius"] * 1.15, rxz))
    solid = (dmin < -0.5 * dx).astype(np.float64)
    # the parameters are f32 in the uniform block: allow their rounding
    np.testing.assert_allclose(got[..., 3], dmin, rtol=0, atol=1e-6)
    np.testing.assert_allclose(got[..., 0], band, rtol=0, atol=1e-5)
    np.testing.assert_allclose(got[..., 1], bed, rtol=0, atol=1e-5)
    assert np.array_equal(got[..., 2], solid)
    assert band.max() > 0.9 and solid.sum() > 50 and bed.max() > 0.9


# ============================================================ forces, one cell at a time
OFF = dict(buoyancy=0.0, vortEps=0.0, turbAmp=0.0, radCool=0.0, coolRate=0.0, sootForm=0.0, sootOx=0.0,
           sootDissip=0.0, smokeDissip=0.0, vitDecay=0.0, fuelDissip=0.0, velDissip=0.0, sponge=0.0, windX=0.0,
           windZ=0.0, gust=0.0, visc=0.0, lamAge=0.0, smolder=0.0, kMix=0.0)
FRES = (8, 24, 8)       # dx = 1 cm; cell (4, 12, 4) sits at y = 0.125 m, far from the top fade
CELL = (4, 12, 4)


def _forces(S, S2, vel=(0, 0, 0), dt=0.01, vort=None, precision=O.FP64, **params):
    """One forces pass on a uniform state (no sources: geo zero).  Returns the centre cell's vel, S,
    S2 and the parameter block as the kernel sees it (f32 values, plus the f32 dt)."""
    o = _sim(FRES, 0.01, precision, **{**OFF, **params})
    nx, ny, nz = FRES
    shape = (nz, ny, nx, 4)
    V = np.zeros(shape)
    V[..., :3] = vel
    o.write(O.VELY, V)
    o.write(O.SY, np.broadcast_to(np.asarray(S, np.float64), shape))
    o.write(O.S2Y, np.broadcast_to(np.asarray(S2, np.float64), shape))
    o.write(O.VORT, np.zeros(shape) if vort is None else vort)
    o.write(O.GEO, np.zeros(shape))
    o.stage(O.ST_FORCES, dt)
    x, y, z = CELL
    P = dict(zip(O.param_names(), o.get_params().astype(np.float64)))
    P["dt"] = f32(dt)                     # the grid entries (nx .. nLogs) of the block are the handle's
    return o.read(O.VELX)[z, y, x], o.read(O.S)[z, y, x], o.read(O.S2)[z, y, x], P


def _burn(theta, Y, X, P):
    """Mixing-limited burning (src/sim.js:300-309) with the block's kMix, rs, heatOut, ignT, expansion."""
    dt = P["dt"]
    T0 = 293 + 1000 * theta
    o = np.clip(1 - Y - X, 0, 1)
    ign = _smoothstep(P["ignT"] - 50, P["ignT"] + 50, T0)
    b = min(Y, o / P["rs"]) * (1 - np.exp(-P["kMix"] * ign * dt))
    dT = min(P["heatOut"] * b, max(2100 - T0, 0))
    q = dT / dt
    T = T0 + dT
    return dict(b=b, Y=Y - b, X=min(X + (1 + P["rs"]) * b, 1.0), T=T, q=q, phi=min(P["expansion"] * q / T, 60.0), o=o)


@pytest.mark.parametrize("theta,Y,X", [(0.6, 0.08, 0.1), (0.9, 0.3, 0.05), (0.2, 0.1, 0.0), (1.5, 0.05, 0.85)])
def test_forces_mixing_limited_burning_alone(theta, Y, X):
    v, s, s2, P = _forces((theta, Y, 0.0, 0.0), (X, 0.02, 0.0, 0.0), dt=0.01, kMix=30.0)
    r = _burn(theta, Y, X, P)
    assert abs(s[1] - r["Y"]) < 1e-12 and abs(s2[0] - r["X"]) < 1e-12
    assert abs(s[0] - (r["T"] - 293) * 1e-3) < 1e-12
    assert abs(s[3] - r["q"] * 1e-4) < 1e-9
    assert abs(v[3] - r["phi"]) < 1e-9
    # the age channel re-weighted by the new Z = fuel + products
    ageF = min(0.02 / max(Y + X, 1e-4) + P["dt"], 4.0)
    assert abs(s2[1] - ageF * np.clip(r["Y"] + r["X"], 0, 1)) < 1e-12
    if theta == 0.2:
        assert r["b"] == 0 and s[1] == Y           # below ignition: nothing burns
    else:
        assert r["b"] > 0


def test_forces_soot_formation_oxidation_and_burnout():
    theta, Y, X, soot, age, co = 0.95, 0.06, 0.2, 2.0, 0.01, 0.1
    v, s, s2, P = _forces((theta, Y, soot, 0.0), (X, age, 0.0, co), dt=0.004, kMix=30.0, sootForm=25.0, sootOx=30.0,
                          sootDissip=0.4)
    dt = P["dt"]
    r = _burn(theta, Y, X, P)
    ageF = min(age / max(Y + X, 1e-4) + dt, 4.0)
    T = r["T"]
    z = soot + P["sootForm"] * r["b"] * _smoothstep(0.015, 0.05, ageF) * _smoothstep(1000, 1400, T) * (1 - co)
    z *= np.exp(-P["sootOx"] * (0.25 + r["o"]) * _smoothstep(1100, 1400, T) * dt)
    z *= np.exp(-P["sootOx"] * 0.5 * r["o"] * _smoothstep(850, 1100, T) * (1 - _smoothstep(1100, 1300, T)) * dt)
    z *= np.exp(-P["sootDissip"] * dt)
    assert abs(s[2] - z) < 1e-12 * max(1, z)
    assert abs(z - soot) > 1e-3


@pytest.mark.parametrize("theta,soot", [(1.2, 3.0), (0.5, 10.0), (1.6, 400.0)])
def test_forces_radiative_cooling_alone(theta, soot):
    v, s, s2, P = _forces((theta, 0.0, soot, 0.0), (0.0, 0.0, 0.0, 0.0), dt=0.005, radCool=0.25)
    T = 293 + 1000 * theta
    th = T * 1e-3
    T -= min(P["radCool"] * 970 * soot * th * th * (th ** 4 - 0.00737), 6000) * P["dt"]
    assert abs(s[0] - (T - 293) * 1e-3) < 1e-12


def test_forces_entrainment_cooling_and_dissipation_rates():
    v, s, s2, P = _forces((0.8, 0.05, 0.0, 0.0), (0.3, 0.0, 0.4, 0.0), dt=0.01, coolRate=3.0, fuelDissip=0.3,
                          vitDecay=1.5, smokeDissip=0.3)
    dt = P["dt"]
    T = 293 + 800 * np.exp(-P["coolRate"] * 1.0 * dt)          # nothing burns: notBurning = 1
    assert abs(s[0] - (T - 293) * 1e-3) < 1e-12
    assert abs(s[1] - 0.05 * np.exp(-P["fuelDissip"] * dt)) < 1e-12
    assert abs(s2[0] - 0.3 * np.exp(-P["vitDecay"] * dt)) < 1e-12
    assert abs(s2[2] - 0.4 * np.exp(-P["smokeDissip"] * dt)) < 1e-12


def test_forces_buoyancy_and_velocity_dissipation():
    v, s, s2, P = _forces((0.5, 0.0, 0.0, 0.0), (0, 0, 0, 0), vel=(0.2, 0.3, -0.1), dt=0.01, buoyancy=26.8,
                          velDissip=0.05)
    dt = P["dt"]
    e = np.exp(-P["velDissip"] * dt)
    # the wind relaxation acts with zero wind too: it pulls v.x and v.z toward 0 above the floor
    py = (CELL[1] + 0.5) * f32(0.01)
    wk = _smoothstep(0.05, 0.5, py)
    ex = 0.2 + (0.0 - 0.2) * dt * 0.35 * wk
    ez = -0.1 + (0.0 + 0.1) * dt * 0.35 * wk
    np.testing.assert_allclose(v[:3], np.array([ex, 0.3 + P["buoyancy"] * 0.5 * dt, ez]) * e, rtol=1e-12)


def test_forces_wind_relaxation_with_gust():
    v, _, _, P = _forces((0.0, 0, 0, 0), (0, 0, 0, 0), vel=(0.1, 0.0, 0.4), dt=0.01, windX=2.0, windZ=-1.0, gust=0.5)
    dt = P["dt"]
    py = (CELL[1] + 0.5) * f32(0.01)
    wk = _smoothstep(0.05, 0.5, py) * (1 + P["gust"])
    ex = 0.1 + (P["windX"] * wk - 0.1) * dt * 0.35 * wk
    ez = 0.4 + (P["windZ"] * wk - 0.4) * dt * 0.35 * wk
    np.testing.assert_allclose(v[[0, 2]], [ex, ez], rtol=1e-12)


def test_forces_speed_limit():
    v, _, _, P = _forces((0.0, 0, 0, 0), (0, 0, 0, 0), vel=(9.0, 12.0, 20.0))
    r = 1 - P["dt"] * 0.35 * _smoothstep(0.05, 0.5, (CELL[1] + 0.5) * f32(0.01))   # zero-wind relaxation first
    w = np.array([9.0 * r, 12.0, 20.0 * r])
    np.testing.assert_allclose(v[:3], w * 15 / np.linalg.norm(w), rtol=1e-12)


def test_forces_hot_gas_viscosity_on_a_quadratic_profile():
    """v.x = a x^2 (cell units): the 7-point Laplacian is 2a inside; nu = visc 1.5e-5 (T/293)^1.7."""
    a = 0.004
    nx, ny, nz = FRES
    i, j, k = _grid(FRES)
    V = np.zeros((nz, ny, nx, 4))
    V[..., 0] = a * i * i
    theta = 0.9
    o = _sim(FRES, 0.01, O.FP64, **{**OFF, "visc": 2.0})
    o.write(O.VELY, V)
    o.write(O.SY, np.broadcast_to(np.array([theta, 0, 0, 0.0]), V.shape))
    o.write(O.S2Y, np.zeros_like(V))
    o.write(O.VORT, np.zeros_like(V))
    o.write(O.GEO, np.zeros_like(V))
    o.stage(O.ST_FORCES, 0.01)
    P = dict(zip(O.param_names(), o.get_params().astype(np.float64)))
    dt, dx = f32(0.01), f32(0.01)
    got = o.read(O.VELX)[4, 12, 4, 0]
    T = 293 + 1000 * theta
    nu = P["visc"] * 1.5e-5 * (T / 293) ** 1.7
    r = 1 - dt * 0.35 * _smoothstep(0.05, 0.5, (CELL[1] + 0.5) * dx)        # zero-wind relaxation after it
    assert abs(got - (a * 16 + min(nu * dt / dx ** 2, 0.15) * 2 * a) * r) < 1e-12
    assert got != a * 16


def _vortex_state():
    """w = (0, 0, wz) everywhere with |w| stored as 0.2 + 0.05 x: grad |w| = (0.1, 0, 0) (undivided,
    over two cells), so cross(g/|g|, w) = (0, -wz, 0)."""
    nx, ny, nz = FRES
    i, _, _ = _grid(FRES)
    vort = np.zeros((nz, ny, nx, 4))
    vort[..., 2] = 0.3
    vort[..., 3] = 0.2 + 0.05 * i
    return vort


@pytest.mark.parametrize("lamAge,age", [(0.0, 0.0), (0.2, 0.015), (0.2, 0.11), (0.2, 0.5)])
def test_forces_vorticity_confinement_on_a_known_vortex_and_the_laminar_gate(lamAge, age):
    theta = 1.0                                       # T = 1293 K: the temperature gate is 1
    Y = 0.01
    v, _, _, P = _forces((theta, Y, 0, 0), (0.0, age * Y, 0, 0), dt=0.01, vort=_vortex_state(), vortEps=4.0,
                         lamAge=lamAge)
    dt = P["dt"]
    ageF = min(age * Y / max(Y, 1e-4) + dt, 4.0)
    lam = 1.0 if lamAge == 0 else _smoothstep(0.5 * P["lamAge"], max(P["lamAge"], 1e-3), ageF)
    py = (CELL[1] + 0.5) * f32(0.01)
    gate = _smoothstep(450, 900, 1293.0) * _smoothstep(0.02, 0.08, py) * lam
    np.testing.assert_allclose(v[:3], [0.0, -P["vortEps"] * gate * 0.3 * dt, 0.0], rtol=1e-12, atol=1e-15)
    if lamAge and age < 0.05:
        assert v[1] == 0.0                            # young gas: no confinement
    if lamAge and 0.1 < ageF < 0.2:
        assert 0 < lam < 1


# ============================================================ f32 math
def _ulps(got, ref):
    """Error in units of the f32 spacing at the reference value."""
    ref = np.asarray(ref, np.float64)
    sp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    return np.abs(np.asarray(got, np.float64) - ref) / sp


def test_f32_exp_log_pow_sin_cos_within_ulp_bounds():
    rng = np.random.default_rng(11)
    x = rng.uniform(-87.0, 88.0, 200000).astype(np.float32)
    assert _ulps(O.fmath("exp", x), np.exp(x.astype(np.float64))).max() <= 2.0
    a = np.exp(rng.uniform(-60, 60, 200000)).astype(np.float32)
    lg = np.log(a.astype(np.float64))
    assert _ulps(O.fmath("log", a), lg).max() <= 2.0
    # pow = exp(b log a): the error scales with |b log a| (kernels use b log a within +-10)
    a = rng.uniform(0.05, 8.0, 200000).astype(np.float32)
    b = rng.uniform(-4.0, 4.0, 200000).astype(np.float32)
    y = b.astype(np.float64) * np.log(a.astype(np.float64))
    keep = np.abs(y) < 10
    ref = np.power(a.astype(np.float64), b.astype(np.float64))
    assert _ulps(O.fmath("pow", a, b)[keep], ref[keep]).max() <= 4.0 + 2.0 * np.abs(y[keep]).max()
    t = rng.uniform(-50, 50, 200000).astype(np.float32)
    for name, fn in (("sin", np.sin), ("cos", np.cos)):
        ref = fn(t.astype(np.float64))
        err = np.abs(O.fmath(name, t) - ref)
        assert np.all(err <= 4 * np.spacing(np.float32(1.0)) * np.maximum(np.abs(ref), 2 ** -4)), name
    # special values
    assert O.fmath("exp", [0.0])[0] == 1.0 and O.fmath("exp", [-200.0])[0] == 0.0
    assert np.isinf(O.fmath("exp", [100.0])[0]) and O.fmath("log", [1.0])[0] == 0.0
    assert O.fmath("pow", [1.0], [7.5])[0] == 1.0 and O.fmath("pow", [3.0], [0.0])[0] == 1.0
