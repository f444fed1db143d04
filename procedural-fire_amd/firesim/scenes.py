"""Synthetic spot-fed fire scenes for the measured workloads.

The reference feeds its fire from log capsules (src/scene.js:21-42 -> geometry kernel,
src/sim.js:498-523).  The workloads measured here feed it instead from pyrolysis spots
(fuel mass fluxes in kg/s on small spheres) and, for the adjoint scene, a floor fuel-flux
field; both enter the source kernel through fs_set_spots / fs_set_face_flux.

Positions and fluxes are drawn from the reference's own generator, the 32-bit LCG of
src/scene.js:13-16 (s = s * 1664525 + 1013904223 mod 2^32, u = s / 2^32), so a seed
names one layout on every host.

Parameter mapping of the scene descriptions:
  * "curl noise amplitude A m/s"      -> turbAmp = A (src/sim.js:368 scales it by dt*10/12)
  * "vorticity confinement eps"       -> vortEps = eps (src/sim.js:360)
  * "quiescent air at T K"            -> S.x = (T - 293) / 1000 everywhere (T = 293 + 1000 S.x,
                                         src/sim.js:300)
  * no log geometry: n_logs = 0 and bedRadius -> 1e-4 m so no coal bed is emitted
    (the bed mask of src/sim.js:512-520 does not depend on the logs).
"""
import math

import numpy as np

SPOT_TEMP = 0.75          # the reference's srcTemp default: pyrolysis gas enters at 1043 K
SPOT_RADIUS = 0.02        # m
SPOT_HEIGHT = 0.10        # m
DISC_DIAMETER = 0.30      # m

SCENES = {
    # 64x96x64 at 1 cm, dt 2 ms, 100 steps: 4 spots of 2 g/s on a 0.3 m disc, curl noise 0.2,
    # 4-level multigrid with 2 V-cycles, air at 300 K
    "spots4_64": dict(res=(64, 96, 64), dx=0.01, dt=0.002, steps=100, n_spots=4, flux=(2e-3, 2e-3),
                      layout="disc", params={"turbAmp": 0.2}, max_levels=4, vcycles=2, ambient_K=300.0),
    # 128x192x128 at 5 mm, dt 1 ms, 300 steps: 8 spots of U(1.5, 3) g/s, vorticity confinement 0.3
    "spots8_128": dict(res=(128, 192, 128), dx=0.005, dt=0.001, steps=300, n_spots=8, flux=(1.5e-3, 3e-3),
                       layout="disc", params={"vortEps": 0.3}, max_levels=0, vcycles=2, ambient_K=None),
    # 64x96x64 at 1 cm, dt 2 ms, 200 steps: 8 spots plus a 64x64 floor fuel flux
    "adjoint8_64": dict(res=(64, 96, 64), dx=0.01, dt=0.002, steps=200, n_spots=8, flux=(1.5e-3, 3e-3),
                        layout="disc", params={}, max_levels=0, vcycles=2, ambient_K=None, face=True),
    # 192x288x192 at 5.4 mm, dt 0.75 ms, 600 steps: 12 spots on 3 log-like segments, U(1, 3) g/s,
    # curl noise 0.3, 6-level multigrid
    "segments12_192": dict(res=(192, 288, 192), dx=0.0054, dt=0.00075, steps=600, n_spots=12, flux=(1e-3, 3e-3),
                           layout="segments", params={"turbAmp": 0.3}, max_levels=6, vcycles=2, ambient_K=None),
}

FACE_FLUX = 5e-3          # kg/m^2/s on the floor cells within the disc (adjoint scene)


def lcg(seed):
    """The reference's uniform generator (src/scene.js:13-16)."""
    s = seed & 0xFFFFFFFF

    def u():
        nonlocal s
        s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
        return s / 4294967296.0
    return u


def _disc(R, n, flux):
    rows = []
    for _ in range(n):
        r = 0.5 * DISC_DIAMETER * math.sqrt(R())
        a = 2.0 * math.pi * R()
        f = flux[0] + (flux[1] - flux[0]) * R()
        rows.append([r * math.cos(a), SPOT_HEIGHT, r * math.sin(a), SPOT_RADIUS, f, SPOT_TEMP])
    return rows


def _segments(R, n, flux, n_seg=3):
    """n spots evenly along n_seg chords 0.36 m long through the centre, crossed like a log pile
    (60 degrees apart plus a seeded twist), each chord at its own height."""
    rows = []
    per = n // n_seg
    twist = (R() - 0.5) * 0.4
    for k in range(n_seg):
        a = twist + k * math.pi / n_seg
        y = 0.06 + 0.03 * k
        for j in range(per):
            t = -0.18 + 0.36 * (j + 0.5) / per
            f = flux[0] + (flux[1] - flux[0]) * R()
            rows.append([t * math.cos(a), y, t * math.sin(a), SPOT_RADIUS, f, SPOT_TEMP])
    return rows


def scene_spots(name, seed=7):
    """Spot rows (x, y, z, radius, flux_kg_s, temp) of a scene, float32, shape (n, 6)."""
    sc = SCENES[name]
    R = lcg(seed)
    rows = _disc(R, sc["n_spots"], sc["flux"]) if sc["layout"] == "disc" else _segments(R, sc["n_spots"], sc["flux"])
    return np.asarray(rows, np.float32)


def scene_size(name):
    sc = SCENES[name]
    return tuple(float(np.float32(n * sc["dx"])) for n in sc["res"])


def scene_face_flux(name, res=None, seed=11):
    """Floor fuel flux (nz, nx) of a scene (on grid `res`, same dx), or None: within the
    disc, FACE_FLUX times a per-cell factor U(0.5, 1.5) from the reference LCG seeded with
    `seed` (row-major over the floor), zero outside.  The texture keeps neighbouring floor
    cells from holding identical fuel values: with a uniform flux every limiter bound on
    the floor is a tie between texels, where the step has no derivative."""
    sc = SCENES[name]
    if not sc.get("face"):
        return None
    nx, _, nz = res or sc["res"]
    dx = sc["dx"]
    x = (np.arange(nx) + 0.5) * dx - 0.5 * nx * dx
    z = (np.arange(nz) + 0.5) * dx - 0.5 * nz * dx
    r = np.hypot(x[None, :], z[:, None])
    R = lcg(seed)
    u = np.array([R() for _ in range(nx * nz)], np.float64).reshape(nz, nx)
    return np.where(r < 0.5 * DISC_DIAMETER, FACE_FLUX * (0.5 + u), 0.0).astype(np.float32)


def scene_params(name):
    """Parameter overrides of a scene (no coal bed, plus the scene's own settings)."""
    p = {"bedRadius": 1e-4}
    p.update(SCENES[name]["params"])
    return p


def apply_ambient(sim, name):
    """Writes the scene's ambient temperature into S.x (after create/reset)."""
    K = SCENES[name].get("ambient_K")
    if K is None:
        return
    S = sim.read("S")
    S[..., 0] = (K - 293.0) / 1000.0
    sim.write("S", S)


def make_scene(name, seed=7, res=None, **kw):
    """A FireSim holding the scene: no logs, the scene's spots, face flux, parameters and ambient.
    `res` overrides the grid (same dx); keyword arguments go to FireSim (precision, solve, ...)."""
    from . import FireSim
    sc = SCENES[name]
    res = tuple(res or sc["res"])
    size = tuple(float(np.float32(n * sc["dx"])) for n in res)
    kw.setdefault("vcycles", sc["vcycles"])
    kw.setdefault("max_levels", sc["max_levels"])
    sim = FireSim(res=res, size=size, logs=np.zeros(0, np.float32), **kw)
    sim.set_params(scene_params(name))
    sim.set_spots(scene_spots(name, seed))
    if sc.get("face") and res == tuple(sc["res"]):
        sim.set_face_flux(scene_face_flux(name))
    apply_ambient(sim, name)
    return sim


# ---- isolated pressure solve workload ------------------------------------------------
def normals(R):
    """Standard normals from the uniform generator by Box-Muller (pairs)."""
    while True:
        u1 = max(R(), 1e-300)
        u2 = R()
        r = math.sqrt(-2.0 * math.log(u1))
        yield r * math.cos(2.0 * math.pi * u2)
        yield r * math.sin(2.0 * math.pi * u2)


def blob_rhs(dims, n_blobs=64, sigma=(2.0, 8.0), noise=1e-2, seed=11):
    """Poisson right-hand side, shape (nz, ny, nx): n_blobs Gaussian blobs (centres uniform in
    cell space, widths U(sigma) cells, amplitudes N(0, 1)) plus white noise N(0, noise^2),
    then made zero-mean.  Blob parameters come from the reference LCG, the white noise from
    numpy's PCG64 seeded with `seed` (25 M draws)."""
    nx, ny, nz = dims
    R = lcg(seed)
    G = normals(R)
    b = np.zeros((nz, ny, nx), np.float64)
    for _ in range(n_blobs):
        c = (R() * nx, R() * ny, R() * nz)
        s = sigma[0] + (sigma[1] - sigma[0]) * R()
        a = next(G)
        lo = [max(0, int(c[k] - 4 * s)) for k in range(3)]
        hi = [min(d, int(c[k] + 4 * s) + 2) for k, d in enumerate(dims)]
        x = (np.arange(lo[0], hi[0]) + 0.5 - c[0]) ** 2
        y = (np.arange(lo[1], hi[1]) + 0.5 - c[1]) ** 2
        z = (np.arange(lo[2], hi[2]) + 0.5 - c[2]) ** 2
        g = np.exp(-(z[:, None, None] + y[None, :, None] + x[None, None, :]) / (2 * s * s))
        b[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] += a * g
    b += np.random.default_rng(seed).normal(0.0, noise, b.shape)
    b -= b.mean()
    return b
