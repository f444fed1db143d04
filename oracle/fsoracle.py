"""TEST INFRASTRUCTURE ONLY — ctypes binding of the CPU oracle (oracle/_build/libfsoracle.so).

The oracle is a C++ restatement of the reference's FireSim step (src/sim.js:627-648) and
its inputs (src/noise.js, src/scene.js).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this module, and only as the checker or the timed
CPU baseline.  The product (procedural-fire_amd/) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libfsoracle.so")

FP32, FP16, BF16, FP64 = 0, 1, 2, 3
SOLVE_VCYCLES, SOLVE_REFINE64, SOLVE_TOL32 = 0, 1, 2
(VEL, S, S2, VORT, GEO, VELX, SX, S2X, VELY, SY, S2Y, P, B, NOISE, CURL) = range(15)
(ST_GEO, ST_ADV1, ST_ADV2, ST_CURL, ST_FORCES, ST_DIV, ST_MG, ST_PROJECT) = range(8)
NUM_PARAMS = 60


class FsConfig(ctypes.Structure):
    """Mirror of fs_config in include/firesim.h."""
    _fields_ = [
        ("res", ctypes.c_int * 3),
        ("size", ctypes.c_float * 3),
        ("vcycles", ctypes.c_int),
        ("max_levels", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("solve", ctypes.c_int),
        ("div_tol", ctypes.c_double),
        ("max_vcycles", ctypes.c_int),
        ("noise_n", ctypes.c_int),
        ("noise_seed", ctypes.c_int),
        ("device", ctypes.c_int),
        ("stream", ctypes.c_void_p),
        ("use_graph", ctypes.c_int),
        ("adaptive", ctypes.c_int),
        ("omega", ctypes.c_float),
        ("pre_sweeps", ctypes.c_int),
        ("post_sweeps", ctypes.c_int),
        ("direct_level", ctypes.c_int),
        ("first_cycles", ctypes.c_int),
        ("reserved", ctypes.c_int),
    ]


def make_config(res=(80, 120, 80), size=(1.2, 1.8, 1.2), vcycles=2, max_levels=0, precision=FP16,
                solve=SOLVE_VCYCLES, div_tol=1e-5, max_vcycles=8, noise_n=128, noise_seed=1, omega=0.0, pre_sweeps=0,
                post_sweeps=0, direct_level=0, first_cycles=0):
    c = FsConfig()
    c.res[:] = list(res)
    c.size[:] = list(size)
    c.vcycles = vcycles
    c.max_levels = max_levels
    c.precision = precision
    c.solve = solve
    c.div_tol = div_tol
    c.max_vcycles = max_vcycles
    c.noise_n = noise_n
    c.noise_seed = noise_seed
    c.omega = omega
    c.pre_sweeps = pre_sweeps
    c.post_sweeps = post_sweeps
    c.direct_level = direct_level
    c.first_cycles = first_cycles
    return c


class FsSpot(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_float * 3), ("radius", ctypes.c_float), ("flux", ctypes.c_float),
                ("temp", ctypes.c_float), ("norm", ctypes.c_float), ("pad", ctypes.c_float)]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "-j4"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, dp, fp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.POINTER(FsConfig), fp, ctypes.c_int]
        L.orc_destroy.argtypes = [vp]
        for n in ("orc_set_params", "orc_get_params"):
            getattr(L, n).argtypes = [vp, fp]
        L.orc_param_defaults.argtypes = [fp]
        L.orc_param_name.restype = ctypes.c_char_p
        L.orc_param_name.argtypes = [ctypes.c_int]
        L.orc_set_logs.argtypes = [vp, fp, ctypes.c_int]
        L.orc_reset.argtypes = [vp]
        L.orc_step.argtypes = [vp, ctypes.c_double, ctypes.c_int]
        L.orc_stage.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.orc_read_field.argtypes = [vp, ctypes.c_int, dp]
        L.orc_write_field.argtypes = [vp, ctypes.c_int, dp]
        L.orc_field_dims.argtypes = [vp, ctypes.c_int, ip, ip]
        L.orc_time.restype = ctypes.c_double
        L.orc_time.argtypes = [vp]
        L.orc_set_clock.argtypes = [vp, ctypes.c_double, ctypes.c_long]
        L.orc_frame.restype = ctypes.c_long
        L.orc_frame.argtypes = [vp]
        L.orc_mg_info.argtypes = [vp, ip]
        L.orc_coarse_iters.argtypes = [vp]
        L.orc_residual.restype = ctypes.c_double
        L.orc_residual.argtypes = [vp, ip]
        L.orc_build_campfire.argtypes = [ctypes.c_int, fp, ctypes.c_int]
        L.orc_bake_noise.argtypes = [ctypes.c_int, ctypes.c_int, fp, fp]
        L.orc_hash3f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, fp]
        L.orc_f16_round.argtypes = [fp, fp, ctypes.c_long]
        L.orc_bf16_round.argtypes = [fp, fp, ctypes.c_long]
        L.orc_mg_create.restype = vp
        L.orc_mg_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_mg_destroy.argtypes = [vp]
        L.orc_mg_set_omega.argtypes = [vp, ctypes.c_float]
        L.orc_mg_levels.argtypes = [vp, ip, ip]
        L.orc_mg_set.argtypes = [vp, dp, dp]
        L.orc_mg_get.argtypes = [vp, dp]
        L.orc_mg_vcycles.argtypes = [vp, ctypes.c_int, dp]
        L.orc_mg_op.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_mg_level_get.argtypes = [vp, ctypes.c_int, dp, dp]
        L.orc_mg_level_set.argtypes = [vp, ctypes.c_int, dp, dp]
        L.orc_mg_refine.argtypes = [vp, ctypes.c_double, ctypes.c_int, dp]
        L.orc_mg_get_x.argtypes = [vp, dp]
        L.orc_mg_residual.restype = ctypes.c_double
        L.orc_mg_residual.argtypes = [vp]
        L.orc_mjs_study.restype = ctypes.c_double
        L.orc_mjs_study.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, dp]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_spots.argtypes = [vp, ctypes.POINTER(FsSpot), ctypes.c_int, ctypes.POINTER(FsSpot)]
        L.orc_set_face_flux.argtypes = [vp, fp]
        L.orc_set_flux64.argtypes = [vp, dp, ctypes.c_int, dp]
        L.orc_sparks_create.restype = vp
        L.orc_sparks_create.argtypes = [vp, ctypes.c_int]
        L.orc_sparks_destroy.argtypes = [vp]
        L.orc_sparks_reset.argtypes = [vp]
        L.orc_sparks_set.argtypes = [vp, ctypes.c_char_p, ctypes.c_float]
        L.orc_sparks_burst.argtypes = [vp, ctypes.c_int, fp, ctypes.c_float, ctypes.c_float, ctypes.c_double,
                                       ctypes.c_float, ctypes.c_float, ctypes.c_int, fp]
        L.orc_sparks_step.argtypes = [vp, ctypes.c_double]
        L.orc_sparks_read.argtypes = [vp, fp]
        L.orc_sparks_write.argtypes = [vp, fp]
        L.orc_sparks_time.restype = ctypes.c_double
        L.orc_sparks_time.argtypes = [vp]
        _lib = L
    return _lib


def num_threads():
    return lib().orc_num_threads()


def set_threads(n):
    """Thread count of the oracle's OpenMP loops for the rest of the process."""
    lib().orc_set_threads.argtypes = [ctypes.c_int]
    lib().orc_set_threads(int(n))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def param_defaults():
    out = np.zeros(NUM_PARAMS, np.float32)
    lib().orc_param_defaults(_fp(out))
    return out


def param_names():
    return [lib().orc_param_name(i).decode() for i in range(NUM_PARAMS)]


def build_campfire(seed=7, max_logs=16):
    out = np.zeros(max_logs * 8, np.float32)
    n = lib().orc_build_campfire(seed, _fp(out), max_logs)
    return out, n


def bake_noise(n=128, seed=1):
    noise = np.zeros(n ** 3 * 4, np.float32)
    curl = np.zeros(n ** 3 * 4, np.float32)
    lib().orc_bake_noise(n, seed, _fp(noise), _fp(curl))
    return noise.reshape(n, n, n, 4), curl.reshape(n, n, n, 4)


def hash3f(x, y, z):
    out = np.zeros(3, np.float32)
    lib().orc_hash3f(int(x), int(y), int(z), _fp(out))
    return out


def f16_round(a):
    a = np.ascontiguousarray(a, np.float32)
    out = np.empty_like(a)
    lib().orc_f16_round(_fp(a), _fp(out), a.size)
    return out


def bf16_round(a):
    a = np.ascontiguousarray(a, np.float32)
    out = np.empty_like(a)
    lib().orc_bf16_round(_fp(a), _fp(out), a.size)
    return out


_FMATH = {"exp": 0, "log": 1, "pow": 2, "sin": 3, "cos": 4}


def fmath(name, a, b=None):
    """The oracle's f32 exp / log / pow / sin / cos (the textbook constructions the kernels share)."""
    L = lib()
    fp = ctypes.POINTER(ctypes.c_float)
    L.orc_fmath.argtypes = [ctypes.c_int, fp, fp, fp, ctypes.c_long]
    a = np.ascontiguousarray(a, np.float32).ravel()
    bb = np.ascontiguousarray(b if b is not None else np.zeros_like(a), np.float32).ravel()
    out = np.empty_like(a)
    L.orc_fmath(_FMATH[name], _fp(a), _fp(bb), _fp(out), a.size)
    return out


class OracleSim:
    """CPU FireSim restatement.  precision: FP32/FP16/BF16 (f32 math) or FP64."""

    def __init__(self, res=(64, 96, 64), size=None, logs=None, n_logs=None, precision=FP16, vcycles=2,
                 max_levels=0, solve=SOLVE_VCYCLES, div_tol=1e-5, max_vcycles=8, noise_n=128, noise_seed=1, omega=0.0,
                 pre_sweeps=0, post_sweeps=0, direct_level=0, first_cycles=0):
        if size is None:
            size = (res[0] * 0.015, res[1] * 0.015, res[2] * 0.015)
        if logs is None:
            logs, n = build_campfire(7)
            n_logs = 16 if n_logs is None else n_logs
        logs = np.ascontiguousarray(logs, np.float32)
        if n_logs is None:
            n_logs = logs.size // 8
        self.res = tuple(res)
        self.cfg = make_config(res, size, vcycles, max_levels, precision, solve, div_tol, max_vcycles, noise_n, noise_seed,
                               omega, pre_sweeps, post_sweeps, direct_level, first_cycles)
        self.h = lib().orc_create(ctypes.byref(self.cfg), _fp(logs), n_logs)
        if not self.h:
            raise RuntimeError("oracle create failed")

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, **kw):
        names = param_names()
        p = self.get_params()
        for k, v in kw.items():
            p[names.index(k)] = v
        lib().orc_set_params(self.h, _fp(p))

    def set_param_block(self, p):
        p = np.ascontiguousarray(p, np.float32)
        lib().orc_set_params(self.h, _fp(p))

    def set_logs(self, logs):
        """Rewrites the packed logs (8 floats per log) and invalidates the geometry (src/main.js:147-149)."""
        a = np.ascontiguousarray(logs, np.float32)
        lib().orc_set_logs(self.h, _fp(a), a.size // 8)

    def get_params(self):
        out = np.zeros(NUM_PARAMS, np.float32)
        lib().orc_get_params(self.h, _fp(out))
        return out

    def set_spots(self, spots):
        """rows of (x, y, z, radius, flux_kg_s, temp); returns the per-spot mass normalisation"""
        a = np.asarray(spots, np.float32).reshape(-1, 6)
        arr = (FsSpot * max(len(a), 1))()
        out = (FsSpot * max(len(a), 1))()
        for i, r in enumerate(a):
            arr[i].pos[:] = [float(r[0]), float(r[1]), float(r[2])]
            arr[i].radius, arr[i].flux, arr[i].temp = float(r[3]), float(r[4]), float(r[5])
        lib().orc_set_spots(self.h, arr, len(a), out)
        return np.array([out[i].norm for i in range(len(a))], np.float64)

    def set_face_flux(self, flux):
        if flux is None:
            lib().orc_set_face_flux(self.h, None)
            return
        nx, _, nz = self.res
        a = np.ascontiguousarray(flux, np.float32)
        assert a.shape == (nz, nx)
        lib().orc_set_face_flux(self.h, _fp(a))

    def set_flux64(self, spot=None, face=None):
        """Double-precision spot fluxes (kg/s, one per spot) and/or floor flux (nz, nx) in place of
        the float values given to set_spots / set_face_flux (None: back to those)."""
        sp = None if spot is None else np.ascontiguousarray(spot, np.float64)
        fc = None if face is None else np.ascontiguousarray(face, np.float64)
        if fc is not None:
            assert fc.shape == (self.res[2], self.res[0])
        lib().orc_set_flux64(self.h, None if sp is None else _dp(sp), 0 if sp is None else sp.size,
                             None if fc is None else _dp(fc))

    def budget_enable(self, on=True):
        lib().orc_budget_enable.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib().orc_budget_enable(self.h, 1 if on else 0)

    def budget(self):
        """The FS_BUDGET_* record (same layout as the library's fs_budget_read)."""
        lib().orc_budget_read.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        out = np.zeros(24, np.float64)
        lib().orc_budget_read(self.h, _dp(out))
        return out

    def reset(self):
        lib().orc_reset(self.h)

    def step(self, dt, n=1):
        lib().orc_step(self.h, float(dt), int(n))

    def stage(self, st, dt=0.0):
        lib().orc_stage(self.h, int(st), float(dt))

    def dims(self, field):
        d = (ctypes.c_int * 3)()
        ch = ctypes.c_int()
        lib().orc_field_dims(self.h, field, d, ctypes.byref(ch))
        return tuple(d), ch.value

    def read(self, field):
        (nx, ny, nz), ch = self.dims(field)
        out = np.zeros(nx * ny * nz * ch, np.float64)
        rc = lib().orc_read_field(self.h, field, _dp(out))
        if rc:
            raise ValueError("bad field")
        return out.reshape((nz, ny, nx, ch) if ch > 1 else (nz, ny, nx))

    def write(self, field, arr):
        a = np.ascontiguousarray(arr, np.float64).ravel()
        (nx, ny, nz), ch = self.dims(field)
        assert a.size == nx * ny * nz * ch
        rc = lib().orc_write_field(self.h, field, _dp(a))
        if rc:
            raise ValueError("bad field")

    @property
    def time(self):
        return lib().orc_time(self.h)

    @property
    def frame(self):
        return lib().orc_frame(self.h)

    def set_clock(self, time, frame):
        lib().orc_set_clock(self.h, float(time), int(frame))

    def mg_info(self):
        d = (ctypes.c_int * 48)()
        n = lib().orc_mg_info(self.h, d)
        return [tuple(d[i * 3:i * 3 + 3]) for i in range(n)], lib().orc_coarse_iters(self.h)

    def residual(self):
        u = ctypes.c_int()
        r = lib().orc_residual(self.h, ctypes.byref(u))
        return r, u.value


class OracleSparks:
    """CPU restatement of the spark particles (src/sparks.js) over an OracleSim's velocity and curl
    texture; dead particles take the spawn slots in index order, like the library."""

    def __init__(self, sim, count=8192):
        self.sim = sim
        self.count = int(count)
        self.h = lib().orc_sparks_create(sim.h, self.count)

    def __del__(self):
        try:
            if self.h:
                lib().orc_sparks_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def reset(self):
        lib().orc_sparks_reset(self.h)

    def set(self, **kw):
        for k, v in kw.items():
            if lib().orc_sparks_set(self.h, k.encode(), float(v)):
                raise ValueError(k)

    def burst(self, n=60, pos=(0.0, 0.1, 0.0), speed=3.0, radius=0.05, dur=0.12, fan=0.6, size=1.3, aim=None):
        p = np.asarray(pos, np.float32)
        a = np.asarray(aim if aim is not None else (0, 0, 0), np.float32)
        lib().orc_sparks_burst(self.h, int(n), _fp(p), float(speed), float(radius), float(dur), float(fan), float(size),
                               1 if aim is not None else 0, _fp(a))

    def step(self, dt):
        lib().orc_sparks_step(self.h, float(dt))

    def read(self):
        out = np.zeros((self.count, 16), np.float32)
        lib().orc_sparks_read(self.h, _fp(out))
        return out

    def write(self, arr):
        a = np.ascontiguousarray(arr, np.float32)
        lib().orc_sparks_write(self.h, _fp(a))

    @property
    def time(self):
        return lib().orc_sparks_time(self.h)


class OracleMG:
    """Stand-alone sim.js multigrid (src/sim.js:416-495, 588-599) on an nx*ny*nz grid."""

    def __init__(self, dims, h, max_levels=0, fp64=False, omega=1.0):
        self.dims = tuple(dims)
        self.h = lib().orc_mg_create(dims[0], dims[1], dims[2], float(h), int(max_levels), 1 if fp64 else 0)
        if omega != 1.0:
            lib().orc_mg_set_omega(self.h, float(omega))

    def __del__(self):
        try:
            if self.h:
                lib().orc_mg_destroy(self.h)
        except Exception:
            pass

    def levels(self):
        d = (ctypes.c_int * 48)()
        c = ctypes.c_int()
        n = lib().orc_mg_levels(self.h, d, ctypes.byref(c))
        return [tuple(d[i * 3:i * 3 + 3]) for i in range(n)], c.value

    def set(self, b=None, p=None):
        bb = None if b is None else np.ascontiguousarray(b, np.float64).ravel()
        pp = None if p is None else np.ascontiguousarray(p, np.float64).ravel()
        lib().orc_mg_set(self.h, _dp(bb) if bb is not None else None, _dp(pp) if pp is not None else None)

    def get(self):
        out = np.zeros(int(np.prod(self.dims)), np.float64)
        lib().orc_mg_get(self.h, _dp(out))
        return out.reshape(self.dims[::-1])

    def get_x(self):
        out = np.zeros(int(np.prod(self.dims)), np.float64)
        lib().orc_mg_get_x(self.h, _dp(out))
        return out.reshape(self.dims[::-1])

    def vcycles(self, n):
        hist = np.zeros(max(n, 1), np.float64)
        lib().orc_mg_vcycles(self.h, n, _dp(hist))
        return hist[:n]

    def refine(self, tol, max_cycles):
        hist = np.zeros(max_cycles, np.float64)
        used = lib().orc_mg_refine(self.h, tol, max_cycles, _dp(hist))
        return hist[:used]

    def op(self, op, level, arg=0):
        lib().orc_mg_op(self.h, op, level, arg)

    def level_get(self, level):
        dims, _ = self.levels()
        nx, ny, nz = dims[level]
        p = np.zeros(nx * ny * nz)
        b = np.zeros(nx * ny * nz)
        lib().orc_mg_level_get(self.h, level, _dp(p), _dp(b))
        return p.reshape(nz, ny, nx), b.reshape(nz, ny, nx)

    def level_set(self, level, p=None, b=None):
        pp = None if p is None else np.ascontiguousarray(p, np.float64).ravel()
        bb = None if b is None else np.ascontiguousarray(b, np.float64).ravel()
        lib().orc_mg_level_set(self.h, level, _dp(pp) if pp is not None else None, _dp(bb) if bb is not None else None)

    def residual(self):
        return lib().orc_mg_residual(self.h)


def direct_solve(b, h, fp64=False):
    """The exact level solve of the tolerance solves' coarse level (b: (nz, ny, nx)); None if the
    library's kernels do not take the shape."""
    L = lib()
    L.orc_direct_solve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    b = np.ascontiguousarray(b, np.float64)
    nz, ny, nx = b.shape
    p = np.zeros_like(b)
    if L.orc_direct_solve(nx, ny, nz, float(h), 1 if fp64 else 0, _dp(b), _dp(p)):
        return None
    return p


def mjs_study(N, face=True, mode=2, omega=1.0, pre=2, post=2):
    """Re-run research/webgpu-bench/poisson.mjs's study; returns (r0, residual/r0 history)."""
    hist = np.zeros(8, np.float64)
    r0 = lib().orc_mjs_study(N, 1 if face else 0, mode, omega, pre, post, _dp(hist))
    return r0, hist[: (6 if mode in (0, 1) else 8)]


# ---- render-volume consumers (oracle/render_oracle.cpp) --------------------------------
def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def render_lights(U, S, lut, lstate):
    """One light-extraction dispatch (src/renderer.js:112-164).  U: the 88-float uniform block;
    S: (nz, ny, nx, 4) decoded floats; lut: (2, n, 4); lstate: (3, 8, 4) filter state, updated
    in place.  Returns the (8, 8) lights."""
    L = lib()
    L.orc_render_lights.argtypes = [ctypes.POINTER(ctypes.c_float)] * 5
    U, S, lut = _f32(U), _f32(S), _f32(lut)
    assert lstate.dtype == np.float32 and lstate.flags.c_contiguous and lstate.size == 96
    out = np.zeros((8, 8), np.float32)
    L.orc_render_lights(_fptr(U), _fptr(S), _fptr(lut), _fptr(lstate), _fptr(out))
    return out


def render_sharpen(k, S):
    """The flame-field prefilter (src/renderer.js:176-191) of S (nz, ny, nx, 4), rounded to binary16."""
    L = lib()
    L.orc_render_sharpen.argtypes = [ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    S = _f32(S)
    nz, ny, nx, _ = S.shape
    out = np.zeros_like(S)
    L.orc_render_sharpen(float(k), nx, ny, nz, _fptr(S), _fptr(out))
    return out


def render_volume(U, S, S2, lut, lights, curl, depth=None):
    """The volume march (src/renderer.js:235-326) over every pixel: returns vol (vh, vw, 4) (L, Tr)
    rounded to binary16 and aux (vh, vw, 2) (haze, depth).  S is the field marched (the sim's S
    or the prefiltered copy), curl the (n, n, n, 4) curl texture; depth None = sky (1e6)."""
    L = lib()
    fp = ctypes.POINTER(ctypes.c_float)
    L.orc_render_volume.argtypes = [fp, fp, fp, fp, fp, fp, ctypes.c_int, ctypes.c_int, fp, ctypes.c_int, fp, fp]
    U, S, S2, lut, lights, curl = _f32(U), _f32(S), _f32(S2), _f32(lut), _f32(lights), _f32(curl)
    depth = np.full((1, 1), 1e6, np.float32) if depth is None else _f32(depth)
    vw, vh = int(U[2]), int(U[3])
    vol = np.zeros((vh, vw, 4), np.float32)
    aux = np.zeros((vh, vw, 2), np.float32)
    L.orc_render_volume(_fptr(U), _fptr(S), _fptr(S2), _fptr(lut), _fptr(lights), _fptr(depth), depth.shape[1],
                        depth.shape[0], _fptr(curl), curl.shape[0], _fptr(vol), _fptr(aux))
    return vol, aux


# ---- blackbody LUT (src/blackbody.js:4-88), restated with numpy in float64 ----------------
_H, _C, _KB = 6.62607015e-34, 2.99792458e8, 1.380649e-23
_XYZ2RGB = np.array([[3.2404542, -1.5371385, -0.4985314], [-0.9692660, 1.8760108, 0.0415560],
                     [0.0556434, -0.2040259, 1.0572252]])
_BRADFORD = np.array([[0.8951, 0.2664, -0.1614], [-0.7502, 1.7135, 0.0367], [0.0389, -0.0685, 1.0296]])
_D65 = np.array([0.95047, 1.0, 1.08883])


def cie_fit(lnm):
    """The multi-lobe piecewise-Gaussian fit of the CIE 1931 2-degree observer (Wyman, Sloan &
    Shirley 2013), as src/blackbody.js:20-25 writes it."""
    l = np.asarray(lnm, np.float64)

    def g(mu, s1, s2):
        t = (l - mu) / np.where(l < mu, s1, s2)
        return np.exp(-0.5 * t * t)
    x = 1.056 * g(599.8, 37.9, 31.0) + 0.362 * g(442.0, 16.0, 26.7) - 0.065 * g(501.1, 20.4, 26.2)
    y = 0.821 * g(568.8, 46.9, 40.5) + 0.286 * g(530.9, 16.3, 31.1)
    z = 1.217 * g(437.0, 11.8, 36.0) + 0.681 * g(459.0, 26.0, 13.8)
    return np.stack([x, y, z], -1)


def planck(lnm, T):
    """Planck's law, spectral radiance per metre of wavelength."""
    l = np.asarray(lnm, np.float64) * 1e-9
    return (2 * _H * _C * _C) / (l ** 5 * np.expm1((_H * _C) / (l * _KB * T)))


def blackbody_xyz(T, soot=False):
    l = np.arange(380, 781, 2, dtype=np.float64)
    p = planck(l, T) * ((l / 550.0) ** -1.39 if soot else 1.0)
    return (p[:, None] * cie_fit(l)).sum(0)


def blackbody_lut(n=1024, tmin=400.0, tmax=3000.0, twb=5000.0):
    ref = blackbody_xyz(1500.0)[1]
    if twb:
        w = blackbody_xyz(twb)
        src, dst = _BRADFORD @ (w / w[1]), _BRADFORD @ _D65
        M = _XYZ2RGB @ np.linalg.inv(_BRADFORD) @ np.diag(dst / src) @ _BRADFORD
    else:
        M = _XYZ2RGB
    out = np.zeros((2, n, 4), np.float32)
    for row in range(2):
        for i in range(n):
            T = tmin + (tmax - tmin) * i / (n - 1)
            xyz = blackbody_xyz(T, row == 1) / ref
            out[row, i, :3] = np.maximum(M @ xyz, 0.0)
            out[row, i, 3] = xyz[1]
    return out
