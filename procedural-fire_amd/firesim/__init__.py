"""Python host binding of the gfx950 fire-simulation library (libfiresim.so, C ABI in include/firesim.h).

Mirrors the reference's `FireSim` surface (src/sim.js:48-649): construction options
(res, size, logs, vcycles), `set_params` / `setParams` (src/sim.js:601-604), `reset`
(src/sim.js:606-618) and `step` (src/sim.js:627-648), plus field readback in the
texel order of the reference's 3D textures.  Unknown parameter names raise, like
UniformStruct.set (src/gpu.js:65).

There is no CPU fallback: if libfiresim.so is missing or no GPU is present, the
constructors raise.
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libfiresim.so")

PREC_FP32, PREC_FP16, PREC_BF16 = 0, 1, 2
SOLVE_VCYCLES, SOLVE_REFINE64, SOLVE_TOL32, SOLVE_SPECTRAL64 = 0, 1, 2, 3
(VEL, S, S2, VORT, GEO, VELX, SX, S2X, VELY, SY, S2Y, P, B, NOISE, CURL) = range(15)
FIELD_NAMES = {"vel": VEL, "S": S, "S2": S2, "vort": VORT, "geo": GEO, "velX": VELX, "SX": SX, "S2X": S2X,
               "velY": VELY, "SY": SY, "S2Y": S2Y, "p": P, "b": B, "noise": NOISE, "curl": CURL}
DT_F32, DT_F64, DT_NATIVE = 0, 1, 2
(ST_GEO, ST_ADV1, ST_ADV2, ST_CURL, ST_FORCES, ST_DIV, ST_MG, ST_PROJECT) = range(8)
STAGE_LABELS = ("adv1", "adv2", "curl", "forces", "div", "mg", "project")
NUM_PARAMS = 60
# the FS_BUDGET_* record of include/firesim.h, in order
BUDGET_KEYS = ("steps", "fuel0", "fuel", "fuel_transport", "fuel_sources", "fuel_injected", "fuel_burned",
               "fuel_dissipated", "fuel_removed", "fuel_rounding", "heat0", "heat", "heat_transport", "heat_sources",
               "heat_released", "heat_radiated", "heat_entrained", "heat_removed", "heat_rounding", "fuel_expected",
               "expansion", "divergence", "bad_cells")
BUDGET_N = 24


class FsError(RuntimeError):
    pass


class FsConfig(ctypes.Structure):
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


class FsSpot(ctypes.Structure):
    """A pyrolysis spot: mass flux `flux` (kg/s) spread over a sphere of `radius` (m) at `pos`."""
    _fields_ = [("pos", ctypes.c_float * 3), ("radius", ctypes.c_float), ("flux", ctypes.c_float),
                ("temp", ctypes.c_float), ("norm", ctypes.c_float), ("pad", ctypes.c_float)]


class FsSparkBurst(ctypes.Structure):
    """Sparks.burst({...}) options (src/sparks.js:44-46)."""
    _fields_ = [("n", ctypes.c_int), ("pos", ctypes.c_float * 3), ("speed", ctypes.c_float),
                ("radius", ctypes.c_float), ("dur", ctypes.c_double), ("fan", ctypes.c_float),
                ("size", ctypes.c_float), ("has_aim", ctypes.c_int), ("aim", ctypes.c_float * 3)]


SPARK_FLOATS = 16
MAX_SPOTS = 64
RHO0 = 1.204


def spots_array(spots):
    """(n, 6) rows of (x, y, z, radius, flux_kg_s, temp) -> ctypes array of FsSpot."""
    a = np.asarray(spots, np.float32).reshape(-1, 6)
    arr = (FsSpot * max(len(a), 1))()
    for i, r in enumerate(a):
        arr[i].pos[:] = [float(r[0]), float(r[1]), float(r[2])]
        arr[i].radius, arr[i].flux, arr[i].temp = float(r[3]), float(r[4]), float(r[5])
    return arr, len(a)


_lib = None


def lib():
    """Loads libfiresim.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FsError(f"{LIB_PATH} is missing: build it with `make -C procedural-fire_amd/csrc` "
                      "(or __graft_entry__.build()); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, fp, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)
    sz = ctypes.c_size_t
    sig = {
        "fs_config_default": (None, [ctypes.POINTER(FsConfig)]),
        "fs_last_error": (ctypes.c_char_p, []),
        "fs_version": (ctypes.c_char_p, []),
        "fs_create": (ip, [ctypes.POINTER(FsConfig), fp, ip, ctypes.POINTER(vp)]),
        "fs_destroy": (ip, [vp]),
        "fs_param_count": (ip, []),
        "fs_param_name": (ctypes.c_char_p, [ip]),
        "fs_param_index": (ip, [ctypes.c_char_p]),
        "fs_param_defaults": (ip, [fp]),
        "fs_set_params": (ip, [vp, ctypes.POINTER(ctypes.c_char_p), fp, ip]),
        "fs_get_params": (ip, [vp, fp]),
        "fs_set_logs": (ip, [vp, fp, ip]),
        "fs_set_spots": (ip, [vp, ctypes.POINTER(FsSpot), ip]),
        "fs_get_spots": (ip, [vp, ctypes.POINTER(FsSpot), ip, ctypes.POINTER(ctypes.c_int)]),
        "fs_set_face_flux": (ip, [vp, fp]),
        "fs_reset": (ip, [vp]),
        "fs_step": (ip, [vp, ctypes.c_double, ip]),
        "fs_sync": (ip, [vp]),
        "fs_run_stage": (ip, [vp, ip, ctypes.c_double]),
        "fs_read_field": (ip, [vp, ip, ip, vp, sz]),
        "fs_write_field": (ip, [vp, ip, ip, vp, sz]),
        "fs_field_info": (ip, [vp, ip, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
        "fs_device_ptr": (vp, [vp, ip]),
        "fs_get_clock": (ip, [vp, dp, ctypes.POINTER(ctypes.c_int64)]),
        "fs_grid_info": (ip, [vp, ctypes.POINTER(ctypes.c_int), dp, dp]),
        "fs_mg_info": (ip, [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
        "fs_residual": (ip, [vp, dp, ctypes.POINTER(ctypes.c_int)]),
        "fs_profile_enable": (ip, [vp, ip]),
        "fs_timings": (ip, [vp, fp]),
        "fs_build_campfire": (ip, [ip, fp, ip]),
        "fs_bench_copy": (ip, [ip, ctypes.c_size_t, ip, fp]),
        "fs_poisson_create": (ip, [ctypes.POINTER(ctypes.c_int), ctypes.c_double, ip, ip, vp, ctypes.POINTER(vp)]),
        "fs_poisson_destroy": (ip, [vp]),
        "fs_poisson_set_rhs": (ip, [vp, ip, vp]),
        "fs_poisson_set_x": (ip, [vp, ip, vp]),
        "fs_poisson_solve": (ip, [vp, ctypes.c_double, ip, dp, ctypes.POINTER(ctypes.c_int)]),
        "fs_poisson_vcycles": (ip, [vp, ip, dp]),
        "fs_poisson_read_x": (ip, [vp, ip, vp]),
        "fs_poisson_set_omega": (ip, [vp, ctypes.c_float]),
        "fs_stats": (ip, [vp, ip, dp]),
        "fs_stats_noalloc": (ip, [vp, ip, dp]),
        "fs_bench_stream": (ip, [ip, ctypes.c_size_t, ip, ip, ctypes.POINTER(ctypes.c_float)]),
        "fs_profile": (ip, [vp, dp, dp, ip, ctypes.POINTER(ctypes.c_int)]),
        "fs_diag_enable": (ip, [vp, ip, ctypes.c_double]),
        "fs_diag_read": (ip, [vp, dp, ip, ctypes.POINTER(ctypes.c_int)]),
        "fs_budget_enable": (ip, [vp, ip]),
        "fs_budget_read": (ip, [vp, dp]),
        "fs_poisson_info": (ip, [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
        "fs_set_clock": (ip, [vp, ctypes.c_double, ctypes.c_int64]),
        "fs_spark_burst_default": (None, [ctypes.POINTER(FsSparkBurst)]),
        "fs_sparks_create": (ip, [vp, ip, ctypes.POINTER(vp)]),
        "fs_sparks_destroy": (ip, [vp]),
        "fs_sparks_reset": (ip, [vp]),
        "fs_sparks_set": (ip, [vp, ctypes.c_char_p, ctypes.c_float]),
        "fs_sparks_burst": (ip, [vp, ctypes.POINTER(FsSparkBurst)]),
        "fs_sparks_step": (ip, [vp, ctypes.c_double]),
        "fs_sparks_read": (ip, [vp, vp, sz]),
        "fs_sparks_write": (ip, [vp, vp, sz]),
        "fs_sparks_clock": (ip, [vp, dp, ctypes.POINTER(ctypes.c_int64)]),
        "fs_sparks_device_ptr": (vp, [vp]),
        "fs_adjoint_create": (ip, [vp, ip, ip, ctypes.POINTER(vp)]),
        "fs_adjoint_destroy": (ip, [vp]),
        "fs_adjoint_forward": (ip, [vp, ctypes.c_double, ip]),
        "fs_adjoint_seed": (ip, [vp, ip, ip, vp, sz]),
        "fs_adjoint_seed_mismatch": (ip, [vp, fp, fp, dp]),
        "fs_adjoint_backward": (ip, [vp, ctypes.c_double, ip]),
        "fs_adjoint_tangent": (ip, [vp, ctypes.c_double, ip, dp, fp, fp, fp, fp, ctypes.c_double, ip]),
        "fs_adjoint_read": (ip, [vp, ip, ip, vp, sz]),
        "fs_adjoint_grads": (ip, [vp, dp, ip, dp]),
        "fs_render_param_count": (ip, []),
        "fs_render_param_index": (ip, [ctypes.c_char_p]),
        "fs_render_param_defaults": (ip, [fp]),
        "fs_render_create": (ip, [vp, ip, ip, ctypes.c_double, ctypes.POINTER(vp)]),
        "fs_render_destroy": (ip, [vp]),
        "fs_render_set": (ip, [vp, ctypes.c_char_p, ctypes.c_float]),
        "fs_render_get_uniforms": (ip, [vp, fp]),
        "fs_render_set_camera": (ip, [vp, dp, dp, dp, dp, ctypes.c_double]),
        "fs_render_set_white_balance": (ip, [vp, ctypes.c_double]),
        "fs_render_set_depth": (ip, [vp, fp, ip, ip]),
        "fs_render_frame": (ip, [vp, ctypes.c_double, ctypes.c_double, ctypes.c_float, ctypes.c_float]),
        "fs_render_size": (ip, [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
        "fs_render_read": (ip, [vp, ip, vp, sz]),
        "fs_render_device_ptr": (vp, [vp, ip]),
        "fs_blackbody_lut": (ip, [ip, ctypes.c_double, ctypes.c_double, ctypes.c_double, fp]),
        "fs_spectral_basis": (ip, [ip, ip, dp, dp]),
        "fs_render_profile_enable": (ip, [vp, ip]),
        "fs_render_empty_skip": (ip, [vp, ip]),
        "fs_render_timings": (ip, [vp, fp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise FsError(f"firesim error {rc}: {lib().fs_last_error().decode()}")


def param_names():
    return [lib().fs_param_name(i).decode() for i in range(lib().fs_param_count())]


def param_defaults():
    out = np.zeros(NUM_PARAMS, np.float32)
    _check(lib().fs_param_defaults(out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return out


def build_campfire(seed=7, max_logs=16):
    """buildCampfire(seed).logs packed by packLogs (src/scene.js:21-42, 92-98)."""
    out = np.zeros(max_logs * 8, np.float32)
    n = lib().fs_build_campfire(seed, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), max_logs)
    if n < 0:
        _check(n)
    return out, n


def bench_copy(nbytes=256 << 20, iters=10, device=0):
    """ms per device-to-device copy of `nbytes` with a plain streaming kernel."""
    ms = ctypes.c_float()
    _check(lib().fs_bench_copy(int(device), int(nbytes), int(iters), ctypes.byref(ms)))
    return ms.value


STREAM_KINDS = {"copy_plain": 0, "copy": 1, "read": 2, "write": 3}


def bench_stream(kind="copy", nbytes=1 << 30, iters=10, device=0):
    """The achievable-bandwidth probe: GB/s of one pass of `kind` over `nbytes` (copies count the bytes
    read and written)."""
    ms = ctypes.c_float()
    k = STREAM_KINDS[kind]
    _check(lib().fs_bench_stream(int(device), int(nbytes), k, int(iters), ctypes.byref(ms)))
    moved = nbytes * (2 if k in (0, 1) else 1)
    return moved / (ms.value * 1e-3) / 1e9


class FireSim:
    """One simulation grid on one GPU: the C-ABI counterpart of `new FireSim(device, {...})`."""

    def __init__(self, res=(80, 120, 80), size=(1.2, 1.8, 1.2), logs=None, vcycles=2, max_levels=0,
                 precision=PREC_FP16, solve=SOLVE_VCYCLES, div_tol=1e-5, max_vcycles=8, noise_n=128,
                 noise_seed=1, device=0, stream=None, use_graph=False, adaptive=True, omega=0.0, pre_sweeps=0,
                 post_sweeps=0, direct_level=0, first_cycles=0):
        L = lib()
        cfg = FsConfig()
        L.fs_config_default(ctypes.byref(cfg))
        cfg.res[:] = [int(v) for v in res]
        cfg.size[:] = [float(v) for v in size]
        cfg.vcycles = int(vcycles)
        cfg.max_levels = int(max_levels)
        cfg.precision = int(precision)
        cfg.solve = int(solve)
        cfg.div_tol = float(div_tol)
        cfg.max_vcycles = int(max_vcycles)
        cfg.noise_n = int(noise_n)
        cfg.noise_seed = int(noise_seed)
        cfg.device = int(device)
        cfg.stream = stream
        cfg.use_graph = 1 if use_graph else 0
        cfg.adaptive = 1 if adaptive else 0
        cfg.omega = float(omega)
        cfg.pre_sweeps = int(pre_sweeps)
        cfg.post_sweeps = int(post_sweeps)
        cfg.direct_level = int(direct_level)
        cfg.first_cycles = int(first_cycles)
        if logs is None:
            logs, _ = build_campfire(7)   # src/main.js:34-35: buildCampfire(7) packed to 16 slots
        logs = np.ascontiguousarray(logs, np.float32)
        self._h = ctypes.c_void_p()
        _check(L.fs_create(ctypes.byref(cfg), logs.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), logs.size // 8,
                           ctypes.byref(self._h)))
        self.res = tuple(int(v) for v in res)
        self.size = tuple(float(v) for v in size)
        self.precision = precision
        self.solve = solve
        self.vcycles = vcycles

    # ---- lifetime
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().fs_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- parameters (src/sim.js:601-604)
    def set_params(self, params=None, **kw):
        d = dict(params or {})
        d.update(kw)
        if not d:
            return
        names = (ctypes.c_char_p * len(d))(*[k.encode() for k in d])
        vals = np.array([float(v) for v in d.values()], np.float32)
        _check(lib().fs_set_params(self._h, names, vals.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(d)))

    setParams = set_params

    @property
    def params(self):
        out = np.zeros(NUM_PARAMS, np.float32)
        _check(lib().fs_get_params(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return dict(zip(param_names(), out.tolist()))

    def set_logs(self, logs):
        logs = np.ascontiguousarray(logs, np.float32)
        _check(lib().fs_set_logs(self._h, logs.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), logs.size // 8))

    # ---- build-extension sources
    def set_spots(self, spots):
        """Pyrolysis spots, rows of (x, y, z, radius, flux_kg_s, temp); [] removes them."""
        arr, n = spots_array(spots)
        _check(lib().fs_set_spots(self._h, arr, n))

    def spots(self):
        """The spots as set, with their mass normalisation (columns x, y, z, r, flux, temp, norm)."""
        arr = (FsSpot * MAX_SPOTS)()
        n = ctypes.c_int()
        _check(lib().fs_get_spots(self._h, arr, MAX_SPOTS, ctypes.byref(n)))
        return np.array([[*a.pos, a.radius, a.flux, a.temp, a.norm] for a in arr[:n.value]], np.float64).reshape(-1, 7)

    def set_face_flux(self, flux):
        """Floor fuel mass flux (kg/m^2/s), shape (nz, nx); None removes it."""
        if flux is None:
            _check(lib().fs_set_face_flux(self._h, None))
            return
        a = np.ascontiguousarray(flux, np.float32)
        nx, _, nz = self.res
        if a.shape != (nz, nx):
            raise ValueError(f"face flux must have shape {(nz, nx)}")
        _check(lib().fs_set_face_flux(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    # ---- diagnostics (src/main.js:268-355)
    def stats(self, field="S"):
        """FIRE.stats(name) reduced on the device: per channel min, max, mean, bad (non-finite
        count) and argmax (i, j, k) of the first maximum, unrounded."""
        out = np.zeros(20, np.float64)
        _check(lib().fs_stats(self._h, FIELD_NAMES.get(field, field), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        nx, ny, _ = self.res
        res = []
        for c in range(4):
            mn, mx, mean, bad, arg = out[5 * c:5 * c + 5]
            a = int(arg)
            res.append({"min": mn, "max": mx, "mean": mean, "bad": int(bad),
                        "argmax": None if a < 0 else (a % nx, (a // nx) % ny, a // (nx * ny))})
        return res

    def profile(self):
        """FIRE.profile() reduced on the device: heat release (kW) and rows every 4th y-plane
        (h, Tc, Tmax, soot, fuel, hotCells), unrounded."""
        nrow = (self.res[1] + 3) // 4
        rows = np.zeros(nrow * 6, np.float64)
        q = ctypes.c_double()
        n = ctypes.c_int()
        _check(lib().fs_profile(self._h, ctypes.byref(q), rows.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nrow,
                                ctypes.byref(n)))
        keys = ("h", "Tc", "Tmax", "soot", "fuel", "hotCells")
        return {"Q_kW": q.value / 1000.0, "rows": [dict(zip(keys, rows[6 * r:6 * r + 6])) for r in range(n.value)]}

    def diag_enable(self, capacity=4096, t_flame=1000.0):
        """Per-step record (time, heat release, max T, flame tip, axis T per plane) into a
        device ring of `capacity` steps; 0 turns it off."""
        _check(lib().fs_diag_enable(self._h, int(capacity), float(t_flame)))
        self._diag_cap = int(capacity)

    def diag_series(self):
        """The held per-step records, oldest first: dict of arrays time, Q (W), Tmax (K), tip (m),
        axisT (steps x ny, K)."""
        cap = getattr(self, "_diag_cap", 0)
        ny = self.res[1]
        buf = np.zeros(max(cap, 1) * (4 + ny), np.float64)
        n = ctypes.c_int()
        _check(lib().fs_diag_read(self._h, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), cap, ctypes.byref(n)))
        a = buf[:n.value * (4 + ny)].reshape(n.value, 4 + ny)
        return {"time": a[:, 0].copy(), "Q": a[:, 1].copy(), "Tmax": a[:, 2].copy(), "tip": a[:, 3].copy(),
                "axisT": a[:, 4:].copy()}

    def budget_enable(self, on=True):
        """FP64 mass and energy budget of the following steps, reduced on the device (fs_budget_*);
        on=True restarts the record from the current state."""
        _check(lib().fs_budget_enable(self._h, 1 if on else 0))

    def budget(self):
        """The budget record as a dict (BUDGET_KEYS; kg, J, m^3/s)."""
        buf = np.zeros(BUDGET_N, np.float64)
        _check(lib().fs_budget_read(self._h, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return dict(zip(BUDGET_KEYS, buf.tolist()))

    # ---- stepping
    def reset(self):
        _check(lib().fs_reset(self._h))

    def step(self, dt, n=1):
        _check(lib().fs_step(self._h, float(dt), int(n)))

    def sync(self):
        _check(lib().fs_sync(self._h))

    def run_stage(self, stage, dt=0.0):
        _check(lib().fs_run_stage(self._h, int(stage), float(dt)))

    # ---- fields
    def field_info(self, field):
        field = FIELD_NAMES.get(field, field)
        d = (ctypes.c_int * 3)()
        ch, nb = ctypes.c_int(), ctypes.c_int()
        _check(lib().fs_field_info(self._h, field, d, ctypes.byref(ch), ctypes.byref(nb)))
        return tuple(d), ch.value, nb.value

    def read(self, field, dtype=np.float32):
        """Field as an array of shape (nz, ny, nx[, 4]) (x fastest, like the texture)."""
        field = FIELD_NAMES.get(field, field)
        (nx, ny, nz), ch, nb = self.field_info(field)
        shape = (nz, ny, nx, ch) if ch > 1 else (nz, ny, nx)
        if dtype == "native":
            raw = {2: np.uint16, 4: np.float32, 8: np.float64}[nb]
            out = np.empty(shape, raw)
            code = DT_NATIVE
        else:
            out = np.empty(shape, np.dtype(dtype))
            code = DT_F64 if out.dtype == np.float64 else DT_F32
        _check(lib().fs_read_field(self._h, field, code, out.ctypes.data, out.nbytes))
        return out

    def write(self, field, arr, native=False):
        field = FIELD_NAMES.get(field, field)
        a = np.ascontiguousarray(arr)
        if native:
            code = DT_NATIVE
        elif a.dtype == np.float64:
            code = DT_F64
        else:
            a = np.ascontiguousarray(a, np.float32)
            code = DT_F32
        _check(lib().fs_write_field(self._h, field, code, a.ctypes.data, a.nbytes))

    def device_ptr(self, field):
        return lib().fs_device_ptr(self._h, FIELD_NAMES.get(field, field))

    # ---- state / diagnostics
    @property
    def time(self):
        t = ctypes.c_double()
        _check(lib().fs_get_clock(self._h, ctypes.byref(t), None))
        return t.value

    @property
    def frame(self):
        f = ctypes.c_int64()
        _check(lib().fs_get_clock(self._h, None, ctypes.byref(f)))
        return f.value

    def set_clock(self, time, frame):
        _check(lib().fs_set_clock(self._h, float(time), int(frame)))

    def save_state(self):
        """(vel, S, S2, p, time, frame): everything a later step reads besides the
        parameters, spots and geometry (the FP64 pressure is the solve's starting guess)."""
        return (self.read("vel"), self.read("S"), self.read("S2"), self.read("p", np.float64), self.time, self.frame)

    def load_state(self, st):
        vel, S_, S2_, p, t, f = st
        self.write("vel", vel)
        self.write("S", S_)
        self.write("S2", S2_)
        self.write("p", p)
        self.set_clock(t, f)

    def grid_info(self):
        r = (ctypes.c_int * 3)()
        dx = ctypes.c_double()
        o = (ctypes.c_double * 3)()
        _check(lib().fs_grid_info(self._h, r, ctypes.byref(dx), o))
        return tuple(r), dx.value, tuple(o)

    @property
    def dx(self):
        return self.grid_info()[1]

    def mg_info(self):
        n = ctypes.c_int()
        d = (ctypes.c_int * 48)()
        c = ctypes.c_int()
        _check(lib().fs_mg_info(self._h, ctypes.byref(n), d, ctypes.byref(c)))
        return [tuple(d[3 * i:3 * i + 3]) for i in range(n.value)], c.value

    def residual(self):
        r = ctypes.c_double()
        u = ctypes.c_int()
        _check(lib().fs_residual(self._h, ctypes.byref(r), ctypes.byref(u)))
        return r.value, u.value

    def profile_stages(self, on=True):
        """Per-stage device timing of the following steps (the Profiler labels, src/sim.js:629-644)."""
        _check(lib().fs_profile_enable(self._h, 1 if on else 0))

    def timings(self):
        ms = np.zeros(len(STAGE_LABELS), np.float32)
        _check(lib().fs_timings(self._h, ms.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return dict(zip(STAGE_LABELS, ms.tolist()))


class Adjoint:
    """Reverse-mode derivative of a run of steps of one FireSim (fs_adjoint_*).

    forward(dt, n) runs the steps (the sim ends in the final state) keeping the state every
    `checkpoint_every` steps; seed the adjoint of the final state (seed / seed_mismatch), then
    backward() gives the adjoint of the initial state (read) and the gradients of the spot
    fluxes and the floor flux (grads).  Needs FP32 storage and the FP64-refined solve.
    """

    def __init__(self, sim, max_steps, checkpoint_every=20):
        self.sim = sim
        self._h = ctypes.c_void_p()
        _check(lib().fs_adjoint_create(sim._h, int(max_steps), int(checkpoint_every), ctypes.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().fs_adjoint_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, dt, n):
        _check(lib().fs_adjoint_forward(self._h, float(dt), int(n)))

    def seed(self, field, arr):
        field = FIELD_NAMES.get(field, field)
        a = np.ascontiguousarray(arr, np.float32)
        _check(lib().fs_adjoint_seed(self._h, field, DT_F32, a.ctypes.data, a.nbytes))

    def seed_mismatch(self, target_theta, target_soot):
        """Seeds L = sum (theta - theta*)^2 + (soot - soot*)^2 over the final state; returns L."""
        t = np.ascontiguousarray(target_theta, np.float32)
        s = np.ascontiguousarray(target_soot, np.float32)
        loss = ctypes.c_double()
        _check(lib().fs_adjoint_seed_mismatch(self._h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                              s.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(loss)))
        return loss.value

    def backward(self, solve_tol=0.0, max_cycles=0):
        _check(lib().fs_adjoint_backward(self._h, float(solve_tol), int(max_cycles)))

    def tangent(self, dt, n, spot_dir=None, face_dir=None, state_dir=None, solve_tol=0.0, max_cycles=0):
        """Forward mode: n steps (recorded like forward()) carrying the derivative along the
        direction (spot fluxes [n_spots], floor flux (nz, nx), initial (vel, S, S2)); returns the
        final state's tangent {"vel", "S", "S2"}."""
        P = ctypes.POINTER(ctypes.c_float)
        keep = []

        def fptr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.float32)
            keep.append(a)
            return a.ctypes.data_as(P)
        sd = None if spot_dir is None else np.ascontiguousarray(spot_dir, np.float64)
        vd, Sd, S2d = state_dir if state_dir is not None else (None, None, None)
        _check(lib().fs_adjoint_tangent(self._h, float(dt), int(n),
                                        None if sd is None else sd.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        fptr(face_dir), fptr(vd), fptr(Sd), fptr(S2d), float(solve_tol),
                                        int(max_cycles)))
        return {f: self.read(f) for f in ("vel", "S", "S2")}

    def read(self, field):
        """After backward(): the adjoint of the initial state's field; after tangent(): the final
        state's tangent.  Shape (nz, ny, nx, 4)."""
        field = FIELD_NAMES.get(field, field)
        nx, ny, nz = self.sim.res
        out = np.empty((nz, ny, nx, 4), np.float32)
        _check(lib().fs_adjoint_read(self._h, field, DT_F32, out.ctypes.data, out.nbytes))
        return out

    def grads(self, n_spots=None, face=False):
        """(dL/d spot flux [n_spots], dL/d floor flux [nz, nx] or None), FP64."""
        if n_spots is None:
            n_spots = len(self.sim.spots())
        g = np.zeros(max(n_spots, 1), np.float64)
        nx, _, nz = self.sim.res
        gf = np.zeros((nz, nx), np.float64) if face else None
        _check(lib().fs_adjoint_grads(self._h, g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), int(n_spots),
                                      gf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if face else None))
        return g[:n_spots], gf


class Sparks:
    """Spark particles carried by a FireSim's velocity (the reference's Sparks, src/sparks.js:17-271):
    step(dt) right after the sim's step with the same dt; read() gives (count, 16) float32 rows
    (position, age | velocity, burn time | previous position, T | diameter mm, seed, alive, flicker)."""

    def __init__(self, sim, count=8192):
        self.sim = sim
        self.count = int(count)
        self._h = ctypes.c_void_p()
        _check(lib().fs_sparks_create(sim._h, self.count, ctypes.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().fs_sparks_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        _check(lib().fs_sparks_reset(self._h))

    def set(self, **kw):
        """rate, turb, lifeK, sizeK, upK, brightness"""
        for k, v in kw.items():
            _check(lib().fs_sparks_set(self._h, k.encode(), float(v)))

    def burst(self, n=60, pos=(0.0, 0.1, 0.0), speed=3.0, radius=0.05, dur=0.12, fan=0.6, size=1.3, aim=None):
        b = FsSparkBurst()
        lib().fs_spark_burst_default(ctypes.byref(b))
        b.n = int(n)
        b.pos[:] = [float(v) for v in pos]
        b.speed, b.radius, b.dur, b.fan, b.size = float(speed), float(radius), float(dur), float(fan), float(size)
        if aim is not None:
            b.has_aim = 1
            b.aim[:] = [float(v) for v in aim]
        _check(lib().fs_sparks_burst(self._h, ctypes.byref(b)))

    def step(self, dt):
        _check(lib().fs_sparks_step(self._h, float(dt)))

    def read(self):
        out = np.empty((self.count, SPARK_FLOATS), np.float32)
        _check(lib().fs_sparks_read(self._h, out.ctypes.data, out.nbytes))
        return out

    def write(self, arr):
        a = np.ascontiguousarray(arr, np.float32)
        _check(lib().fs_sparks_write(self._h, a.ctypes.data, a.nbytes))

    @property
    def time(self):
        t = ctypes.c_double()
        _check(lib().fs_sparks_clock(self._h, ctypes.byref(t), None))
        return t.value


def spectral_basis(n, kind):
    """Bases of the exact pressure solve (host only): kind 0 = DST-II (Dirichlet faces at both ends),
    1 = DCT-IV (Neumann low end, Dirichlet high end).  Returns (T, lam): T[k] the orthonormal mode k,
    lam[k] the eigenvalue of the 1D second difference with that boundary rule (h = 1)."""
    T = np.zeros((n, n), np.float64)
    lam = np.zeros(n, np.float64)
    _check(lib().fs_spectral_basis(int(n), int(kind), T.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   lam.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return T, lam


RP_COUNT = 88
RENDER_VOL, RENDER_AUX, RENDER_LIGHTS, RENDER_SHARP, RENDER_LUT = range(5)
RENDER_NLIGHTS = 8


def render_param_index(name):
    return lib().fs_render_param_index(name.encode())


def render_param_defaults():
    out = np.zeros(RP_COUNT, np.float32)
    _check(lib().fs_render_param_defaults(out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return out


def blackbody_lut(n=1024, tmin=400.0, tmax=3000.0, twb=5000.0):
    """blackbodyLUT({n, Tmin, Tmax, Twb}) (src/blackbody.js:78-88): (2, n, 4) float32, row 0 the
    blackbody, row 1 soot; columns linear sRGB radiance relative to a 1500 K blackbody's luminance,
    then the unbalanced luminance Y."""
    out = np.zeros((2, n, 4), np.float32)
    _check(lib().fs_blackbody_lut(int(n), float(tmin), float(tmax), float(twb),
                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return out


def _vnorm(a):
    l = math.hypot(*a) or 1.0
    return [a[0] / l, a[1] / l, a[2] / l]


def look_at(pos, target, fov_y_deg=40.0, roll=0.0):
    """lookAt (src/camera.js:6-20): dict with pos, target, fwd, right, up, fovY (rad)."""
    fwd = _vnorm([target[i] - pos[i] for i in range(3)])
    right = _vnorm([fwd[1] * 0 - fwd[2] * 1, fwd[2] * 0 - fwd[0] * 0, fwd[0] * 1 - fwd[1] * 0])
    up = [right[1] * fwd[2] - right[2] * fwd[1], right[2] * fwd[0] - right[0] * fwd[2], right[0] * fwd[1] - right[1] * fwd[0]]
    if roll:
        c, s_ = math.cos(roll), math.sin(roll)
        right, up = [right[i] * c + up[i] * s_ for i in range(3)], [up[i] * c - right[i] * s_ for i in range(3)]
    return dict(pos=list(pos), target=list(target), fwd=fwd, right=right, up=up, fovY=fov_y_deg * math.pi / 180)


def orbit(yaw=0.0, pitch=0.15, dist=2.4, target=(0.0, 0.45, 0.0), fov=40.0):
    """orbit (src/camera.js:22-29)."""
    pos = [target[0] + dist * math.cos(pitch) * math.sin(yaw), target[1] + dist * math.sin(pitch),
           target[2] + dist * math.cos(pitch) * math.cos(yaw)]
    return look_at(pos, target, fov)


class Render:
    """The Renderer passes that consume a FireSim's S / S2 (src/renderer.js:104-326): fire-light
    extraction, flame-field prefilter and the volume ray-march, on the sim's GPU and stream.
    frame() renders the sim's current fields; read_vol() gives (vh, vw, 4) float32 (L.rgb, Tr),
    read_aux() (vh, vw, 2) (haze, surface depth), read_lights() (8, 8) (pos, luminance, colour,
    spread)."""

    def __init__(self, sim, width=960, height=1080, vol_scale=0.5):
        self.sim = sim
        self._h = ctypes.c_void_p()
        _check(lib().fs_render_create(sim._h, int(width), int(height), float(vol_scale), ctypes.byref(self._h)))
        vw, vh = ctypes.c_int(), ctypes.c_int()
        _check(lib().fs_render_size(self._h, ctypes.byref(vw), ctypes.byref(vh)))
        self.vw, self.vh = vw.value, vh.value

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().fs_render_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, **kw):
        """this.params entries: exposure, stepMul, sootAbs, cubic, flameSharp, detailAmt, ..."""
        for k, v in kw.items():
            _check(lib().fs_render_set(self._h, k.encode(), float(v)))

    def set_camera(self, cam):
        """cam: a look_at / orbit dict (pos, fwd, right, up, fovY)."""
        arr = [(ctypes.c_double * 3)(*[float(x) for x in cam[k]]) for k in ("pos", "fwd", "right", "up")]
        _check(lib().fs_render_set_camera(self._h, *arr, float(cam["fovY"])))

    def set_white_balance(self, twb):
        _check(lib().fs_render_set_white_balance(self._h, float(twb)))

    def set_depth(self, depth=None):
        """Surface depth (dh, dw) float32 in metres along the ray; None = no surfaces (sky)."""
        if depth is None:
            _check(lib().fs_render_set_depth(self._h, None, 0, 0))
            return
        d = np.ascontiguousarray(depth, np.float32)
        _check(lib().fs_render_set_depth(self._h, d.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), d.shape[1], d.shape[0]))

    def frame(self, time, frame, jitter=(0.0, 0.0)):
        _check(lib().fs_render_frame(self._h, float(time), float(frame), float(jitter[0]), float(jitter[1])))

    def profile(self, on=True):
        _check(lib().fs_render_profile_enable(self._h, int(bool(on))))

    def empty_skip(self, on=True):
        """Skip march samples that provably contribute nothing (default on; bit-identical results)."""
        _check(lib().fs_render_empty_skip(self._h, int(bool(on))))

    def timings(self):
        """Device ms of the last frame's passes: {"lights", "sharpen", "volume"} (sharpen 0 if skipped)."""
        ms = (ctypes.c_float * 3)()
        _check(lib().fs_render_timings(self._h, ms))
        return {"lights": ms[0], "sharpen": ms[1], "volume": ms[2]}

    def uniforms(self):
        out = np.zeros(RP_COUNT, np.float32)
        _check(lib().fs_render_get_uniforms(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def _read(self, what, shape):
        out = np.empty(shape, np.float32)
        _check(lib().fs_render_read(self._h, what, out.ctypes.data, out.nbytes))
        return out

    def read_vol(self):
        return self._read(RENDER_VOL, (self.vh, self.vw, 4))

    def read_aux(self):
        return self._read(RENDER_AUX, (self.vh, self.vw, 2))

    def read_lights(self):
        return self._read(RENDER_LIGHTS, (RENDER_NLIGHTS, 8))

    def read_sharp(self):
        nx, ny, nz = self.sim.res
        return self._read(RENDER_SHARP, (nz, ny, nx, 4))

    def read_lut(self):
        return self._read(RENDER_LUT, (2, 1024, 4))

    def device_ptr(self, what):
        return lib().fs_render_device_ptr(self._h, what)


class Poisson:
    """Isolated pressure solve with the sim's operator and multigrid."""

    def __init__(self, dims, h, max_levels=0, device=0, stream=None, omega=1.0):
        L = lib()
        d = (ctypes.c_int * 3)(*[int(v) for v in dims])
        self._h = ctypes.c_void_p()
        _check(L.fs_poisson_create(d, float(h), int(max_levels), int(device), stream, ctypes.byref(self._h)))
        self.dims = tuple(int(v) for v in dims)
        if omega != 1.0:
            _check(L.fs_poisson_set_omega(self._h, float(omega)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().fs_poisson_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_rhs(self, b):
        a = np.ascontiguousarray(b, np.float64)
        _check(lib().fs_poisson_set_rhs(self._h, DT_F64, a.ctypes.data))

    def set_x(self, x):
        a = np.ascontiguousarray(x, np.float64)
        _check(lib().fs_poisson_set_x(self._h, DT_F64, a.ctypes.data))

    def solve(self, tol, max_cycles=32):
        hist = np.zeros(max_cycles + 1, np.float64)
        n = ctypes.c_int()
        _check(lib().fs_poisson_solve(self._h, float(tol), int(max_cycles), hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      ctypes.byref(n)))
        return hist[: n.value + 1]

    def vcycles(self, cycles, history=True):
        hist = np.zeros(max(cycles, 1), np.float64)
        _check(lib().fs_poisson_vcycles(self._h, int(cycles),
                                        hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if history else None))
        return hist[:cycles] if history else None

    def read_x(self, fp64=True):
        nx, ny, nz = self.dims
        out = np.empty((nz, ny, nx), np.float64 if fp64 else np.float32)
        _check(lib().fs_poisson_read_x(self._h, DT_F64 if fp64 else DT_F32, out.ctypes.data))
        return out

    def info(self):
        n = ctypes.c_int()
        d = (ctypes.c_int * 48)()
        c = ctypes.c_int()
        _check(lib().fs_poisson_info(self._h, ctypes.byref(n), d, ctypes.byref(c)))
        return [tuple(d[3 * i:3 * i + 3]) for i in range(n.value)], c.value
