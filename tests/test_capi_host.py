"""The C-ABI library loads and exports every entry point include/firesim.h declares;
host-only entry points (no GPU needed) behave like the reference's host code."""
import ctypes
import os
import re

import numpy as np
import pytest

import firesim as F
import fsoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# SIM_FIELDS, src/sim.js:14-23
SIM_FIELDS = ['nx', 'ny', 'nz', 'dx', 'dt', 'time', 'frame', 'seed', 'ox', 'oy', 'oz', 'nLogs',
              'fuelSrc', 'bedFuel', 'srcTemp', 'bedTemp', 'srcVel', 'srcFreq', 'srcSpeed', 'srcThresh',
              'ignT', 'kMix', 'rs', 'heatOut', 'expansion', 'sootForm', 'sootOx', 'radCool',
              'coolRate', 'sootDissip', 'smokeDissip', 'vitDecay', 'fuelDissip', 'buoyancy', 'vortEps', 'turbAmp',
              'turbFreq', 'turbRise', 'turbEvolve', 'windX', 'windZ', 'gust', 'intensity', 'velDissip',
              'bedRadius', 'logSrc', 'burstX', 'burstY', 'burstZ', 'burstAmt', 'smolder', 'sponge', 'hotR0', 'hotR1',
              'hotY1', 'visc', 'lamAge', 'anchorFreq', 'srcFloor', 'anchorGain']


def _declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(fs_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(F.LIB_PATH)
    syms = _declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_param_block_layout_matches_sim_fields():
    assert F.param_names() == SIM_FIELDS
    assert F.lib().fs_param_index(b"vortEps") == SIM_FIELDS.index("vortEps")
    assert F.lib().fs_param_index(b"nope") == -1
    d = F.param_defaults()
    # SIM_DEFAULTS (src/sim.js:25-44)
    want = {'fuelSrc': 0.13, 'bedFuel': 0.16, 'srcTemp': 0.75, 'bedTemp': 0.82, 'srcVel': 0.25, 'srcFreq': 35.0,
            'srcSpeed': 0.06, 'ignT': 550, 'kMix': 30, 'rs': 4.7, 'heatOut': 7700, 'expansion': 0.7, 'sootForm': 25,
            'sootOx': 30, 'radCool': 0.25, 'coolRate': 3.0, 'sootDissip': 0.4, 'smokeDissip': 0.3, 'vitDecay': 1.5,
            'fuelDissip': 0.3, 'buoyancy': 26.8, 'vortEps': 8, 'turbAmp': 0.8, 'turbFreq': 1.5, 'turbRise': 1.5,
            'turbEvolve': 0.6, 'intensity': 1, 'velDissip': 0.05, 'bedRadius': 0.22, 'hotR0': 0.12, 'hotR1': 0.45,
            'hotY1': 0.5, 'logSrc': 1, 'smolder': 0.02, 'sponge': 1, 'anchorFreq': 1.6, 'srcFloor': 0.25,
            'anchorGain': 0.95}
    for k, v in want.items():
        assert d[SIM_FIELDS.index(k)] == np.float32(v), k
    assert np.array_equal(d, O.param_defaults())


def test_campfire_builder_matches_oracle_restatement():
    a, na = F.build_campfire(7)
    b, nb = O.build_campfire(7)
    assert na == nb == 8 and np.array_equal(a, b)
    a, _ = F.build_campfire(3, max_logs=4)
    assert a.shape == (32,)


def test_config_defaults_are_the_reference_defaults():
    cfg = F.FsConfig()
    F.lib().fs_config_default(ctypes.byref(cfg))
    assert list(cfg.res) == [80, 120, 80] and cfg.vcycles == 2 and cfg.noise_n == 128 and cfg.noise_seed == 1
    assert cfg.precision == F.PREC_FP16


def test_errors_are_status_codes_not_exceptions():
    L = F.lib()
    h = ctypes.c_void_p()
    cfg = F.FsConfig()
    L.fs_config_default(ctypes.byref(cfg))
    cfg.noise_n = 100          # not a power of two: rejected before any device call
    rc = L.fs_create(ctypes.byref(cfg), None, 0, ctypes.byref(h))
    assert rc == -1 and b"power of two" in L.fs_last_error()
    assert L.fs_step(None, 0.01, 1) == -1
    assert L.fs_destroy(None) == 0


def test_field_size_limits_are_checked_before_any_device_call():
    L = F.lib()
    h = ctypes.c_void_p()
    cfg = F.FsConfig()
    L.fs_config_default(ctypes.byref(cfg))
    cfg.precision = F.PREC_FP32
    cfg.res[0], cfg.res[1], cfg.res[2] = 512, 1024, 512          # 2^28 cells x 16 B = 4 GiB per field
    assert L.fs_create(ctypes.byref(cfg), None, 0, ctypes.byref(h)) == -1
    assert b"4 GiB" in L.fs_last_error()
    cfg.precision = F.PREC_FP16                                  # 2 GiB per field: past this check
    cfg.noise_n = 100                                            # and stopped by the next one, still on the host
    assert L.fs_create(ctypes.byref(cfg), None, 0, ctypes.byref(h)) == -1
    assert b"power of two" in L.fs_last_error()
