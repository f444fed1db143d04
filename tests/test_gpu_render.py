"""GPU parity of the render-volume consumers (fs_render_*, src/renderer.js:104-326) with the oracle.

A campfire on a 32x48x32 grid is stepped on the GPU; each frame's light extraction, prefilter and
volume march are recomputed by the oracle from the GPU's own fields (S, S2, the curl texture) and
the library's uniform block, and must agree bit for bit: lights and aux as float32, the volume
(L, Tr) and the prefiltered field as binary16.  The uniform block itself is checked entry by entry
against uploadParams as written in render_common.make_uniforms.
"""
import numpy as np
import pytest

import firesim as F
import fsoracle as O
from render_common import RP, make_uniforms

pytestmark = pytest.mark.gpu

RES, SIZE, DT = (32, 48, 32), (0.6, 0.9, 0.6), 1.0 / 90.0
CAM = F.orbit(yaw=0.35, pitch=0.12, dist=2.2, target=(0.0, 0.5, 0.0), fov=38.0)   # src/main.js:45-46 defaults
FINAL = dict(cubic=1, flameSharp=1, detailAmt=0.3, sootEdge=0.12, lightK=0.6)        # tools/render_final.mjs:23

CASES = {
    "defaults": dict(params={}, jitter=(0.0, 0.0), depth=False),
    "final": dict(params=FINAL, jitter=(0.31, -0.27), depth=False),
    "lens_depth": dict(params=dict(FINAL, focus=2.2, lensX=0.012, lensY=-0.007, smokeScat=0.5, stepMul=0.5),
                       jitter=(-0.4, 0.15), depth=True),
}


def bits_equal(a, b):
    a, b = np.ascontiguousarray(a, np.float32), np.ascontiguousarray(b, np.float32)
    diff = np.argwhere(a.view(np.uint32) != b.view(np.uint32))
    return len(diff), (diff[:3], a[tuple(diff[0])] if len(diff) else None, b[tuple(diff[0])] if len(diff) else None)


def fire(prec, steps=60):
    logs, _ = F.build_campfire(7)
    sim = F.FireSim(res=RES, size=SIZE, logs=logs, precision=prec, noise_n=32)
    sim.step(DT, steps)
    sim.sync()
    return sim, len(logs) // 8          # log slots the sim holds (its nLogs uniform)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("prec", [F.PREC_FP32, F.PREC_FP16, F.PREC_BF16], ids=["fp32", "fp16", "bf16"])
def test_render_bit_exact_with_oracle(prec, case):
    c = CASES[case]
    print(f"render parity prec {prec} case {case}", flush=True)
    sim, n_logs = fire(prec)
    r = F.Render(sim, width=96, height=128, vol_scale=0.5)
    assert (r.vw, r.vh) == (48, 64)
    r.set(**c["params"])
    r.set_camera(CAM)
    depth = None
    if c["depth"]:
        rng = np.random.default_rng(5)
        depth = (1.6 + 1.2 * rng.random((37, 29))).astype(np.float32)
        r.set_depth(depth)
    lut = r.read_lut()
    np.testing.assert_array_equal(lut, F.blackbody_lut())
    curl = sim.read(F.CURL, np.float32)
    lstate = np.zeros((3, 8, 4), np.float32)
    t_prev = None
    for f in range(2):
        t = sim.time
        r.frame(t, 10 + f, c["jitter"])
        U = r.uniforms()
        want = make_uniforms(CAM, RES, SIZE, c["params"], width=96, height=128, vol_scale=0.5, time=t, frame=10 + f,
                             jitter=c["jitter"], light_dt=-1.0 if t_prev is None else t - t_prev, n_logs=n_logs)
        t_prev = t
        np.testing.assert_array_equal(U, want)
        S, S2 = sim.read(F.S, np.float32), sim.read(F.S2, np.float32)
        L_o = O.render_lights(U, S, lut, lstate)
        L_g = r.read_lights()
        nd, info = bits_equal(L_g, L_o)
        print(f"frame {f}: lights on {int((L_g[:, 3] > 0).sum())}, differing {nd}", flush=True)
        assert nd == 0, info
        field = S
        if c["params"].get("flameSharp", 0) > 0 and c["params"].get("cubic", 0) > 0.5:
            field = O.render_sharpen(U[RP["flameSharp"]], S)
            nd, info = bits_equal(r.read_sharp(), field)
            assert nd == 0, info
        vol_o, aux_o = O.render_volume(U, field, S2, lut, L_g, curl, depth)
        vol_g, aux_g = r.read_vol(), r.read_aux()
        nv, info_v = bits_equal(vol_g, vol_o)
        na, info_a = bits_equal(aux_g, aux_o)
        print(f"frame {f}: mean L {vol_g[..., :3].mean():.4g}, min Tr {vol_g[..., 3].min():.3f}, "
              f"differing vol {nv} aux {na}", flush=True)
        assert nv == 0, info_v
        assert na == 0, info_a
        sim.step(DT)
    assert vol_g[..., :3].max() > 0 and vol_g[..., 3].min() < 1   # the fire is visible


def test_render_sharpen_once_per_state_and_errors():
    print("render state / errors", flush=True)
    sim, _ = fire(F.PREC_FP16, steps=20)
    r = F.Render(sim, width=64, height=64, vol_scale=0.25)
    with pytest.raises(F.FsError, match="camera"):
        r.frame(sim.time, 0)
    r.set_camera(CAM)
    with pytest.raises(F.FsError, match="unknown render parameter"):
        r.set(nope=1.0)
    with pytest.raises(F.FsError, match="set per frame"):
        r.set(vw=3.0)
    with pytest.raises(F.FsError, match="no prefiltered"):
        r.read_sharp()
    r.set(**FINAL)
    r.frame(sim.time, 0)
    a = r.read_sharp()
    S = sim.read(F.S, np.float32)
    sim.write(F.S, S * 0.5)                    # a new state at the same clock is prefiltered again
    r.frame(sim.time, 1)
    b = r.read_sharp()
    assert np.abs(b - 0.5 * a).max() <= 1e-2 * np.abs(a).max()
    small = F.FireSim(res=(16, 8, 16), size=(0.3, 0.15, 0.3), noise_n=16)
    with pytest.raises(F.FsError, match="ny >= 16"):
        F.Render(small, 32, 32)


@pytest.mark.parametrize("params", [dict(FINAL), dict(FINAL, cubic=0, flameSharp=0, smokeScat=0.5), dict(FINAL, detailAmt=6.0)],
                         ids=["final", "trilinear_smoke", "wide_warp"])
def test_empty_space_skip_is_exact(params):
    print(f"empty-space skip {params}", flush=True)
    sim, _ = fire(F.PREC_FP16, steps=90)
    r = F.Render(sim, width=96, height=128, vol_scale=0.5)
    r.set(**params)
    r.set_camera(CAM)
    out = {}
    for on in (True, False):
        r.empty_skip(on)
        r.frame(sim.time, 3)
        out[on] = (r.read_vol(), r.read_aux())
    for a, b in zip(out[True], out[False]):
        nd, info = bits_equal(a, b)
        assert nd == 0, info
    assert out[True][0][..., 3].min() < 1


EDGE_CAMS = {
    "inside_box": F.look_at([0.05, 0.3, 0.02], [0.0, 0.5, 0.0], 70.0),        # rays start inside the volume
    "looking_away": F.look_at([0.0, 0.4, 2.0], [0.0, 0.4, 5.0], 38.0),        # every ray misses the box
    "straight_down": F.look_at([0.001, 2.5, 0.0], [0.0, 0.0, 0.0], 30.0),     # near-vertical rays
}


@pytest.mark.parametrize("cam", list(EDGE_CAMS))
def test_render_edge_cases_bit_exact(cam):
    """Odd volume sizes (vol_scale 0.37 of 101x77), a zero-depth surface patch, and cameras inside the box,
    facing away from it and looking straight down."""
    print(f"render edge case {cam}", flush=True)
    sim, _ = fire(F.PREC_FP16, steps=60)
    r = F.Render(sim, width=101, height=77, vol_scale=0.37)
    assert (r.vw, r.vh) == (37, 28)
    r.set(**FINAL)
    r.set_camera(EDGE_CAMS[cam])
    depth = np.full((9, 11), 1e6, np.float32)
    depth[3:6, 4:7] = 0.0                                  # a surface at the eye: no march there
    r.set_depth(depth)
    r.frame(sim.time, 63, (0.5, -0.5))
    U = r.uniforms()
    S, S2 = sim.read(F.S, np.float32), sim.read(F.S2, np.float32)
    lut = r.read_lut()
    L = r.read_lights()
    field = O.render_sharpen(U[RP["flameSharp"]], S)
    vol_o, aux_o = O.render_volume(U, field, S2, lut, L, sim.read(F.CURL, np.float32), depth)
    vol_g, aux_g = r.read_vol(), r.read_aux()
    for a, b in ((vol_g, vol_o), (aux_g, aux_o)):
        nd, info = bits_equal(a, b)
        assert nd == 0, info
    if cam == "looking_away":
        assert np.all(vol_g[..., 3] == 1) and np.all(vol_g[..., :3] == 0)
    assert np.all(np.isfinite(vol_g)) and np.all(np.isfinite(aux_g))
