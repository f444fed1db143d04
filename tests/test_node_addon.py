"""The N-API addon (procedural-fire_amd/lib/firesim.node) and the Node FireSim class over it."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import firesim as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "procedural-fire_amd", "lib", "firesim.node")
JS = os.path.join(ROOT, "procedural-fire_amd", "node", "firesim.js")
NODE = shutil.which("node")

needs_node = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON),
                                reason="node or the built addon (make -C procedural-fire_amd/node) is missing")


def _node(code):
    return subprocess.run([NODE, "-e", code], capture_output=True, text=True, timeout=60, cwd=ROOT)


@needs_node
def test_addon_loads_and_mirrors_sim_fields():
    r = _node(f"const m = require({JS!r}); console.log(JSON.stringify({{fields: m.SIM_FIELDS, "
              f"defaults: m.SIM_DEFAULTS, logs: Array.from(m.buildCampfireLogs(7))}}))")
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    assert d["fields"] == F.param_names()
    assert abs(d["defaults"]["buoyancy"] - 26.8) < 1e-5 and d["defaults"]["vortEps"] == 8
    logs, _ = F.build_campfire(7)
    assert np.array_equal(np.array(d["logs"], np.float32), logs)


@needs_node
def test_addon_reports_errors_as_js_exceptions():
    r = _node(f"const m = require({JS!r}); try {{ new m.FireSim({{noiseN: 100}}); console.log('no'); }} "
              f"catch (e) {{ console.log('threw:' + e.message); }}")
    assert r.returncode == 0, r.stderr
    assert "threw:" in r.stdout and "power of two" in r.stdout


@pytest.mark.gpu
@needs_node
def test_node_steps_match_the_c_abi_bit_for_bit(tmp_path):
    print("node run_steps 32x48x32 x10", flush=True)
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "run_steps.js"), str(tmp_path), "10"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    meta = json.loads((tmp_path / "meta.json").read_text())
    assert meta["frame"] == 10 and meta["threwOnUnknown"]
    sim = F.FireSim(res=(32, 48, 32), size=(1.2, 1.8, 1.2), noise_n=32, use_graph=True)
    sim.set_params(vortEps=4, intensity=1.2)
    sim.step(1 / 90, 10)
    for f, fid in (("vel", F.VEL), ("S", F.S), ("S2", F.S2)):
        a = np.fromfile(tmp_path / f"{f}.bin", np.uint16)
        assert np.array_equal(a, sim.read(fid, "native").ravel()), f
    # FIRE.stats semantics (src/main.js:268-294) against numpy on the same field
    S = sim.read(F.S)
    for c in range(4):
        st = meta["stats"][c]
        assert abs(st["max"] - float(f"{S[..., c].max():.4g}")) <= 1e-12
        assert abs(st["mean"] - float(f"{S[..., c].mean():.4g}")) <= 1e-9 + 1e-3 * abs(st["mean"])
    # FIRE.profile heat release (src/main.js:347-352)
    q = S[..., 3].astype(np.float64) * 1e4
    T = 293 + 1000 * S[..., 0].astype(np.float64)
    Q = np.sum(np.where(q > 0, 354800 * q / T, 0.0)) * sim.dx ** 3 / 1000
    assert abs(meta["profile"]["Q_kW"] - round(Q, 1)) <= 0.11


@pytest.mark.gpu
@needs_node
def test_node_spot_scene_matches_the_c_abi_bit_for_bit(tmp_path):
    """setSpots / setFaceFlux through the addon give the C ABI's fields and normalisations."""
    print("node spot scene adjoint8_64 x8", flush=True)
    from firesim import scenes
    name = "adjoint8_64"
    sc = scenes.SCENES[name]
    spots = scenes.scene_spots(name)
    face = scenes.scene_face_flux(name)
    scene = {"res": list(sc["res"]), "size": list(scenes.scene_size(name)), "params": scenes.scene_params(name),
             "spots": spots.tolist(), "face": face.ravel().tolist(), "dt": sc["dt"]}
    (tmp_path / "scene.json").write_text(json.dumps(scene))
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "run_steps.js"), str(tmp_path), "8",
                        str(tmp_path / "scene.json")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    meta = json.loads((tmp_path / "meta.json").read_text())
    sim = scenes.make_scene(name, noise_n=32, use_graph=True)
    assert np.array_equal(np.array(meta["norms"], np.float32), sim.spots()[:, 6].astype(np.float32))
    sim.step(sc["dt"], 8)
    for f, fid in (("vel", F.VEL), ("S", F.S), ("S2", F.S2)):
        a = np.fromfile(tmp_path / f"{f}.bin", np.uint16)
        assert np.array_equal(a, sim.read(fid, "native").ravel()), f


@pytest.mark.gpu
@needs_node
def test_node_sparks_match_the_c_abi_bit_for_bit(tmp_path):
    """The Node Sparks class (rate, params, bursts, step) gives the C ABI's particle array."""
    print("node sparks 32x48x32 x20", flush=True)
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "run_sparks.js"), str(tmp_path), "20"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    meta = json.loads((tmp_path / "meta.json").read_text())
    assert meta["threwOnUnknown"] and meta["alive"] > 50
    sim = F.FireSim(res=(32, 48, 32), size=(1.2, 1.8, 1.2), noise_n=32, use_graph=True)
    sp = F.Sparks(sim, 1024)
    sp.set(rate=300.0, turb=1.2)
    for k in range(20):
        if k == 0:
            sp.burst(n=140, pos=(0, 0.15, 0), speed=3.2, radius=0.12, dur=0.5)
        if k == 7:
            sp.burst(n=40, pos=(0.05, 0.1, 0), speed=5, radius=0.03, dur=0.05, fan=0.9, size=1.4, aim=(0.3, 0, -0.2))
        sim.step(1 / 90)
        sp.step(1 / 90)
    a = np.fromfile(tmp_path / "sparks.bin", np.float32).reshape(1024, 16)
    assert np.array_equal(a.view(np.uint32), sp.read().view(np.uint32))


@pytest.mark.gpu
@needs_node
def test_node_render_matches_the_c_abi_bit_for_bit(tmp_path):
    """The Node Render class (params, orbit camera, depth, jitter, two frames) gives the C ABI's images."""
    print("node render 32x48x32", flush=True)
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "run_render.js"), str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    meta = json.loads((tmp_path / "meta.json").read_text())
    assert meta["threwOnUnknown"] and (meta["vw"], meta["vh"]) == (48, 64)
    logs, _ = F.build_campfire(7)
    sim = F.FireSim(res=(32, 48, 32), size=(0.6, 0.9, 0.6), logs=logs, noise_n=32)
    sim.step(1 / 90, 40)
    rd = F.Render(sim, 96, 128, 0.5)
    rd.set(cubic=1, flameSharp=1, detailAmt=0.3, sootEdge=0.12, lightK=0.6, focus=2.2, lensX=0.01)
    rd.set_camera(F.orbit(yaw=0.35, pitch=0.12, dist=2.2, target=(0, 0.5, 0), fov=38))
    depth = (1.8 + (np.arange(600) % 7) * 0.1).astype(np.float32).reshape(30, 20)
    rd.set_depth(depth)
    for f in range(2):
        rd.frame(sim.time, 5 + f, (0.25, -0.125))
        sim.step(1 / 90)
    for name, want in (("vol", rd.read_vol()), ("aux", rd.read_aux()), ("lights", rd.read_lights())):
        a = np.fromfile(tmp_path / f"{name}.bin", np.float32).reshape(want.shape)
        assert np.array_equal(a.view(np.uint32), want.view(np.uint32)), name
    assert rd.read_vol()[..., 3].min() < 1


@needs_node
def test_node_camera_and_lut_helpers():
    r = _node(f"const m = require({JS!r}); const c = m.orbit({{yaw: 0.35, pitch: 0.12, dist: 2.2, target: [0, 0.5, 0], "
              f"fov: 38}}); console.log(JSON.stringify({{c, lut: Array.from(m.addon.blackbodyLut(16, 500, 2500, 4000)), "
              f"d: Array.from(m.addon.renderParamDefaults())}}))")
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    c = F.orbit(yaw=0.35, pitch=0.12, dist=2.2, target=(0, 0.5, 0), fov=38)
    for k in ("pos", "fwd", "right", "up"):
        np.testing.assert_allclose(d["c"][k], c[k], rtol=0, atol=1e-15)
    assert abs(d["c"]["fovY"] - c["fovY"]) < 1e-15
    np.testing.assert_array_equal(np.array(d["lut"], np.float32), F.blackbody_lut(16, 500, 2500, 4000).ravel())
    np.testing.assert_array_equal(np.array(d["d"], np.float32), F.render_param_defaults())
