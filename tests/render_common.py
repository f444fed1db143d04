"""Shared pieces of the render-volume tests: the uniform block as uploadParams fills it
(src/renderer.js:634-648), written out independently of the library, and simple scenes."""
import math

import numpy as np

RP_FIELDS = [
    'width', 'height', 'vw', 'vh', 'time', 'frame', 'nLogs', 'nStones',
    'camX', 'camY', 'camZ', 'tanFov', 'fwdX', 'fwdY', 'fwdZ', 'aspect', 'rgtX', 'rgtY', 'rgtZ', 'jitX', 'upX', 'upY',
    'upZ', 'jitY', 'ox', 'oy', 'oz', 'dx', 'sx', 'sy', 'sz', 'tempScale', 'nx', 'ny', 'nz', 'nLights',
    'exposure', 'stepMul', 'sootAbs', 'sootEmit', 'emitK', 'blueK', 'smokeScat', 'smokeG',
    'ambient', 'detailAmt', 'detailFreq', 'detailSpeed', 'hazeK', 'glowK', 'bloomK', 'tonemap',
    'grain', 'vignette', 'saturation', 'contrast', 'bedGlow', 'logGlow', 'lightK', 'accumN',
    'bbTmin', 'bbTmax', 'bbN', 'emberFlicker', 'sootEmitMin', 'flameSharp', 'moon', 'fogK',
    'shadowSoft', 'lift', 'emitTempLo', 'glowTemp', 'inset', 'knee', 'surfDetail', 'matA', 'matB', 'matC', 'sootScat',
    'chroma', 'lensX', 'lensY', 'focus', 'lightDt', 'flickerK', 'cubic', 'aperture', 'sootEdge',
]
RP = {n: i for i, n in enumerate(RP_FIELDS)}
RP_DEFAULTS = dict(
    exposure=15.0, stepMul=0.7, sootAbs=4.0, sootEmit=10.0, emitK=4.0, blueK=0.05, smokeScat=0.02, smokeG=0.5,
    ambient=1.0, detailAmt=0.6, detailFreq=2.2, detailSpeed=0.9, hazeK=1.0, glowK=1.0, bloomK=0.04, tonemap=2,
    grain=0.02, vignette=0.6, saturation=1.0, contrast=1.0, bedGlow=1.0, logGlow=1.0, lightK=1.0,
    emberFlicker=1.0, sootEmitMin=0.0, flameSharp=0.0, moon=1.0, fogK=1.0, shadowSoft=6.0, lift=0.0,
    emitTempLo=900, glowTemp=1.0, tempScale=1000, inset=0.06, knee=0.6, surfDetail=1, matA=0, matB=0, matC=0,
    sootScat=0.015, chroma=0.0025, lensX=0, lensY=0, focus=0, lightDt=-1, flickerK=0.35, cubic=0, aperture=0,
    sootEdge=0,
)


def vol_size(width, height, vol_scale):
    """resize (src/renderer.js:575-580): vw, vh with Math.round (half up)."""
    w, h = max(16, int(math.floor(width))), max(16, int(math.floor(height)))
    return w, h, max(8, math.floor(w * vol_scale + 0.5)), max(8, math.floor(h * vol_scale + 0.5))


def make_uniforms(cam, res, size, params=None, width=128, height=96, vol_scale=0.5, time=0.0, frame=0,
                  jitter=(0.0, 0.0), light_dt=-1.0, n_logs=0, bb=(400.0, 3000.0, 1024)):
    """The 88-float block of uploadParams for a grid `res` of physical `size` (origin at
    (-sx/2, 0, -sz/2)), a look_at camera dict and this.params overrides."""
    w, h, vw, vh = vol_size(width, height, vol_scale)
    p = dict(RP_DEFAULTS)
    p.update(params or {})
    p.update(lightDt=light_dt, width=w, height=h, vw=vw, vh=vh, time=time, frame=frame, nLogs=n_logs, nStones=0,
             nLights=8, camX=cam['pos'][0], camY=cam['pos'][1], camZ=cam['pos'][2], fwdX=cam['fwd'][0],
             fwdY=cam['fwd'][1], fwdZ=cam['fwd'][2], rgtX=cam['right'][0], rgtY=cam['right'][1],
             rgtZ=cam['right'][2], upX=cam['up'][0], upY=cam['up'][1], upZ=cam['up'][2],
             tanFov=math.tan(cam['fovY'] / 2), aspect=w / h, jitX=jitter[0], jitY=jitter[1],
             ox=-float(np.float32(size[0])) / 2, oy=0.0, oz=-float(np.float32(size[2])) / 2,
             dx=float(np.float32(size[0])) / res[0], sx=size[0], sy=size[1], sz=size[2],
             nx=res[0], ny=res[1], nz=res[2], bbTmin=bb[0], bbTmax=bb[1], bbN=bb[2])
    U = np.zeros(len(RP_FIELDS), np.float32)
    for k, v in p.items():
        U[RP[k]] = np.float32(v)
    return U


def axis_camera(x, y, z0, fov_deg=20.0):
    """A look_at camera at (x, y, z0) looking along +z (right = -x, up = +y)."""
    return dict(pos=[x, y, z0], fwd=[0.0, 0.0, 1.0], right=[-1.0, 0.0, 0.0], up=[0.0, 1.0, 0.0],
                fovY=fov_deg * math.pi / 180)


def bb_interp(lut, T, row, tmin=400.0, tmax=3000.0):
    """bbRow in float64 (the LUT lerp)."""
    n = lut.shape[1]
    x = min(max((T - tmin) / (tmax - tmin), 0.0), 1.0) * (n - 1)
    i = int(math.floor(x))
    f = x - i
    a, b = lut[row, i, :3].astype(np.float64), lut[row, min(i + 1, n - 1), :3].astype(np.float64)
    return a * (1 - f) + b * f
