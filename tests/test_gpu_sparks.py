"""GPU parity of the spark particles (fs_sparks_*, src/sparks.js) with the oracle.

The campfire on the 64x96x64 test grid is stepped on the GPU; after every step the
oracle receives the GPU's velocity field (so the comparison isolates the spark update),
and both step the same particle set with the same bursts (an ignite-like burst, an
aimed pop, a log-shift burst) and a high base rate that keeps dead slots being reused.
The particle arrays must agree bit for bit (float32 storage and f32 math on both sides,
slots taken in index order on both).
"""
import numpy as np
import pytest

import firesim as F
import fsoracle as O
from helpers import DT, make_pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [F.PREC_FP32, F.PREC_FP16], ids=["fp32", "fp16"])
def test_sparks_bit_exact_with_oracle(prec):
    print(f"sparks parity prec {prec}", flush=True)
    gpu, orc = make_pair(precision=prec, warm_steps=40)
    n = 4096
    sg, so = F.Sparks(gpu, n), O.OracleSparks(orc, n)
    for s in (sg, so):
        s.set(rate=400.0, turb=1.0, lifeK=1.0, sizeK=1.0, upK=1.0)
    events = {0: dict(n=140, pos=(0.0, 0.15, 0.0), speed=3.2, radius=0.12, dur=0.5),
              12: dict(n=60, pos=(0.05, 0.1, -0.02), speed=5.0, radius=0.03, dur=0.05, fan=0.9, size=1.4,
                       aim=(0.3, 0.0, -0.2)),
              25: dict(n=190, pos=(0.0, 0.2, 0.05), speed=4.5, radius=0.12, dur=0.15, fan=0.7, size=1.5)}
    for k in range(60):
        if k in events:
            sg.burst(**events[k])
            so.burst(**events[k])
        gpu.step(DT)
        orc.write(O.VEL, gpu.read(F.VEL, np.float64))
        sg.step(DT)
        so.step(DT)
        if k % 10 == 9 or k == 59:
            a, b = sg.read(), so.read()
            alive = int((a[:, 14] > 0.5).sum())
            diff = np.argwhere(a.view(np.uint32) != b.view(np.uint32))
            print(f"step {k}: alive {alive}, differing values {len(diff)}", flush=True)
            assert len(diff) == 0, (diff[:5], a[diff[0][0]], b[diff[0][0]])
    assert alive > 100
    assert sg.time == pytest.approx(60 * DT)


def test_sparks_reset_and_errors():
    print("sparks reset / errors", flush=True)
    sim = F.FireSim(res=(32, 48, 32), size=(0.6, 0.9, 0.6), noise_n=32)
    sp = F.Sparks(sim, 256)
    sp.set(rate=1000.0)
    sim.step(DT)
    sp.step(DT)
    assert (sp.read()[:, 14] > 0.5).sum() > 0
    sp.reset()
    assert not sp.read().any() and sp.time == 0.0
    with pytest.raises(F.FsError, match="unknown spark parameter"):
        sp.set(nope=1.0)
    with pytest.raises(F.FsError):
        sp.step(0.0)
    with pytest.raises(F.FsError):
        F.Sparks(sim, 0)
