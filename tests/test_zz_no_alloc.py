"""fs_stats_noalloc (the allocation-free form of fs_stats) against fs_stats, through the C ABI.

Two calls in a row on one GPU, the second on a smaller field (its scratch keeps the first call's
partials beyond the second call's blocks), in every storage precision: the results equal
fs_stats's bit for bit.  A field above FS_STATS_NOALLOC_MAX_CELLS is refused with FS_EINVAL.
"""
import ctypes

import numpy as np
import pytest

import firesim as F

pytestmark = pytest.mark.gpu

MAX_CELLS = 1 << 24


def _stats(sim, field, noalloc):
    out = np.full(20, np.nan, np.float64)
    fn = F.lib().fs_stats_noalloc if noalloc else F.lib().fs_stats
    rc = fn(sim._h, field, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return rc, out


@pytest.mark.parametrize("prec", [F.PREC_FP32, F.PREC_FP16, F.PREC_BF16], ids=["fp32", "fp16", "bf16"])
def test_noalloc_stats_equal_fs_stats_on_two_calls_second_smaller(prec):
    logs, _ = F.build_campfire(7)
    for res in ((64, 96, 64), (24, 40, 20)):
        print(f"noalloc stats prec {prec} res {res}", flush=True)
        sim = F.FireSim(res=res, size=(1.2, 1.8, 1.2), logs=logs, precision=prec, noise_n=32)
        sim.step(1.0 / 90.0, 12)
        sim.sync()
        for field in (F.VEL, F.S, F.S2, F.VORT, F.GEO):
            rc1, a = _stats(sim, field, noalloc=True)
            rc2, b = _stats(sim, field, noalloc=False)
            assert rc1 == 0 and rc2 == 0, (rc1, rc2, F.lib().fs_last_error())
            assert np.array_equal(a, b), (res, field, a, b)
        s = _stats(sim, F.S, True)[1]
        assert s[1] > 0.3                       # a flame: max theta of S
        sim.close()


def test_noalloc_stats_refuses_fields_beyond_the_scratch():
    res = (256, 256, 257)
    assert res[0] * res[1] * res[2] > MAX_CELLS
    print(f"noalloc stats size limit res {res}", flush=True)
    sim = F.FireSim(res=res, size=(1.0, 1.0, 1.0 * 257 / 256), logs=np.zeros(0, np.float32), precision=F.PREC_FP16,
                    noise_n=32)
    rc, out = _stats(sim, F.S, True)
    assert rc == -1                         # FS_EINVAL
    assert b"scratch" in F.lib().fs_last_error()
    assert np.all(np.isnan(out))
    sim.close()
