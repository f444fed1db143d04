"""The exact separable inner solve (FS_SOLVE_SPECTRAL64) on the GPU.

The solve is FP32 transforms on the matrix cores (rocBLAS GEMMs) inside the FP64 defect correction, so it is
checked by what it must achieve rather than bit for bit against a restatement:
  * the divergence residual ||b - A p|| / ||b||, recomputed here in float64 from the device's FP64 pressure and
    its kept right-hand side, is <= div_tol after every step, and one correction per step suffices;
  * graph replay gives the same bits as eager launches, step after step (rocBLAS's kernels are recorded into the
    captured graphs);
  * the step's fields stay within the FP64 oracle's tolerance (tests/test_gpu_fp64_error.py, form
    "mixed_spectral").
"""
import numpy as np
import pytest

import firesim as F
from firesim import scenes

pytestmark = pytest.mark.gpu

NAME = "spots4_64"


def residual64(p, b, h):
    """||b - A p|| / ||b|| with the step's operator (face-Dirichlet x/z sides and top, Neumann floor)."""
    p = np.asarray(p, np.float64)
    b = np.asarray(b, np.float64)
    nz, ny, nx = p.shape
    s = np.zeros_like(p)
    diag = np.full_like(p, 6.0)
    for ax, lo_rule in ((2, 1.0), (0, 1.0), (1, -1.0)):        # x, z: Dirichlet both ends; y: Neumann floor
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax], hi[ax] = slice(1, None), slice(None, -1)
        s[tuple(lo)] += p[tuple(hi)]
        s[tuple(hi)] += p[tuple(lo)]
        first = [slice(None)] * 3
        last = [slice(None)] * 3
        first[ax], last[ax] = 0, -1
        diag[tuple(first)] += lo_rule
        diag[tuple(last)] += 1.0
    r = b - (s - diag * p) / (h * h)
    return np.linalg.norm(r) / np.linalg.norm(b)


def make(use_graph, precision=F.PREC_FP16):
    return scenes.make_scene(NAME, precision=precision, solve=F.SOLVE_SPECTRAL64, div_tol=1e-5, max_vcycles=8,
                             noise_n=32, use_graph=use_graph)


@pytest.mark.parametrize("prec", [F.PREC_FP16, F.PREC_FP32], ids=["fp16", "fp32"])
def test_spectral_solve_meets_the_tolerance_every_step(prec):
    print(f"spectral residual {NAME} prec {prec}", flush=True)
    sc = scenes.SCENES[NAME]
    sim = make(True, prec)
    sim.step(sc["dt"], 30)
    used_all = []
    for k in range(8):
        sim.step(sc["dt"], 1)
        rel, used = sim.residual()
        p, b = sim.read(F.P, np.float64), sim.read(F.B, np.float64)
        mine = residual64(p, b, float(np.float32(sc["dx"])))      # the operator's h is the f32 cell size
        used_all.append(used)
        print(f"step {k}: device {rel:.3e}, recomputed {mine:.3e}, corrections {used}", flush=True)
        assert rel <= 1e-5 and mine <= 1.0001e-5, (rel, mine)
        assert abs(mine - rel) <= 1e-3 * rel + 1e-12
    assert max(used_all) <= 2


def test_spectral_graph_replay_equals_eager():
    print("spectral graph vs eager", flush=True)
    sc = scenes.SCENES[NAME]
    a, b = make(True), make(False)
    for sim in (a, b):
        sim.step(sc["dt"], 20)
    for k in range(6):
        for sim in (a, b):
            sim.step(sc["dt"], 1)
        for f in (F.VEL, F.S, F.S2, F.P):
            x, y = a.read(f, "native"), b.read(f, "native")
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (k, f)
        assert a.residual() == b.residual()


FUSED_GRIDS = [(32, 48, 32), (64, 96, 64), (32, 48, 96), (32, 48, 128)]     # nz = 32, 64, 96, 128
LIBRARY_Z_GRIDS = [(32, 48, 40)]                                            # nz = 40: the rocBLAS z pass


@pytest.mark.parametrize("res", FUSED_GRIDS + LIBRARY_Z_GRIDS,
                         ids=["x".join(map(str, r)) for r in FUSED_GRIDS + LIBRARY_Z_GRIDS])
def test_z_pass_meets_the_tolerance_on_every_shape(res):
    """The fused MFMA z pass (z transform, eigenvalue division, inverse z in one kernel; one instantiation per
    nz) on each nz it takes, and the rocBLAS z pass on a shape it does not take: the residual recomputed in
    float64 from the device's pressure stays <= div_tol with one or two corrections per step."""
    print(f"spectral z pass, res {res}", flush=True)
    sc = scenes.SCENES[NAME]
    sim = scenes.make_scene(NAME, res=res, precision=F.PREC_FP32, solve=F.SOLVE_SPECTRAL64, div_tol=1e-5,
                            max_vcycles=8, noise_n=32, use_graph=False)
    sim.step(sc["dt"], 10)
    for k in range(3):
        sim.step(sc["dt"], 1)
        rel, used = sim.residual()
        mine = residual64(sim.read(F.P, np.float64), sim.read(F.B, np.float64), float(np.float32(sc["dx"])))
        print(f"step {k}: device {rel:.3e}, recomputed {mine:.3e}, corrections {used}", flush=True)
        assert rel <= 1e-5 and mine <= 1.0001e-5 and used <= 2, (rel, mine, used)


def test_spectral_mode_errors():
    print("spectral errors", flush=True)
    with pytest.raises(F.FsError, match="even nx"):
        F.FireSim(res=(33, 48, 32), size=(0.33, 0.48, 0.32), solve=F.SOLVE_SPECTRAL64, noise_n=16)
    sim = F.FireSim(res=(32, 48, 32), size=(0.32, 0.48, 0.32), solve=F.SOLVE_SPECTRAL64, precision=F.PREC_FP32,
                    noise_n=16)
    with pytest.raises(F.FsError, match="REFINE64"):
        F.Adjoint(sim, 10)


def test_spectral_mode_needs_the_host_checked_solve():
    """The spectral corrections are library GEMMs that cannot exit on the device's stop flag, so the fully
    asynchronous schedule (max_vcycles corrections per step) is refused at create."""
    print("spectral refuses adaptive = 0", flush=True)
    with pytest.raises(F.FsError, match="adaptive"):
        F.FireSim(res=(32, 48, 32), size=(0.32, 0.48, 0.32), solve=F.SOLVE_SPECTRAL64, adaptive=False, noise_n=16)
