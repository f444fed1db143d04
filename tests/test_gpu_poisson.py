"""Isolated pressure Poisson solve on the GPU (fs_poisson_*): the sim.js multigrid against the
oracle (bit-exact V-cycles), the FP64-refined solve against the closed-form DST/DCT solution,
and odd grid sizes that exercise the parity-split layout's row padding."""
import numpy as np
import pytest

import firesim as F
import fsoracle as O
from test_oracle_pins import _spectral_solve

pytestmark = pytest.mark.gpu


def _rhs(dims, seed=0):
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    b = np.sin(0.3 * x + 0.2) * np.cos(0.21 * y) * np.sin(0.17 * z + 1.0) + 0.05 * rng.standard_normal((nz, ny, nx))
    return b


# 96^3 keeps three levels (96, 48, 24) outside the single-workgroup tail, so its level 2 takes the
# prolongation fused into the first post-smoothing pass; the smaller grids reach the tail by level 2.
@pytest.mark.parametrize("omega", [1.0, 1.25])
@pytest.mark.parametrize("dims", [(32, 48, 32), (64, 96, 64), (20, 30, 22), (21, 30, 22), (12, 18, 10), (96, 96, 96)])
def test_vcycles_bit_exact_vs_oracle(dims, omega):
    print(f"poisson vcycles {dims} omega {omega}", flush=True)
    h = 1.0 / 128          # exactly representable: the device keeps the spacing as f32
    b = _rhs(dims)
    ps = F.Poisson(dims, h, omega=omega)
    ps.set_rhs(b)
    ps.set_x(np.zeros_like(b))
    hist = ps.vcycles(4)
    orc = O.OracleMG(dims, h, fp64=False, omega=omega)
    orc.set(b=b, p=np.zeros_like(b))
    ohist = orc.vcycles(4)
    assert ps.info()[0] == orc.levels()[0]
    assert np.array_equal(ps.read_x(fp64=False), orc.get().astype(np.float32))
    assert np.allclose(hist, ohist, rtol=1e-9)


@pytest.mark.parametrize("omega", [1.0, 1.25])
@pytest.mark.parametrize("dims", [(32, 48, 32), (64, 96, 64), (24, 36, 20)])
def test_refined_solve_reaches_closed_form(dims, omega):
    print(f"poisson refine {dims} omega {omega}", flush=True)
    h = 1.0 / 256
    b = _rhs(dims, 1)
    b -= b.mean()
    exact = _spectral_solve(b, h)
    ps = F.Poisson(dims, h, omega=omega)
    ps.set_rhs(b)
    ps.set_x(np.zeros_like(b))
    hist = ps.solve(1e-10, 40)
    assert hist[-1] <= 1e-10 and len(hist) < 40
    x = ps.read_x()
    assert np.max(np.abs(x - exact)) <= 1e-8 * np.max(np.abs(exact))
    # same cycle count and iterate as the oracle's FP32-in-FP64 refinement
    orc = O.OracleMG(dims, h, fp64=False, omega=omega)
    orc.set(b=b, p=np.zeros_like(b))
    ohist = orc.refine(1e-10, 40)
    assert len(ohist) == len(hist) - 1
    assert np.max(np.abs(x - orc.get_x())) <= 1e-12 * np.max(np.abs(exact))


def test_blob_workload_refined_to_1e8_matches_fp64_oracle():
    """The isolated-solve workload (seeded Gaussian blobs + noise, zero mean) at 64x96x64:
    FP32 V-cycles (omega 1.25) in FP64 refinement to 1e-8, against the FP64 oracle converged
    to ~1e-14; same cycle count as the oracle's own FP32-in-FP64 refinement."""
    from firesim import scenes
    print("poisson blobs 64x96x64", flush=True)
    dims = (64, 96, 64)
    h = 1.0 / 256
    b = scenes.blob_rhs(dims)
    ps = F.Poisson(dims, h, omega=1.25)
    ps.set_rhs(b)
    ps.set_x(np.zeros_like(b))
    hist = ps.solve(1e-8, 40)
    assert hist[-1] <= 1e-8 and len(hist) - 1 <= 8, hist
    x = ps.read_x()
    ref = O.OracleMG(dims, h, fp64=True, omega=1.25)
    ref.set(b=b, p=np.zeros_like(b))
    rh = ref.vcycles(30)
    assert rh[-1] < 1e-12
    xr = ref.get()
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= 1e-7
    orc = O.OracleMG(dims, h, fp64=False, omega=1.25)
    orc.set(b=b, p=np.zeros_like(b))
    ohist = orc.refine(1e-8, 40)
    assert len(ohist) == len(hist) - 1
