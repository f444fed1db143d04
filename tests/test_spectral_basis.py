"""The bases of the exact pressure solve (FS_SOLVE_SPECTRAL64), pinned on the CPU against the operator itself.

The step's pressure operator (src/sim.js:416-437, restated by the multigrid's residual) is the 7-point Laplacian
whose diagonal gains +1 at an x/z side or the top (Dirichlet face: ghost = -p) and -1 at the floor (Neumann:
ghost = p).  Each 1D factor must be diagonalised by the library's basis (DST-II for Dirichlet-Dirichlet, DCT-IV
for Neumann-Dirichlet) with the library's eigenvalues, and the separable solve assembled from them must equal a
dense solve of the 3D operator.
"""
import numpy as np
import pytest

import firesim as F


def op1d(n, low_neumann):
    L = np.diag(np.full(n, -2.0)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)
    L[0, 0] += 1.0 if low_neumann else -1.0       # ghost = p (Neumann) or -p (Dirichlet)
    L[-1, -1] -= 1.0                               # Dirichlet high end
    return L


@pytest.mark.parametrize("n", [2, 5, 8, 12, 31, 64])
@pytest.mark.parametrize("kind", [0, 1])
def test_basis_is_orthonormal_and_diagonalises_the_1d_operator(n, kind):
    T, lam = F.spectral_basis(n, kind)
    np.testing.assert_allclose(T @ T.T, np.eye(n), atol=1e-13)
    D = T @ op1d(n, kind == 1) @ T.T
    np.testing.assert_allclose(D, np.diag(lam), atol=1e-12)
    assert np.all(lam < 0)                         # no null mode: every axis has a Dirichlet face


def op3d(nx, ny, nz, h):
    """The multigrid residual's operator as a dense matrix, x fastest."""
    n = nx * ny * nz
    A = np.zeros((n, n))
    idx = lambda x, y, z: (z * ny + y) * nx + x
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                i = idx(x, y, z)
                diag = 6.0
                for (xx, yy, zz, bd) in ((x - 1, y, z, 1.0), (x + 1, y, z, 1.0), (x, y, z - 1, 1.0), (x, y, z + 1, 1.0),
                                         (x, y - 1, z, -1.0), (x, y + 1, z, 1.0)):
                    if 0 <= xx < nx and 0 <= yy < ny and 0 <= zz < nz:
                        A[i, idx(xx, yy, zz)] += 1.0
                    else:
                        diag += bd
                A[i, i] -= diag
    return A / (h * h)


def test_separable_solve_equals_the_dense_solve():
    nx, ny, nz, h = 6, 5, 4, 0.01
    A = op3d(nx, ny, nz, h)
    rng = np.random.default_rng(2)
    b = rng.standard_normal(nx * ny * nz)
    Tx, lx = F.spectral_basis(nx, 0)
    Ty, ly = F.spectral_basis(ny, 1)
    Tz, lz = F.spectral_basis(nz, 0)
    B = b.reshape(nz, ny, nx)
    Bh = np.einsum("ax,by,cz,zyx->cba", Tx, Ty, Tz, B)
    lam = (lz[:, None, None] + ly[None, :, None] + lx[None, None, :]) / (h * h)
    X = np.einsum("ax,by,cz,cba->zyx", Tx, Ty, Tz, Bh / lam)
    np.testing.assert_allclose(X.ravel(), np.linalg.solve(A, b), rtol=1e-10, atol=1e-14)


def test_bad_arguments_are_rejected():
    with pytest.raises(F.FsError):
        F.spectral_basis(0, 0)
    with pytest.raises(F.FsError):
        F.spectral_basis(4, 2)
