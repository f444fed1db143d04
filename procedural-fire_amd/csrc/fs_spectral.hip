// Exact inner solve of the pressure equation for FS_SOLVE_SPECTRAL64.
//
// The step's pressure operator (src/sim.js:416-437: the 7-point Laplacian with face-Dirichlet x/z sides
// and top and a Neumann floor, ghost-cell rule on the diagonal) has constant coefficients and separates
// by axis.  The 1D second difference with Dirichlet faces at both ends is diagonalised by the DST-II
// basis sin(pi (k+1)(i+1/2)/N); with a Neumann low end and a Dirichlet high end by the DCT-IV basis
// cos(pi (k+1/2)(j+1/2)/N); the eigenvalues are -(2 - 2 cos theta_k) / h^2.  So A^-1 b is three
// orthogonal transforms, one division by the summed eigenvalues and the three inverse transforms:
// six dense GEMMs over 3.1 M cells (5.6 GFLOP at 128x192x128), which run on the matrix cores through
// rocBLAS, and one elementwise kernel.
//
// Layout: the level-0 arrays are parity-split (each x-row holds its even-x cells, then its odd-x
// cells); the x transform's columns are permuted to that storage order, so no data is reordered.
// Seen column-major, a level-0 array is an (nx x ny nz) matrix whose columns are the x-rows:
//   x:  C = Mx B            (Mx(k, c) = T_x[k][x(c)])
//   y:  C_z = A_z My^T      (batched over z: A_z is nx x ny)
//   z:  C = A Mz^T          (A is nx ny x nz)
// and the inverse applies the transposes in the opposite order (the bases are orthonormal).
// In FP32 the solve is accurate to about cond(A) x 2^-24 in the residual norm; the FP64 defect
// correction around it (fs_mg.hip) takes the step's residual to its tolerance.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cmath>
#include <vector>

#include "fs_internal.h"

struct Spectral {
    int nx = 0, ny = 0, nz = 0;
    rocblas_handle blas = nullptr;
    float* Mx = nullptr;      // nx x nx, columns in parity-split storage order
    float* My = nullptr;      // ny x ny
    float* Mz = nullptr;      // nz x nz
    double* lam = nullptr;    // nx + ny + nz eigenvalues of the 1D operators (h = 1)
    float* tmp = nullptr;     // nx ny nz work array
    double ih2 = 1.0;
};

namespace {

__global__ __launch_bounds__(256) void k_spec_scale(float* __restrict__ f, const double* __restrict__ lam, int nx, int ny,
                                                    int nz, double ih2) {
    const int i = (int)(blockIdx.x * 256 + threadIdx.x);
    if (i >= nx * ny * nz) return;
    const int kx = i % nx, ky = (i / nx) % ny, kz = i / (nx * ny);
    const double l = (lam[kx] + lam[nx + ky] + lam[nx + ny + kz]) * ih2;
    f[i] = (float)((double)f[i] / l);
}

void basis(int n, int kind, std::vector<double>& T, std::vector<double>& lam) {
    const double pi = 3.14159265358979323846;
    T.assign((size_t)n * n, 0.0);
    lam.assign(n, 0.0);
    for (int k = 0; k < n; k++) {
        const double th = kind == 0 ? pi * (k + 1) / n : pi * (k + 0.5) / n;
        lam[k] = -(2.0 - 2.0 * std::cos(th));
        const double nrm = (kind == 0 && k == n - 1) ? std::sqrt(1.0 / n) : std::sqrt(2.0 / n);
        for (int i = 0; i < n; i++)
            T[(size_t)k * n + i] = nrm * (kind == 0 ? std::sin(th * (i + 0.5)) : std::cos(th * (i + 0.5)));
    }
}

#define BLAS_CHECK(call)                                                                  \
    do {                                                                                  \
        rocblas_status st_ = (call);                                                      \
        if (st_ != rocblas_status_success) {                                              \
            fs_set_error(std::string(#call) + ": " + rocblas_status_to_string(st_));     \
            return FS_EDEVICE;                                                            \
        }                                                                                 \
    } while (0)

}  // namespace

extern "C" int fs_spectral_basis(int n, int kind, double* T, double* lam) {
    if (n < 1 || (kind != 0 && kind != 1) || !T || !lam) {
        fs_set_error("fs_spectral_basis: n >= 1, kind 0 (DST-II) or 1 (DCT-IV), non-null outputs");
        return FS_EINVAL;
    }
    std::vector<double> t, l;
    basis(n, kind, t, l);
    std::copy(t.begin(), t.end(), T);
    std::copy(l.begin(), l.end(), lam);
    return FS_OK;
}

void spectral_free(Spectral* sp) {
    if (!sp) return;
    void* p[] = {sp->Mx, sp->My, sp->Mz, sp->lam, sp->tmp};
    for (void* q : p) if (q) (void)hipFree(q);
    if (sp->blas) rocblas_destroy_handle(sp->blas);
    delete sp;
}

int spectral_solve(Spectral* sp, hipStream_t s, float* b, float* p) {
    const int nx = sp->nx, ny = sp->ny, nz = sp->nz, R = ny * nz, P = nx * ny;
    const float one = 1.f, zero = 0.f;
    rocblas_handle h = sp->blas;
    BLAS_CHECK(rocblas_set_stream(h, s));
    const rocblas_operation N = rocblas_operation_none, T = rocblas_operation_transpose;
    // forward: x (b -> tmp), y (tmp -> p), z (p -> tmp)
    BLAS_CHECK(rocblas_sgemm(h, N, N, nx, R, nx, &one, sp->Mx, nx, b, nx, &zero, sp->tmp, nx));
    BLAS_CHECK(rocblas_sgemm_strided_batched(h, N, T, nx, ny, ny, &one, sp->tmp, nx, (rocblas_stride)P, sp->My, ny, 0, &zero, p,
                                             nx, (rocblas_stride)P, nz));
    BLAS_CHECK(rocblas_sgemm(h, N, T, P, nz, nz, &one, p, P, sp->Mz, nz, &zero, sp->tmp, P));
    const int n = nx * ny * nz;
    k_spec_scale<<<(n + 255) / 256, 256, 0, s>>>(sp->tmp, sp->lam, nx, ny, nz, sp->ih2);
    // inverse: z (tmp -> p), y (p -> tmp), x (tmp -> p)
    BLAS_CHECK(rocblas_sgemm(h, N, N, P, nz, nz, &one, sp->tmp, P, sp->Mz, nz, &zero, p, P));
    BLAS_CHECK(rocblas_sgemm_strided_batched(h, N, N, nx, ny, ny, &one, p, nx, (rocblas_stride)P, sp->My, ny, 0, &zero, sp->tmp,
                                             nx, (rocblas_stride)P, nz));
    BLAS_CHECK(rocblas_sgemm(h, T, N, nx, R, nx, &one, sp->Mx, nx, sp->tmp, nx, &zero, p, nx));
    FS_HIP_CHECK(hipGetLastError());
    return FS_OK;
}

int spectral_build(Spectral** out, int nx, int ny, int nz, double h, hipStream_t s, float* scratch_b, float* scratch_p,
                   std::string& err) {
    *out = nullptr;
    if (nx % 2 != 0) { err = "the spectral solve needs an even nx (the parity-split rows have no padding then)"; return FS_EINVAL; }
    Spectral* sp = new Spectral();
    sp->nx = nx; sp->ny = ny; sp->nz = nz;
    sp->ih2 = 1.0 / (h * h);
    std::vector<double> Tx, Ty, Tz, lx, ly, lz;
    basis(nx, 0, Tx, lx);
    basis(ny, 1, Ty, ly);
    basis(nz, 0, Tz, lz);
    const int nh = nx / 2;
    std::vector<float> mx((size_t)nx * nx), my((size_t)ny * ny), mz((size_t)nz * nz);
    for (int c = 0; c < nx; c++) {   // storage position c holds x = 2c (c < nh) or 2(c - nh) + 1
        const int x = c < nh ? 2 * c : 2 * (c - nh) + 1;
        for (int k = 0; k < nx; k++) mx[(size_t)c * nx + k] = (float)Tx[(size_t)k * nx + x];
    }
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < ny; k++) my[(size_t)j * ny + k] = (float)Ty[(size_t)k * ny + j];
    for (int j = 0; j < nz; j++)
        for (int k = 0; k < nz; k++) mz[(size_t)j * nz + k] = (float)Tz[(size_t)k * nz + j];
    std::vector<double> lam;
    lam.insert(lam.end(), lx.begin(), lx.end());
    lam.insert(lam.end(), ly.begin(), ly.end());
    lam.insert(lam.end(), lz.begin(), lz.end());
    const size_t n = (size_t)nx * ny * nz;
    bool ok = hipMalloc(&sp->Mx, mx.size() * 4) == hipSuccess && hipMalloc(&sp->My, my.size() * 4) == hipSuccess &&
              hipMalloc(&sp->Mz, mz.size() * 4) == hipSuccess && hipMalloc(&sp->lam, lam.size() * 8) == hipSuccess &&
              hipMalloc(&sp->tmp, n * 4) == hipSuccess;
    if (!ok) { spectral_free(sp); err = "hipMalloc failed for the spectral solve"; return FS_ENOMEM; }
    ok = hipMemcpy(sp->Mx, mx.data(), mx.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(sp->My, my.data(), my.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(sp->Mz, mz.data(), mz.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(sp->lam, lam.data(), lam.size() * 8, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) { spectral_free(sp); err = "device copy failed for the spectral solve"; return FS_EDEVICE; }
    if (rocblas_create_handle(&sp->blas) != rocblas_status_success) {
        spectral_free(sp);
        err = "rocblas_create_handle failed";
        return FS_EDEVICE;
    }
    // one eager solve: rocBLAS loads its kernels and sizes its workspace here, not inside a captured graph
    launch_fill_zero(s, scratch_b, n * 4);
    const int rc = spectral_solve(sp, s, scratch_b, scratch_p);
    if (rc != FS_OK || hipStreamSynchronize(s) != hipSuccess) {
        spectral_free(sp);
        err = "the first spectral solve failed";
        return FS_EDEVICE;
    }
    *out = sp;
    return FS_OK;
}
