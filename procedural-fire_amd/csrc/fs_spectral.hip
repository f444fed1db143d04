// Exact inner solve of the pressure equation for FS_SOLVE_SPECTRAL64.
//
// The step's pressure operator (src/sim.js:416-437: the 7-point Laplacian with face-Dirichlet x/z sides
// and top and a Neumann floor, ghost-cell rule on the diagonal) has constant coefficients and separates
// by axis.  The 1D second difference with Dirichlet faces at both ends is diagonalised by the DST-II
// basis sin(pi (k+1)(i+1/2)/N); with a Neumann low end and a Dirichlet high end by the DCT-IV basis
// cos(pi (k+1/2)(j+1/2)/N); the eigenvalues are -(2 - 2 cos theta_k) / h^2.  So A^-1 b is three
// orthogonal transforms, one division by the summed eigenvalues and the three inverse transforms:
// six dense products over 3.1 M cells (5.6 GFLOP at 128x192x128), all on the matrix cores: the x and y
// transforms through rocBLAS, the z transform, the division and the inverse z transform in one fused
// kernel (k_spec_z; rocBLAS and an elementwise kernel for shapes it does not take).
//
// Layout: the level-0 arrays are parity-split (each x-row holds its even-x cells, then its odd-x
// cells); the x transform's columns are permuted to that storage order, so no data is reordered.
// Seen column-major, a level-0 array is an (nx x ny nz) matrix whose columns are the x-rows:
//   x:  C = Mx B            (Mx(k, c) = T_x[k][x(c)]; run as (Mx^T)^T B, which rocBLAS tiles better at nx = 128)
//   y:  C_z = A_z My^T      (batched over z: A_z is nx x ny)
//   z:  C = A Mz^T          (A is nx ny x nz)
// and the inverse applies the transposes in the opposite order (the bases are orthonormal).
// In FP32 the solve is accurate to about cond(A) x 2^-24 in the residual norm; the FP64 defect
// correction around it (fs_mg.hip) takes the step's residual to its tolerance.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "fs_internal.h"

struct Spectral {
    int nx = 0, ny = 0, nz = 0;
    bool fused_z = false;     // the z pass runs as one fused kernel (nz = 32, 64, 96 or 128; nx ny = 32 m)
    int zgrid = 0;            // its blocks: one per CU, at most one per panel
    rocblas_handle blas = nullptr;
    float* Mx = nullptr;      // nx x nx, columns in parity-split storage order
    float* MxT = nullptr;     // its transpose (the forward x product runs as a transposed-A GEMM, like the inverse)
    float* My = nullptr;      // ny x ny
    float* Mz = nullptr;      // nz x nz
    float* Tz = nullptr;      // T_z row-major (the fused z pass)
    double* lam = nullptr;    // nx + ny + nz eigenvalues of the 1D operators (h = 1)
    float* tmp = nullptr;     // nx ny nz work array
    double ih2 = 1.0;
};

namespace {

// The z transform, the division by the eigenvalues and the inverse z transform, fused, over panels of 32
// columns (kx, ky) of the array seen as nz rows z of nx ny columns.  Persistent: each block (one per CU)
// walks its panels, prefetching the next panel into registers while the current one is on the matrix
// cores.  v_mfma_f32_16x16x4f32: lane l holds A[l % 16][l / 16] and B[l / 16][l % 16], and gets
// D[4 (l / 16) + j][l % 16].  NRB = nz / 16 waves; wave w owns the panel's columns 16 (w & 1) .. +15 and
// row blocks 2 (w >> 1), +1, so every SIMD gets the same work.  A wave's T_z fragments (its two row
// blocks of T_z for the forward product and of T_z^T for the inverse, 4 nz floats a lane) are read from
// LDS once and stay in registers, so each k-step reads one panel value from LDS.  The forward result is
// divided in registers and written to LDS over the panel; the inverse result is stored from registers.
typedef float f4v __attribute__((ext_vector_type(4)));
constexpr int ZC = 32;            // panel width (columns)
constexpr int ZLDP = 48;          // panel row stride in LDS (rows 4 apart start 16 banks apart)
template <int NRB>
__global__ __launch_bounds__(NRB * 64) void k_spec_z(float* __restrict__ f, const float* __restrict__ Tz,
                                                     const double* __restrict__ lam, int nx, int ny, int npanels,
                                                     double ih2) {
    constexpr int nz = NRB * 16, ldm = nz + 4, NT = NRB * 64, KS = nz / 4;
    constexpr int NM = nz * nz / 4 / NT, NV = nz * ZC / 4 / NT;   // float4 per thread: T_z, one panel
    extern __shared__ float4 sm4[];
    float* sM = (float*)sm4;          // sM[k ldm + z] = T_z[k][z]
    float* sP = sM + nz * ldm;        // sP[z ZLDP + c]: the panel, then the divided forward result
    const int P = nx * ny, t = threadIdx.x;
    {
        float4 m[NM];
#pragma unroll
        for (int q = 0; q < NM; q++) m[q] = ((const float4*)Tz)[q * NT + t];
#pragma unroll
        for (int q = 0; q < NM; q++) {
            const int i = (q * NT + t) * 4;
            *(float4*)&sM[(i / nz) * ldm + i % nz] = m[q];
        }
    }
    const int w = t >> 6, l = t & 63, lr = l & 15, lk = l >> 4;
    const int col = (w & 1) * 16 + lr, r0 = (w >> 1) * 32;   // first row of the wave's two row blocks
    float4 pre[NV];
    int pnl = blockIdx.x;
    if (pnl < npanels)
#pragma unroll
        for (int q = 0; q < NV; q++) {
            const int v = q * NT + t;
            pre[q] = *(const float4*)&f[(size_t)(v / (ZC / 4)) * P + pnl * ZC + (v % (ZC / 4)) * 4];
        }
    double lz[2][4];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int j = 0; j < 4; j++) lz[h][j] = lam[nx + ny + r0 + h * 16 + 4 * lk + j];
    __syncthreads();
    float fa[2][KS], ia[2][KS];       // forward A = T_z[r][k], inverse A = T_z[k][r], r the row, k = 4 s + lk
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int k = 0; k < KS; k++) {
            fa[h][k] = sM[(r0 + h * 16 + lr) * ldm + 4 * k + lk];
            ia[h][k] = sM[(4 * k + lk) * ldm + r0 + h * 16 + lr];
        }
    for (; pnl < npanels; pnl += gridDim.x) {
        const int c0 = pnl * ZC;
        __syncthreads();              // the previous panel's inverse is done reading sP
#pragma unroll
        for (int q = 0; q < NV; q++) {
            const int v = q * NT + t;
            *(float4*)&sP[(v / (ZC / 4)) * ZLDP + (v % (ZC / 4)) * 4] = pre[q];
        }
        __syncthreads();
        const int nxt = pnl + gridDim.x;
        if (nxt < npanels)
#pragma unroll
            for (int q = 0; q < NV; q++) {
                const int v = q * NT + t;
                pre[q] = *(const float4*)&f[(size_t)(v / (ZC / 4)) * P + nxt * ZC + (v % (ZC / 4)) * 4];
            }
        const int kxy = c0 + col;
        const double lxy = lam[kxy % nx] + lam[nx + kxy / nx];
        // forward: C[kz][c] = sum_z T_z[kz][z] panel[z][c]
        f4v a0 = f4v{0.f, 0.f, 0.f, 0.f}, a1 = a0;
#pragma unroll
        for (int k = 0; k < KS; k++) {
            const float bv = sP[(4 * k + lk) * ZLDP + col];
            a0 = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[0][k], bv, a0, 0, 0, 0);
            a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[1][k], bv, a1, 0, 0, 0);
        }
        __syncthreads();              // every wave is done reading the panel
#pragma unroll
        for (int j = 0; j < 4; j++) {
            sP[(r0 + 4 * lk + j) * ZLDP + col] = (float)((double)a0[j] / ((lxy + lz[0][j]) * ih2));
            sP[(r0 + 16 + 4 * lk + j) * ZLDP + col] = (float)((double)a1[j] / ((lxy + lz[1][j]) * ih2));
        }
        __syncthreads();
        // inverse: out[z][c] = sum_kz T_z[kz][z] C[kz][c]
        a0 = f4v{0.f, 0.f, 0.f, 0.f};
        a1 = a0;
#pragma unroll
        for (int k = 0; k < KS; k++) {
            const float bv = sP[(4 * k + lk) * ZLDP + col];
            a0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ia[0][k], bv, a0, 0, 0, 0);
            a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ia[1][k], bv, a1, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            f[(size_t)(r0 + 4 * lk + j) * P + c0 + col] = a0[j];
            f[(size_t)(r0 + 16 + 4 * lk + j) * P + c0 + col] = a1[j];
        }
    }
}

template <int NRB>
constexpr size_t spec_z_lds() { return (size_t)(NRB * 16 * (NRB * 16 + 4) + NRB * 16 * ZLDP) * sizeof(float); }

template <int NRB>
void spec_z_t(hipStream_t s, int grid, float* f, const float* Tz, const double* lam, int nx, int ny, double ih2) {
    k_spec_z<NRB><<<grid, NRB * 64, spec_z_lds<NRB>(), s>>>(f, Tz, lam, nx, ny, nx * ny / ZC, ih2);
}

// Allows the fused z pass its LDS (above the 64 KiB default for nz >= 96); false if the device refuses.
bool spec_z_prepare(int nrb) {
    hipError_t e = hipErrorInvalidValue;
    switch (nrb) {
        case 2: e = hipFuncSetAttribute((const void*)k_spec_z<2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)spec_z_lds<2>()); break;
        case 4: e = hipFuncSetAttribute((const void*)k_spec_z<4>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)spec_z_lds<4>()); break;
        case 6: e = hipFuncSetAttribute((const void*)k_spec_z<6>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)spec_z_lds<6>()); break;
        case 8: e = hipFuncSetAttribute((const void*)k_spec_z<8>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)spec_z_lds<8>()); break;
        default: break;
    }
    if (e != hipSuccess) (void)hipGetLastError();
    return e == hipSuccess;
}

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
    void* p[] = {sp->Mx, sp->MxT, sp->My, sp->Mz, sp->Tz, sp->lam, sp->tmp};
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
    BLAS_CHECK(rocblas_sgemm(h, T, N, nx, R, nx, &one, sp->MxT, nx, b, nx, &zero, sp->tmp, nx));
    BLAS_CHECK(rocblas_sgemm_strided_batched(h, N, T, nx, ny, ny, &one, sp->tmp, nx, (rocblas_stride)P, sp->My, ny, 0, &zero, p,
                                             nx, (rocblas_stride)P, nz));
    if (sp->fused_z) {   // z, division and inverse z in place on p, then y (p -> tmp), x (tmp -> p)
        switch (nz / 16) {
            case 2: spec_z_t<2>(s, sp->zgrid, p, sp->Tz, sp->lam, nx, ny, sp->ih2); break;
            case 4: spec_z_t<4>(s, sp->zgrid, p, sp->Tz, sp->lam, nx, ny, sp->ih2); break;
            case 6: spec_z_t<6>(s, sp->zgrid, p, sp->Tz, sp->lam, nx, ny, sp->ih2); break;
            default: spec_z_t<8>(s, sp->zgrid, p, sp->Tz, sp->lam, nx, ny, sp->ih2); break;
        }
        BLAS_CHECK(rocblas_sgemm_strided_batched(h, N, N, nx, ny, ny, &one, p, nx, (rocblas_stride)P, sp->My, ny, 0, &zero,
                                                 sp->tmp, nx, (rocblas_stride)P, nz));
        BLAS_CHECK(rocblas_sgemm(h, T, N, nx, R, nx, &one, sp->Mx, nx, sp->tmp, nx, &zero, p, nx));
        FS_HIP_CHECK(hipGetLastError());
        return FS_OK;
    }
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
    std::vector<float> mx((size_t)nx * nx), mxt((size_t)nx * nx), my((size_t)ny * ny), mz((size_t)nz * nz);
    for (int c = 0; c < nx; c++) {   // storage position c holds x = 2c (c < nh) or 2(c - nh) + 1
        const int x = c < nh ? 2 * c : 2 * (c - nh) + 1;
        for (int k = 0; k < nx; k++) mx[(size_t)c * nx + k] = mxt[(size_t)k * nx + c] = (float)Tx[(size_t)k * nx + x];
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
    std::vector<float> tz(Tz.begin(), Tz.end());
    const int nrb = nz / 16;
    int dev = 0, ncu = 0;
    sp->fused_z = nz % 32 == 0 && nrb >= 2 && nrb <= 8 && (nx * ny) % ZC == 0 &&
                  hipGetDevice(&dev) == hipSuccess &&
                  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0 &&
                  spec_z_prepare(nrb);
    sp->zgrid = std::min(ncu, nx * ny / ZC);
    bool ok = hipMalloc(&sp->Mx, mx.size() * 4) == hipSuccess && hipMalloc(&sp->MxT, mxt.size() * 4) == hipSuccess &&
              hipMalloc(&sp->My, my.size() * 4) == hipSuccess &&
              hipMalloc(&sp->Mz, mz.size() * 4) == hipSuccess && hipMalloc(&sp->Tz, tz.size() * 4) == hipSuccess &&
              hipMalloc(&sp->lam, lam.size() * 8) == hipSuccess && hipMalloc(&sp->tmp, n * 4) == hipSuccess;
    if (!ok) { spectral_free(sp); err = "hipMalloc failed for the spectral solve"; return FS_ENOMEM; }
    ok = hipMemcpy(sp->Mx, mx.data(), mx.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(sp->MxT, mxt.data(), mxt.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(sp->My, my.data(), my.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(sp->Mz, mz.data(), mz.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(sp->Tz, tz.data(), tz.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
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
