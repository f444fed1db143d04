// Exact solve of one coarse multigrid level by separable transforms (the tolerance solves' coarse-grid
// solve in place of the V-cycle below that level).
//
// The level's operator is the step's 7-point Laplacian with face-Dirichlet x/z sides and top and a
// Neumann floor (src/sim.js:416-437 at that level's h).  It separates by axis: the DST-II basis
// sin(pi (k+1)(i+1/2)/n) diagonalises the 1D second difference with Dirichlet faces at both ends, the
// DCT-IV basis cos(pi (k+1/2)(j+1/2)/n) the one with a Neumann low end and a Dirichlet high end, with
// eigenvalues -(2 - 2 cos theta_k).  So p = A^-1 b is
//   W = Tx b Tz^T per y-plane,   W <- Ty^T diag(1 / lambda) Ty W per (kx, kz) column,   p = Tx^T W Tz,
// three launches: forward x and z transforms per y-plane, the y transform with the division and its
// inverse per column chunk, the inverse z and x transforms per y-plane.
//
// Each output value is one FMA chain in a fixed order over the contraction index (fma(l, r, acc) from
// acc = 0), so a CPU restatement with the same chains gives the same bits.  Products are tiled through
// LDS with both factors laid out along the contraction index (transposed bases are prepared on the
// host), rows padded so that a wave's float4 reads do not collide in the banks; a thread owns a 2 x 4
// (plane kernels) or 2 x 2 (y kernel) block of outputs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "fs_internal.h"

struct DirectSolve {
    int X = 0, Y = 0, Z = 0;      // level dimensions
    int sx = 1, sz = 1;           // kx-row splits of the forward plane kernel, z-column splits of the inverse
    float ih2 = 1.f;              // 1 / h^2 of the level (f32)
    float* Tx = nullptr;          // X x X row-major (row k = mode k), f32; TxT its transpose
    float* TxT = nullptr;
    float* Ty = nullptr;          // Y x Y, TyT
    float* TyT = nullptr;
    float* Tz = nullptr;          // Z x Z, TzT
    float* TzT = nullptr;
    float* lam = nullptr;         // X + Y + Z eigenvalues of the 1D operators (h = 1), f32
    float* W = nullptr;           // X Y Z work array: kx fastest, then y, then kz
};

namespace {

constexpr int DT = 256;           // threads per block of the plane kernels
constexpr int DTY = 384;          // threads per block of the y kernel
constexpr int YC = 16;            // columns (kx) per block of the y kernel

// Padded LDS row length: rows of n floats (n a multiple of 4) start 4 floats apart in the banks, so a
// wave's float4 reads of 8 consecutive rows touch distinct 16-byte bank groups.
__host__ __device__ constexpr int ldp(int n) { return n + 4; }

// n / d for n < 2^16, d < 2^16 by one multiply-high with m = ceil(2^32 / d) (exact in that range)
struct Div {
    unsigned m;
    int d;
};
Div make_div(int d) { return Div{(unsigned)((((unsigned long long)1 << 32) + (unsigned)d - 1) / (unsigned)d), d}; }
__device__ __forceinline__ int dv(int n, Div f) { return (int)__umulhi((unsigned)n, f.m); }

// out[i][j] = sum_k A[i lda + k] B[j ldb + k] over i < M, j < N, k < K in increasing k (one FMA chain per
// output, from 0): both factors are read along k, four steps at a time as float4.  A thread owns
// TI x TJ outputs (M a multiple of TI, N of TJ, K of 4); store(i0, j0, acc[TI][TJ]).
// ti = M / TI as a Div (the tile index splits without a division).
template <int TI, int TJ, int NT, class Store>
__device__ __forceinline__ void lds_gemm_nt(const float* __restrict__ A, int lda, const float* __restrict__ B, int ldb,
                                            Div ti, int N, int K, Store store) {
    const int nt = ti.d * (N / TJ);
    for (int t = threadIdx.x; t < nt; t += NT) {
        const int tj = dv(t, ti);
        const int i0 = (t - tj * ti.d) * TI, j0 = tj * TJ;
        float acc[TI][TJ];
#pragma unroll
        for (int a = 0; a < TI; a++)
#pragma unroll
            for (int c = 0; c < TJ; c++) acc[a][c] = 0.f;
#pragma unroll 2
        for (int k = 0; k < K; k += 4) {
            float4 av[TI], bv[TJ];
#pragma unroll
            for (int a = 0; a < TI; a++) av[a] = *(const float4*)&A[(i0 + a) * lda + k];
#pragma unroll
            for (int c = 0; c < TJ; c++) bv[c] = *(const float4*)&B[(j0 + c) * ldb + k];
#pragma unroll
            for (int a = 0; a < TI; a++)
#pragma unroll
                for (int c = 0; c < TJ; c++) {
                    float v = acc[a][c];
                    v = __builtin_fmaf(av[a].x, bv[c].x, v);
                    v = __builtin_fmaf(av[a].y, bv[c].y, v);
                    v = __builtin_fmaf(av[a].z, bv[c].z, v);
                    v = __builtin_fmaf(av[a].w, bv[c].w, v);
                    acc[a][c] = v;
                }
        }
        store(i0, j0, acc);
    }
}

// Global -> LDS copies in two phases: a thread first loads up to NV float4 into registers (every load in
// flight at once), and stores them to LDS later.  load(i) and store(i, v) map the flat float4 index i.
constexpr int NV = 8;
template <int NT, class Load>
__device__ __forceinline__ void gather(float4 (&v)[NV], int total, Load load) {
#pragma unroll
    for (int u = 0; u < NV; u++) {
        const int i = u * NT + (int)threadIdx.x;
        if (i < total) v[u] = load(i);
    }
}
template <int NT, class Store>
__device__ __forceinline__ void scatter(const float4 (&v)[NV], int total, Store store) {
#pragma unroll
    for (int u = 0; u < NV; u++) {
        const int i = u * NT + (int)threadIdx.x;
        if (i < total) store(i, v[u]);
    }
}
// float4 i of rows r0 .. of a row-major n-column matrix (q = n / 4 as a Div), and its place in padded
// LDS rows
__device__ __forceinline__ float4 row4(const float* __restrict__ src, int r0, int n, Div q, int i) {
    const int r = dv(i, q), c = (i - r * q.d) * 4;
    return *(const float4*)&src[(r0 + r) * n + c];
}
__device__ __forceinline__ void put_row4(float* dst, int n, Div q, int i, float4 v) {
    const int r = dv(i, q), c = (i - r * q.d) * 4;
    *(float4*)&dst[r * ldp(n) + c] = v;
}
// The loaded registers are consumed here, so every load is issued before the skip test waits on them.
__device__ __forceinline__ void pin(const float4 (&v)[NV]) {
#pragma unroll
    for (int u = 0; u < NV; u++) asm volatile("" ::"v"(v[u].x), "v"(v[u].y), "v"(v[u].z), "v"(v[u].w));
}

// Forward x and z transforms of one y-plane, rows kx in [r0, r0 + R):
//   C[kx][z] = sum_x Tx[kx][x] b[x][z];   W[kz][y][kx] = sum_z C[kx][z] Tz[kz][z].
// LDS: sT[r][x] = Tx[r0 + r] (R rows), sB[z][x] = b (Z rows; later sB[kz][z] = Tz), sC[r][z].
__global__ __launch_bounds__(DT) void k_direct_fwd(const float* __restrict__ b, const float* __restrict__ Tx,
                                                   const float* __restrict__ Tz, float* __restrict__ W, int X, int Y, int Z,
                                                   int R, Div qx, Div qz, Div tr, const int* __restrict__ skip) {
    const int sk = skip[fs_opaque(0)];
    extern __shared__ float4 sh4[];
    float* sT = (float*)sh4;
    float* sB = sT + R * ldp(X);
    float* sC = sB + Z * ldp(std::max(X, Z));
    const int y = blockIdx.x, r0 = blockIdx.y * R, nh = X >> 1;
    float4 vt[NV], vb[NV], vz[NV];
    gather<DT>(vt, R * qx.d, [&](int i) { return row4(Tx, r0, X, qx, i); });
    // b is parity-split: row (z Y + y) holds the even x, then the odd x
    gather<DT>(vb, Z * qx.d, [&](int i) { const int z = dv(i, qx); return *(const float4*)&b[(z * Y + y) * X + (i - z * qx.d) * 4]; });
    gather<DT>(vz, Z * qz.d, [&](int i) { return row4(Tz, 0, Z, qz, i); });
    pin(vt); pin(vb); pin(vz);
    asm volatile("" ::"v"(sk));
    if (sk) return;                                    // uniform over the grid
    scatter<DT>(vt, R * qx.d, [&](int i, float4 v) { put_row4(sT, X, qx, i, v); });
    scatter<DT>(vb, Z * qx.d, [&](int i, float4 v) {
        const int z = dv(i, qx), g = (i - z * qx.d) * 4;
        const float e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int gg = g + u, x = gg < nh ? 2 * gg : 2 * (gg - nh) + 1;
            sB[z * ldp(X) + x] = e[u];
        }
    });
    __syncthreads();
    lds_gemm_nt<2, 4, DT>(sT, ldp(X), sB, ldp(X), tr, Z, X, [&](int i0, int j0, float (&acc)[2][4]) {
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int c = 0; c < 4; c++) sC[(i0 + a) * ldp(Z) + j0 + c] = acc[a][c];
    });
    __syncthreads();
    scatter<DT>(vz, Z * qz.d, [&](int i, float4 v) { put_row4(sB, Z, qz, i, v); });   // sB now holds Tz
    __syncthreads();
    lds_gemm_nt<2, 4, DT>(sC, ldp(Z), sB, ldp(Z), tr, Z, Z, [&](int i0, int j0, float (&acc)[2][4]) {
#pragma unroll
        for (int c = 0; c < 4; c++) {
            float* o = W + ((size_t)(j0 + c) * Y + y) * X + r0 + i0;
#pragma unroll
            for (int a = 0; a < 2; a++) o[a] = acc[a][c];
        }
    });
}

// The y transform, the division by the eigenvalues and the inverse y transform of YC columns kx at one kz:
//   F[ky][c] = sum_y Ty[ky][y] V[c][y];  G = F / ((lx + ly + lz) ih2);  W[kz][y][c] = sum_ky Ty[ky][y] G[ky][c].
// LDS: sT = Ty (later TyT: sT[y][ky]), sV[c][y], sG[c][ky].
__global__ __launch_bounds__(DTY) void k_direct_y(float* __restrict__ W, const float* __restrict__ Ty,
                                                 const float* __restrict__ TyT, const float* __restrict__ lam, int X, int Y,
                                                 int Z, float ih2, Div qy, Div ty, const int* __restrict__ skip) {
    const int sk = skip[fs_opaque(0)];
    extern __shared__ float4 sh4[];
    float* sT = (float*)sh4;
    float* sV = sT + Y * ldp(Y);
    float* sG = sV + YC * ldp(Y);
    const int c0 = blockIdx.x * YC, kz = blockIdx.y;
    float4 vt[NV], vtt[NV], vv[NV];
    gather<DTY>(vt, Y * qy.d, [&](int i) { return row4(Ty, 0, Y, qy, i); });
    gather<DTY>(vtt, Y * qy.d, [&](int i) { return row4(TyT, 0, Y, qy, i); });
    gather<DTY>(vv, Y * (YC >> 2), [&](int i) {
        const int yy = i >> 2;                           // YC / 4 = 4 float4 per row
        return *(const float4*)&W[((size_t)kz * Y + yy) * X + c0 + (i & 3) * 4];
    });
    pin(vt); pin(vtt); pin(vv);
    asm volatile("" ::"v"(sk));
    if (sk) return;
    scatter<DTY>(vt, Y * qy.d, [&](int i, float4 v) { put_row4(sT, Y, qy, i, v); });
    scatter<DTY>(vv, Y * (YC >> 2), [&](int i, float4 v) {
        const int yy = i >> 2, c = (i & 3) * 4;
        sV[(c + 0) * ldp(Y) + yy] = v.x;
        sV[(c + 1) * ldp(Y) + yy] = v.y;
        sV[(c + 2) * ldp(Y) + yy] = v.z;
        sV[(c + 3) * ldp(Y) + yy] = v.w;
    });
    __syncthreads();
    const float* lx = lam;
    const float* ly = lam + X;
    const float lzv = lam[X + Y + kz];
    lds_gemm_nt<2, 2, DTY>(sT, ldp(Y), sV, ldp(Y), ty, YC, Y, [&](int i0, int j0, float (&acc)[2][2]) {
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const float d = ((lx[c0 + j0 + c] + ly[i0 + a]) + lzv) * ih2;
                sG[(j0 + c) * ldp(Y) + i0 + a] = acc[a][c] / d;
            }
    });
    __syncthreads();
    scatter<DTY>(vtt, Y * qy.d, [&](int i, float4 v) { put_row4(sT, Y, qy, i, v); });   // sT now holds TyT
    __syncthreads();
    lds_gemm_nt<2, 2, DTY>(sT, ldp(Y), sG, ldp(Y), ty, YC, Y, [&](int i0, int j0, float (&acc)[2][2]) {
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int c = 0; c < 2; c++) W[((size_t)kz * Y + i0 + a) * X + c0 + j0 + c] = acc[a][c];
    });
}

// Inverse z and x transforms of one y-plane, columns z in [z0, z0 + R):
//   Q[kx][z] = sum_kz W[kz][y][kx] Tz[kz][z];   p[x][y][z] = sum_kx Tx[kx][x] Q[kx][z].
// LDS: sW[kx][kz] (X rows; later sW[x][kx] = TxT), sM[r][kz] = TzT[z0 + r] (R rows), sQ[r][kx].
__global__ __launch_bounds__(DT) void k_direct_inv(const float* __restrict__ W, const float* __restrict__ TxT,
                                                   const float* __restrict__ TzT, float* __restrict__ p, int X, int Y, int Z,
                                                   int R, Div qx, Div qz, Div tx, const int* __restrict__ skip) {
    const int sk = skip[fs_opaque(0)];
    extern __shared__ float4 sh4[];
    float* sW = (float*)sh4;
    float* sM = sW + X * ldp(std::max(X, Z));
    float* sQ = sM + R * ldp(Z);
    const int y = blockIdx.x, z0 = blockIdx.y * R, nh = X >> 1;
    float4 vw[NV], vm[NV], vx[NV];
    gather<DT>(vw, Z * qx.d, [&](int i) { const int kz = dv(i, qx); return *(const float4*)&W[((size_t)kz * Y + y) * X + (i - kz * qx.d) * 4]; });
    gather<DT>(vm, R * qz.d, [&](int i) { return row4(TzT, z0, Z, qz, i); });
    gather<DT>(vx, X * qx.d, [&](int i) { return row4(TxT, 0, X, qx, i); });
    pin(vw); pin(vm); pin(vx);
    asm volatile("" ::"v"(sk));
    if (sk) return;
    scatter<DT>(vw, Z * qx.d, [&](int i, float4 v) {
        const int kz = dv(i, qx), kx = (i - kz * qx.d) * 4;
        sW[(kx + 0) * ldp(Z) + kz] = v.x;
        sW[(kx + 1) * ldp(Z) + kz] = v.y;
        sW[(kx + 2) * ldp(Z) + kz] = v.z;
        sW[(kx + 3) * ldp(Z) + kz] = v.w;
    });
    scatter<DT>(vm, R * qz.d, [&](int i, float4 v) { put_row4(sM, Z, qz, i, v); });
    __syncthreads();
    lds_gemm_nt<2, 4, DT>(sW, ldp(Z), sM, ldp(Z), tx, R, Z, [&](int i0, int j0, float (&acc)[2][4]) {
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int c = 0; c < 4; c++) sQ[(j0 + c) * ldp(X) + i0 + a] = acc[a][c];
    });
    __syncthreads();
    scatter<DT>(vx, X * qx.d, [&](int i, float4 v) { put_row4(sW, X, qx, i, v); });   // sW now holds TxT
    __syncthreads();
    lds_gemm_nt<2, 4, DT>(sW, ldp(X), sQ, ldp(X), tx, R, X, [&](int i0, int j0, float (&acc)[2][4]) {
#pragma unroll
        for (int a = 0; a < 2; a++) {
            const int x = i0 + a;
#pragma unroll
            for (int c = 0; c < 4; c++) p[((z0 + j0 + c) * Y + y) * X + (x & 1) * nh + (x >> 1)] = acc[a][c];
        }
    });
}

size_t fwd_lds(const DirectSolve& d, int R) {
    return sizeof(float) * ((size_t)R * ldp(d.X) + (size_t)d.Z * ldp(std::max(d.X, d.Z)) + (size_t)R * ldp(d.Z));
}
size_t y_lds(const DirectSolve& d) { return sizeof(float) * ((size_t)d.Y * ldp(d.Y) + 2 * (size_t)YC * ldp(d.Y)); }
size_t inv_lds(const DirectSolve& d, int R) {
    return sizeof(float) * ((size_t)d.X * ldp(std::max(d.X, d.Z)) + (size_t)R * ldp(d.Z) + (size_t)R * ldp(d.X));
}

constexpr size_t LDS_MAX = 160 * 1024;

void basis_f32(int n, int kind, float* T, float* TT, float* lam) {   // fs_spectral_basis rounded to f32, and T^T
    std::vector<double> t((size_t)n * n), l(n);
    fs_spectral_basis(n, kind, t.data(), l.data());
    for (int k = 0; k < n; k++)
        for (int i = 0; i < n; i++) {
            T[(size_t)k * n + i] = (float)t[(size_t)k * n + i];
            TT[(size_t)i * n + k] = (float)t[(size_t)k * n + i];
        }
    for (int k = 0; k < n; k++) lam[k] = (float)l[k];
}

}  // namespace

bool direct_fits(int X, int Y, int Z) {
    // x in chunks of YC, y and z in steps of 4 (the products read float4 along each contraction)
    if (X < YC || Y < 4 || Z < 8 || X % YC || Y % 4 || Z % 8 || X > 128 || Z > 128 || Y > 128) return false;
    if ((size_t)X * Y * Z > ((size_t)1 << 19)) return false;
    // every global -> LDS copy in one batch of NV float4 per thread
    if (X * Z > 4 * NV * DT || X * X > 4 * NV * DT || Z * Z > 4 * NV * DT || Y * Y > 4 * NV * DTY) return false;
    DirectSolve d;
    d.X = X; d.Y = Y; d.Z = Z;
    return fwd_lds(d, X / 2) <= LDS_MAX && y_lds(d) <= LDS_MAX && inv_lds(d, Z / 2) <= LDS_MAX;
}

int direct_build(DirectSolve** out, int X, int Y, int Z, float h, std::string& err) {
    *out = nullptr;
    if (!direct_fits(X, Y, Z)) { err = "the direct coarse solve does not take this level's shape"; return FS_EINVAL; }
    DirectSolve* d = new DirectSolve;
    d->X = X; d->Y = Y; d->Z = Z;
    d->ih2 = 1.f / (h * h);
    // each plane's work in two blocks (rows of the forward product, columns of the inverse); the halves
    // keep the 2 x 4 output tiles whole
    d->sx = 2;
    d->sz = 2;
    std::vector<float> tx((size_t)X * X), txt(tx.size()), ty((size_t)Y * Y), tyt(ty.size()), tz((size_t)Z * Z),
        tzt(tz.size()), lam(X + Y + Z);
    basis_f32(X, 0, tx.data(), txt.data(), lam.data());
    basis_f32(Y, 1, ty.data(), tyt.data(), lam.data() + X);
    basis_f32(Z, 0, tz.data(), tzt.data(), lam.data() + X + Y);
    bool ok = true;
    const std::pair<float**, const std::vector<float>*> up[] = {{&d->Tx, &tx}, {&d->TxT, &txt}, {&d->Ty, &ty},
                                                              {&d->TyT, &tyt}, {&d->Tz, &tz},   {&d->TzT, &tzt},
                                                              {&d->lam, &lam}};
    for (const auto& u : up)
        ok = ok && hipMalloc(u.first, u.second->size() * 4) == hipSuccess &&
             hipMemcpy(*u.first, u.second->data(), u.second->size() * 4, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMalloc(&d->W, (size_t)X * Y * Z * 4) == hipSuccess;
    // LDS above the 64 KiB default where a kernel needs it
    ok = ok && hipFuncSetAttribute((const void*)k_direct_fwd, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)fwd_lds(*d, X / d->sx)) == hipSuccess;
    ok = ok && hipFuncSetAttribute((const void*)k_direct_y, hipFuncAttributeMaxDynamicSharedMemorySize, (int)y_lds(*d)) ==
                   hipSuccess;
    ok = ok && hipFuncSetAttribute((const void*)k_direct_inv, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)inv_lds(*d, Z / d->sz)) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        direct_free(d);
        err = "device setup of the direct coarse solve failed";
        return FS_EDEVICE;
    }
    *out = d;
    return FS_OK;
}

void direct_free(DirectSolve* d) {
    if (!d) return;
    void* p[] = {d->Tx, d->TxT, d->Ty, d->TyT, d->Tz, d->TzT, d->lam, d->W};
    for (void* q : p) if (q) (void)hipFree(q);
    delete d;
}

void direct_solve(const DirectSolve* d, hipStream_t s, const float* b, float* p, const int* skip) {
    const int X = d->X, Y = d->Y, Z = d->Z;
    const int Rf = X / d->sx, Ri = Z / d->sz;
    const Div qx = make_div(X >> 2), qy = make_div(Y >> 2), qz = make_div(Z >> 2);
    k_direct_fwd<<<dim3(Y, d->sx), DT, fwd_lds(*d, Rf), s>>>(b, d->Tx, d->Tz, d->W, X, Y, Z, Rf, qx, qz, make_div(Rf / 2),
                                                             skip);
    k_direct_y<<<dim3(X / YC, Z), DTY, y_lds(*d), s>>>(d->W, d->Ty, d->TyT, d->lam, X, Y, Z, d->ih2, qy, make_div(Y / 2),
                                                       skip);
    k_direct_inv<<<dim3(Y, d->sz), DT, inv_lds(*d, Ri), s>>>(d->W, d->TxT, d->TzT, p, X, Y, Z, Ri, qx, qz, make_div(X / 2),
                                                             skip);
}
