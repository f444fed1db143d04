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
// acc = 0), so a CPU restatement with the same chains gives the same bits.  The products run on the
// matrix cores (v_mfma_f32_16x16x4_f32, which adds its four products to the accumulator in index order
// with one rounding each: the same chain) from LDS, where both factors lie along the contraction index
// (transposed bases are prepared on the host) in rows padded against bank conflicts; each wave takes
// 16 x 16 output tiles.  A plane kernel block takes 16 rows (or columns) of one y-plane, so the
// level's 64 x 96 x 64 shape runs 384 blocks of each plane kernel and 256 of the y kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "fs_internal.h"

struct DirectSolve {
    int X = 0, Y = 0, Z = 0;      // level dimensions
    int rf = 16, ri = 16;         // kx rows per block of the forward plane kernel, z columns per block of the inverse
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

typedef float v4f __attribute__((ext_vector_type(4)));
__host__ __device__ constexpr int rup16(int n) { return (n + 15) & ~15; }

// out[i][j] = sum_k A[i lda + k] B[j ldb + k] for i < M, j < N: one FMA chain per output over k = 0, 1, ...,
// K - 1 from 0 (K a multiple of 4).  v_mfma_f32_16x16x4_f32 adds A[i][k..k+3] B[j][k..k+3] to the
// accumulator in k order, one rounding per product, so a tile's outputs equal the scalar chains bit for
// bit.  Lane l supplies A[i0 + l % 16][k + l / 16] and B[j0 + l % 16][k + l / 16] and receives
// out[i0 + 4 (l / 16) + r][j0 + l % 16] in acc[r].  A and B hold rows up to M and N rounded up to 16
// (rows past M or N feed only outputs that are dropped).  The block's waves take the tiles in turn;
// store(ib, j, acc) gets rows ib .. ib + 3 of column j < N (rows >= M are for the caller to drop).
template <int NT, class Store>
__device__ __forceinline__ void mfma_nt(const float* A, int lda, const float* B, int ldb, int M, int N, int K, Store store) {
    const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    const int tm = rup16(M) >> 4, nt = tm * (rup16(N) >> 4);
    for (int t = threadIdx.x >> 6; t < nt; t += NT / 64) {
        const int tj = t / tm, i0 = (t - tj * tm) * 16, j0 = tj * 16;
        const float* a = A + (i0 + li) * lda + lg;
        const float* b = B + (j0 + li) * ldb + lg;
        v4f acc = {0.f, 0.f, 0.f, 0.f};
        int k = 0;
        for (; k + 16 <= K; k += 16) {                 // four steps' operands read ahead of their products
            float av[4], bv[4];
#pragma unroll
            for (int u = 0; u < 4; u++) { av[u] = a[k + 4 * u]; bv[u] = b[k + 4 * u]; }
#pragma unroll
            for (int u = 0; u < 4; u++) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u], bv[u], acc, 0, 0, 0);
        }
        for (; k < K; k += 4) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[k], b[k], acc, 0, 0, 0);
        if (j0 + li < N) store(i0 + 4 * lg, j0 + li, acc);
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
                                                   int R, Div qx, Div qz, const int* __restrict__ skip) {
    const int sk = skip[fs_opaque(0)];
    extern __shared__ float4 sh4[];
    float* sT = (float*)sh4;
    float* sB = sT + rup16(R) * ldp(X);
    float* sC = sB + rup16(Z) * ldp(std::max(X, Z));
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
    mfma_nt<DT>(sT, ldp(X), sB, ldp(X), R, Z, X, [&](int ib, int j, v4f acc) {
#pragma unroll
        for (int r = 0; r < 4; r++)
            if (ib + r < R) sC[(ib + r) * ldp(Z) + j] = acc[r];
    });
    __syncthreads();
    scatter<DT>(vz, Z * qz.d, [&](int i, float4 v) { put_row4(sB, Z, qz, i, v); });   // sB now holds Tz
    __syncthreads();
    mfma_nt<DT>(sC, ldp(Z), sB, ldp(Z), R, Z, Z, [&](int ib, int j, v4f acc) {
        float* o = W + ((size_t)j * Y + y) * X + r0 + ib;    // rows kx = r0 + ib .. + 3, contiguous
        if (ib + 3 < R) {
            *(float4*)o = make_float4(acc[0], acc[1], acc[2], acc[3]);
        } else {
#pragma unroll
            for (int r = 0; r < 4; r++)
                if (ib + r < R) o[r] = acc[r];
        }
    });
}

// The y transform, the division by the eigenvalues and the inverse y transform of YC columns kx at one kz:
//   F[ky][c] = sum_y Ty[ky][y] V[c][y];  G = F / ((lx + ly + lz) ih2);  W[kz][y][c] = sum_ky Ty[ky][y] G[ky][c].
// LDS: sT = Ty (later TyT: sT[y][ky]), sV[c][y], sG[c][ky], sL = the X + Y + Z eigenvalues.
__global__ __launch_bounds__(DTY) void k_direct_y(float* __restrict__ W, const float* __restrict__ Ty,
                                                 const float* __restrict__ TyT, const float* __restrict__ lam, int X, int Y,
                                                 int Z, float ih2, Div qy, const int* __restrict__ skip) {
    const int sk = skip[fs_opaque(0)];
    extern __shared__ float4 sh4[];
    float* sT = (float*)sh4;
    float* sV = sT + rup16(Y) * ldp(Y);
    float* sG = sV + YC * ldp(Y);
    float* sL = sG + YC * ldp(Y);
    const int c0 = blockIdx.x * YC, kz = blockIdx.y;
    const int nl = X + Y + Z;                          // <= DTY (each dimension <= 128)
    float4 vt[NV], vtt[NV], vv[NV];
    gather<DTY>(vt, Y * qy.d, [&](int i) { return row4(Ty, 0, Y, qy, i); });
    gather<DTY>(vtt, Y * qy.d, [&](int i) { return row4(TyT, 0, Y, qy, i); });
    gather<DTY>(vv, Y * (YC >> 2), [&](int i) {
        const int yy = i >> 2;                           // YC / 4 = 4 float4 per row
        return *(const float4*)&W[((size_t)kz * Y + yy) * X + c0 + (i & 3) * 4];
    });
    const float lv = (int)threadIdx.x < nl ? lam[threadIdx.x] : 0.f;
    pin(vt); pin(vtt); pin(vv);
    asm volatile("" ::"v"(sk), "v"(lv));
    if (sk) return;
    scatter<DTY>(vt, Y * qy.d, [&](int i, float4 v) { put_row4(sT, Y, qy, i, v); });
    scatter<DTY>(vv, Y * (YC >> 2), [&](int i, float4 v) {
        const int yy = i >> 2, c = (i & 3) * 4;
        sV[(c + 0) * ldp(Y) + yy] = v.x;
        sV[(c + 1) * ldp(Y) + yy] = v.y;
        sV[(c + 2) * ldp(Y) + yy] = v.z;
        sV[(c + 3) * ldp(Y) + yy] = v.w;
    });
    if ((int)threadIdx.x < nl) sL[threadIdx.x] = lv;
    __syncthreads();
    const float lzv = sL[X + Y + kz];
    mfma_nt<DTY>(sT, ldp(Y), sV, ldp(Y), Y, YC, Y, [&](int ib, int j, v4f acc) {
        const float lx = sL[c0 + j];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int ky = ib + r;
            if (ky < Y) {
                const float d = ((lx + sL[X + ky]) + lzv) * ih2;
                sG[j * ldp(Y) + ky] = acc[r] / d;
            }
        }
    });
    __syncthreads();
    scatter<DTY>(vtt, Y * qy.d, [&](int i, float4 v) { put_row4(sT, Y, qy, i, v); });   // sT now holds TyT
    __syncthreads();
    mfma_nt<DTY>(sT, ldp(Y), sG, ldp(Y), Y, YC, Y, [&](int ib, int j, v4f acc) {
#pragma unroll
        for (int r = 0; r < 4; r++)
            if (ib + r < Y) W[((size_t)kz * Y + ib + r) * X + c0 + j] = acc[r];
    });
}

// Inverse z and x transforms of one y-plane, columns z in [z0, z0 + R):
//   Q[kx][z] = sum_kz W[kz][y][kx] Tz[kz][z];   p[x][y][z] = sum_kx Tx[kx][x] Q[kx][z].
// LDS: sW[kx][kz] (X rows; later sW[x][kx] = TxT), sM[r][kz] = TzT[z0 + r] (R rows), sQ[r][kx].  The last
// product is formed as p^T (rows z, columns x): the same chains, with a wave's stores along x.
__global__ __launch_bounds__(DT) void k_direct_inv(const float* __restrict__ W, const float* __restrict__ TxT,
                                                   const float* __restrict__ TzT, float* __restrict__ p, int X, int Y, int Z,
                                                   int R, Div qx, Div qz, const int* __restrict__ skip) {
    const int sk = skip[fs_opaque(0)];
    extern __shared__ float4 sh4[];
    float* sW = (float*)sh4;
    float* sM = sW + X * ldp(std::max(X, Z));
    float* sQ = sM + rup16(R) * ldp(Z);
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
    mfma_nt<DT>(sW, ldp(Z), sM, ldp(Z), X, R, Z, [&](int ib, int j, v4f acc) {
#pragma unroll
        for (int r = 0; r < 4; r++) sQ[j * ldp(X) + ib + r] = acc[r];   // X is a multiple of 16
    });
    __syncthreads();
    scatter<DT>(vx, X * qx.d, [&](int i, float4 v) { put_row4(sW, X, qx, i, v); });   // sW now holds TxT
    __syncthreads();
    mfma_nt<DT>(sQ, ldp(X), sW, ldp(X), R, X, X, [&](int ib, int x, v4f acc) {
        float* o = p + (x & 1) * nh + (x >> 1);
#pragma unroll
        for (int r = 0; r < 4; r++)
            if (ib + r < R) o[((size_t)(z0 + ib + r) * Y + y) * X] = acc[r];
    });
}

size_t fwd_lds(const DirectSolve& d, int R) {
    return sizeof(float) * ((size_t)rup16(R) * ldp(d.X) + (size_t)rup16(d.Z) * ldp(std::max(d.X, d.Z)) +
                            (size_t)rup16(R) * ldp(d.Z));
}
size_t y_lds(const DirectSolve& d) {
    return sizeof(float) * ((size_t)rup16(d.Y) * ldp(d.Y) + 2 * (size_t)YC * ldp(d.Y) + (size_t)(d.X + d.Y + d.Z));
}
size_t inv_lds(const DirectSolve& d, int R) {
    return sizeof(float) * ((size_t)d.X * ldp(std::max(d.X, d.Z)) + (size_t)rup16(R) * ldp(d.Z) + (size_t)rup16(R) * ldp(d.X));
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

// z columns per block of the inverse plane kernel: one output tile wide where Z allows
static int inv_cols(int Z) { return Z % 16 == 0 ? 16 : 8; }

bool direct_fits(int X, int Y, int Z) {
    // x in chunks of YC, y and z in steps of 4 (the matrix-core products take the contraction four at a time)
    if (X < YC || Y < 4 || Z < 8 || X % YC || Y % 4 || Z % 8 || X > 128 || Z > 128 || Y > 128) return false;
    if ((size_t)X * Y * Z > ((size_t)1 << 19)) return false;
    // every global -> LDS copy in one batch of NV float4 per thread
    if (X * Z > 4 * NV * DT || X * X > 4 * NV * DT || Z * Z > 4 * NV * DT || Y * Y > 4 * NV * DTY) return false;
    DirectSolve d;
    d.X = X; d.Y = Y; d.Z = Z;
    return fwd_lds(d, 16) <= LDS_MAX && y_lds(d) <= LDS_MAX && inv_lds(d, inv_cols(Z)) <= LDS_MAX;
}

int direct_build(DirectSolve** out, int X, int Y, int Z, float h, std::string& err) {
    *out = nullptr;
    if (!direct_fits(X, Y, Z)) { err = "the direct coarse solve does not take this level's shape"; return FS_EINVAL; }
    DirectSolve* d = new DirectSolve;
    d->X = X; d->Y = Y; d->Z = Z;
    d->ih2 = 1.f / (h * h);
    // each plane's work in blocks of 16 rows of the forward product and 16 (or 8) columns of the inverse:
    // one output tile per wave and product, and X / 16 (Z / 16) blocks per plane to fill the chip
    d->rf = 16;
    d->ri = inv_cols(Z);
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
                                   (int)fwd_lds(*d, d->rf)) == hipSuccess;
    ok = ok && hipFuncSetAttribute((const void*)k_direct_y, hipFuncAttributeMaxDynamicSharedMemorySize, (int)y_lds(*d)) ==
                   hipSuccess;
    ok = ok && hipFuncSetAttribute((const void*)k_direct_inv, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)inv_lds(*d, d->ri)) == hipSuccess;
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
    const int Rf = d->rf, Ri = d->ri;
    const Div qx = make_div(X >> 2), qy = make_div(Y >> 2), qz = make_div(Z >> 2);
    k_direct_fwd<<<dim3(Y, X / Rf), DT, fwd_lds(*d, Rf), s>>>(b, d->Tx, d->Tz, d->W, X, Y, Z, Rf, qx, qz, skip);
    k_direct_y<<<dim3(X / YC, Z), DTY, y_lds(*d), s>>>(d->W, d->Ty, d->TyT, d->lam, X, Y, Z, d->ih2, qy, skip);
    k_direct_inv<<<dim3(Y, Z / Ri), DT, inv_lds(*d, Ri), s>>>(d->W, d->TxT, d->TzT, p, X, Y, Z, Ri, qx, qz, skip);
}
