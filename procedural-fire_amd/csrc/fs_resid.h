// Cell functions of the multigrid's level-0 operator shared by the multigrid kernels and the projection kernel
// that also forms the refinement's closing residual: the FP64 residual at the two cells of a parity-split pair,
// and the first colour-0 update of a level whose iterate is zero.
#pragma once
#include <hip/hip_runtime.h>

#include "fs_codec.h"
#include "fs_internal.h"

// The colour-0 Gauss-Seidel update of cell (x, y, z) on a level whose iterate is all
// zero: every neighbour contributes +0, so the update needs only the cell's own b.  The
// kernels that zero a level (restriction, the refinement's residual) write it directly;
// the operations are k_smooth's, in its order, so the values are identical.
__device__ __forceinline__ float red_from_zero(float bv, int x, int y, int z, int nx, int ny, int nz, float h, float omega) {
    float diag = 6.f;
    if (x == 0) diag += 1.f;
    if (x == nx - 1) diag += 1.f;
    if (z == 0) diag += 1.f;
    if (z == nz - 1) diag += 1.f;
    if (y == 0) diag -= 1.f;
    if (y == ny - 1) diag += 1.f;
    const float sum = 0.f;
    float pn = (sum - h * h * bv) / diag;
    if (omega != 1.f) {
        const float po = 0.f;
        pn = po + omega * (pn - po);
    }
    return pn;
}

// A level-0 array of the multigrid read through a buffer resource: 32-bit element offsets, no 64-bit address
// arithmetic per load (the array stays below 2 GiB).
template <class X> struct LevelBuf;
template <> struct LevelBuf<double> {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ LevelBuf(const double* p, int n)
        : r(__builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, n * 8, FS_BUF_WORD3)) {}
    __device__ __forceinline__ double operator()(int i) const {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, i * 8, 0, 0));
    }
};
template <> struct LevelBuf<float> {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ LevelBuf(const float* p, int n)
        : r(__builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, n * 4, FS_BUF_WORD3)) {}
    __device__ __forceinline__ double operator()(int i) const {
        return (double)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, i * 4, 0, 0));
    }
};

// The iterate around the parity-split pair of cells x = 2 hx (even half-row, slot o0) and 2 hx + 1 (odd half-row,
// slot o1) of row (y, z): each cell's own value, its outer x neighbour (the pair's cells are each other's inner
// one) and its y and z neighbours, twelve loads for the two seven-point stencils.  A neighbour outside the grid is
// loaded from the cell itself (unused), so every load issues at once.
struct PairStencil {
    double m[2], w0, e1, k[2], f[2], s[2], n[2];
    int o0, o1;
};
template <class L>
__device__ __forceinline__ PairStencil pair_stencil(const L& x, int hx, int y, int z, int nx, int ny, int nz, int nh, int rs) {
    PairStencil p;
    const int row = (z * ny + y) * rs, dz = ny * rs;
    p.o0 = row + hx;
    p.o1 = row + nh + hx;
    const int xa = 2 * hx;
    const int o[2] = {p.o0, p.o1};
#pragma unroll
    for (int c = 0; c < 2; c++) {
        p.m[c] = x(o[c]);
        p.k[c] = x(fs_opaque(z > 0 ? o[c] - dz : o[c]));
        p.f[c] = x(fs_opaque(z < nz - 1 ? o[c] + dz : o[c]));
        p.s[c] = x(fs_opaque(y > 0 ? o[c] - rs : o[c]));
        p.n[c] = x(fs_opaque(y < ny - 1 ? o[c] + rs : o[c]));
    }
    p.w0 = x(fs_opaque(xa > 0 ? p.o1 - 1 : p.o0));          // x - 1 of cell 0: the odd cell hx - 1
    p.e1 = x(fs_opaque(xa + 2 < nx ? p.o0 + 1 : p.o1));     // x + 1 of cell 1: the even cell hx + 1
    return p;
}
// The FP64 residual bv - A x of the level-0 operator (src/sim.js:416-437 ghost rules: -p on the x/z faces and the
// top, +p below the floor) at cell c of the pair, from its stencil (c = 1 of a padded row's odd slot: in-bounds
// values, a result to be dropped).
__device__ __forceinline__ double pair_resid(const PairStencil& p, int c, double bv, int hx, int y, int z, int nx, int ny,
                                             int nz, double ih2) {
    const int xx = 2 * hx + c;
    const double vw = c == 0 ? p.w0 : p.m[0], ve = c == 0 ? p.m[1] : p.e1;
    // the operator's `if (inside) sum += v; else diag += 1` (-= 1 below the floor) in its order: a face outside
    // adds -0.0, which leaves every sum as it is (+0 included), and diag is an exact small integer
    const bool bw = xx > 0, be = xx < nx - 1, bk = z > 0, bf = z < nz - 1, bs = y > 0, bn = y < ny - 1;
    double sum = 0.0;
    sum += bw ? vw : -0.0;
    sum += be ? ve : -0.0;
    sum += bk ? p.k[c] : -0.0;
    sum += bf ? p.f[c] : -0.0;
    sum += bs ? p.s[c] : -0.0;
    sum += bn ? p.n[c] : -0.0;
    const double diag = (double)(6 + !bw + !be + !bk + !bf - !bs + !bn);
    return bv - (sum - diag * p.m[c]) * ih2;
}

// The sum of v over the 64 lanes of a wave, in a fixed butterfly order; lane 0's value is reproducible.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
