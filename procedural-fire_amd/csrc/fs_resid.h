// Cell functions of the multigrid's level-0 operator shared by the multigrid kernels and the divergence
// kernel that also forms the tolerance solve's starting residual: the FP64 residual at one cell of the
// parity-split iterate, and the first colour-0 update of a level whose iterate is zero.
#pragma once
#include <hip/hip_runtime.h>

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

// FP64 residual bv - A x of the level-0 operator (src/sim.js:416-437 ghost rules: -p on the x/z faces and
// the top, +p below the floor) at cell xx = 2 hx + sgn of row (y, z); x is parity-split with half-row nh and
// row stride rs.  Every neighbour is loaded (an outside one from the cell itself, unused) so the loads issue
// together.
template <class X>
__device__ __forceinline__ double resid_at(const X* __restrict__ x, double bv, int xx, int y, int z, int sgn, int hx, int nx,
                                           int ny, int nz, int nh, int rs, double ih2) {
    const int row = (z * ny + y) * rs;
    const int own = row + sgn * nh + hx, oth = row + (1 - sgn) * nh;
    const double me = (double)x[own];
    const double vw = (double)x[fs_opaque(xx > 0 ? oth + hx - 1 + sgn : own)];
    const double ve = (double)x[fs_opaque(xx < nx - 1 ? oth + hx + sgn : own)];
    const double vk = (double)x[fs_opaque(z > 0 ? own - ny * rs : own)];
    const double vf = (double)x[fs_opaque(z < nz - 1 ? own + ny * rs : own)];
    const double vs = (double)x[fs_opaque(y > 0 ? own - rs : own)];
    const double vn = (double)x[fs_opaque(y < ny - 1 ? own + rs : own)];
    double sum = 0.0, diag = 6.0;
    if (xx > 0) sum += vw; else diag += 1.0;
    if (xx < nx - 1) sum += ve; else diag += 1.0;
    if (z > 0) sum += vk; else diag += 1.0;
    if (z < nz - 1) sum += vf; else diag += 1.0;
    if (y > 0) sum += vs; else diag -= 1.0;
    if (y < ny - 1) sum += vn; else diag += 1.0;
    return bv - (sum - diag * me) * ih2;
}

// The sum of v over the 64 lanes of a wave, in a fixed butterfly order; lane 0's value is reproducible.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
