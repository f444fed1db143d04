// Geometric multigrid for the pressure Poisson equation on gfx950.
//
// Operator (src/sim.js:416-495): cell-centred 7-point Laplacian, p = 0 on the
// x/z faces and the top (ghost = -p), Neumann floor (ghost = +p).  V(2,2) with
// red-black Gauss-Seidel (omega = 1), residual fused into the 8-child restriction,
// mirrored trilinear prolongation, `coarse_iters` sweeps on the coarsest level —
// the same arithmetic, in the same order, as the reference's dispatches.
//
// Layout: every level's p and b (and the FP64 iterate) are stored parity-split:
// each x-row holds its even-x cells, then its odd-x cells (row stride 2*ceil(nx/2)).
// A colour pass then reads and writes one contiguous half-row per cell class —
// 12 bytes per updated cell instead of the 24 a stride-2 pass over x-fastest rows
// moves — and the y/z neighbours sit in the same half of the adjacent rows.
//
// Mixed precision: FP32 V-cycles on the correction inside FP64 defect correction
// (x <- x + V(b - A x)).  The residual and its norms are FP64; the stopping test
// runs on the device and later kernels of the same step early-exit on a device
// flag, so a whole step needs no host round trip (and is graph-capturable).
// Norm reductions use per-block partials summed in a fixed order: reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "fs_internal.h"
#include "fs_resid.h"

namespace {

constexpr int RB = 256;   // threads per block of every multigrid kernel (64 x 4)


// The skip test of a V-cycle kernel comes after its data loads: an empty volatile asm
// that consumes the loaded values (and the flag) pins them before the branch, so the
// flag's load and the data loads are in flight together (one memory latency, not two).
// A plain memory clobber does not: loads through const __restrict__ pointers may move
// across it.

// grid over (half-row index or x, y, z) with 64 x 4 threads per block
inline dim3 grid3(int w, int ny, int nz) { return dim3((unsigned)((w + 63) / 64), (unsigned)((ny + 3) / 4), (unsigned)nz); }

// p_fine += mirrored trilinear interpolation of the coarse p (sim.js:473-495);
// thread (hx, y, z) updates the fine cells x = 2hx and 2hx+1
__device__ __forceinline__ float prolong_at(const float* __restrict__ pc, int x, int y, int z, int cx, int cy, int cz) {
    const int nhc = (cx + 1) >> 1, rsc = 2 * nhc;
    const float gx = ((float)x + 0.5f) * 0.5f - 0.5f, gy = ((float)y + 0.5f) * 0.5f - 0.5f, gz = ((float)z + 0.5f) * 0.5f - 0.5f;
    const float flx = floorf(gx), fly = floorf(gy), flz = floorf(gz);
    const int g0x = (int)flx, g0y = (int)fly, g0z = (int)flz;
    const float fx = gx - flx, fy = gy - fly, fz = gz - flz;
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
        const int qx = g0x + ox, qy = g0y + oy, qz = g0z + oz;
        const int sx = min(max(qx, 0), cx - 1), sy = min(max(qy, 0), cy - 1), sz = min(max(qz, 0), cz - 1);
        float val = pc[(sz * cy + sy) * rsc + (sx & 1) * nhc + (sx >> 1)];
        if (qx < 0 || qx >= cx) val = -val;
        if (qz < 0 || qz >= cz) val = -val;
        if (qy >= cy) val = -val;
        const float wx = ox ? fx : 1.f - fx, wy = oy ? fy : 1.f - fy, wz = oz ? fz : 1.f - fz;
        acc += val * wx * wy * wz;
    }
    return acc;
}
constexpr int FIN = RB;     // threads of the one block that sums the residual's partials

// Sums the per-block partials of the last residual kernel in a fixed order, once.  st->pending
// says whether a residual awaits its reduction and which: a solve's starting residual (phase 0),
// or the residual after one more V-cycle (the phase count lives in the device state, so one
// captured iteration can be replayed any number of times).  Sets st->done once rel <= tol or the
// phase reaches max_cycles.  Every thread reads st->pending before the barrier that precedes its
// only write, so the block takes one branch.
__device__ void finalize_sums(RefineState* st, const double* partials, int npart, double tol, int max_cycles) {
    __shared__ double sr[FIN / 64], sb[FIN / 64];
    double rr = 0.0, bb = 0.0;
    const int pend = st->pending, done = st->done;     // in flight with the partials' loads
    // each thread sums partials j, j + FIN, j + 2 FIN, ... in that order, loading 8 at a time
    int j = threadIdx.x;
    for (; j + 7 * FIN < npart; j += 8 * FIN) {
        double a[8], c[8];
#pragma unroll
        for (int u = 0; u < 8; u++) { a[u] = partials[2 * (j + u * FIN)]; c[u] = partials[2 * (j + u * FIN) + 1]; }
#pragma unroll
        for (int u = 0; u < 8; u++) { rr += a[u]; bb += c[u]; }
    }
    for (; j < npart; j += FIN) { rr += partials[2 * j]; bb += partials[2 * j + 1]; }
    asm volatile("" ::"v"(rr), "v"(bb), "s"(pend), "s"(done));
    if (pend == 0) return;                             // uniform over the block
    const int init = pend == 2;
    rr = wave_sum(rr);
    bb = wave_sum(bb);
    if ((threadIdx.x & 63) == 0) { sr[threadIdx.x >> 6] = rr; sb[threadIdx.x >> 6] = bb; }
    __syncthreads();
    if (threadIdx.x == 0) {
        st->pending = 0;
        if (!init && done) return;                     // (a residual is not formed once the solve has stopped)
        double r2 = 0.0, b2 = 0.0;
        for (int w = 0; w < FIN / 64; w++) { r2 += sr[w]; b2 += sb[w]; }
        const double rel = b2 > 0.0 ? sqrt(r2 / b2) : 0.0;
        const int phase = init ? 0 : st->used + max(st->inc, 1);   // V-cycles applied so far in this solve
        st->rel[phase] = rel;
        st->used = phase;
        st->bnorm2 = b2;
        st->done = (rel <= tol || phase >= max_cycles) ? 1 : 0;
        st->skipv = st->done | st->hold;
    }
}

// one colour pass of red-black GS (sim.js:422-437).  Thread (hx, y, z) owns the cell
// x = 2hx + s of this colour, s = (y + z + color) & 1.  With `xupd` set (the last pass
// of a correction V-cycle) it also adds the final p of both cells x = 2hx, 2hx+1 to
// the FP64 iterate.  omega != 1 over-relaxes: p <- p + omega (p_GS - p) (the
// tolerance-driven solves; the reference's solve is omega = 1 exactly).  WithFin: block
// (0, 0, 0) first runs the pending reduction of the previous residual kernel (the pass is
// then launched unconditionally: the stop flag it would test is the one being formed).
template <bool WithFin>
__global__ __launch_bounds__(RB) void k_smooth(float* __restrict__ p, const float* __restrict__ b, int nx, int ny, int nz,
                                               float h, int color, const int* __restrict__ skip, double* __restrict__ xupd,
                                               float omega, const float* __restrict__ pc_own, FinArgs fin) {
    if (WithFin && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0)
        finalize_sums(fin.st, fin.partials, fin.npart, fin.tol, fin.max_cycles);
    const int sk = skip[fs_opaque(0)];   // in flight with the data loads (a vector load: its wait does not hold up the
                                         // scalar kernel arguments); tested before any store
    const int nh = (nx + 1) >> 1, rs = 2 * nh;
    const int hx = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (hx >= nh || y >= ny) return;
    const int sgn = (y + z + color) & 1;
    const int x = 2 * hx + sgn;
    const int row = (z * ny + y) * rs;
    const int own = row + sgn * nh + hx;
    const int oth = row + (1 - sgn) * nh;             // the other parity's half of this row
    // every neighbour is loaded (an outside one from the cell itself, unused), so the
    // loads issue together; the boundary rule then picks the same f32 operations
    const float vw = p[fs_opaque(x > 0 ? oth + hx - 1 + sgn : own)];
    const float ve = p[fs_opaque(x < nx - 1 ? oth + hx + sgn : own)];
    const float vk = p[fs_opaque(z > 0 ? own - ny * rs : own)];
    const float vf = p[fs_opaque(z < nz - 1 ? own + ny * rs : own)];
    const float vs = p[fs_opaque(y > 0 ? own - rs : own)];
    const float vn = p[fs_opaque(y < ny - 1 ? own + rs : own)];
    const float bv = b[own];
    // pc_own: this colour's cells have not been prolonged yet (k_prolong_red left them); their own value is
    // p + P(pc), formed as k_prolong forms it
    const float po = omega != 1.f ? (pc_own ? p[own] + prolong_at(pc_own, x, y, z, nx >> 1, ny >> 1, nz >> 1) : p[own]) : 0.f;
    const bool pair = xupd && 2 * hx + 1 - sgn < nx;  // the pair's other cell, final already
    const float pother = pair ? p[oth + hx] : 0.f;
    const double xo = xupd ? xupd[own] : 0.0;
    const double xp = pair ? xupd[oth + hx] : 0.0;
    float sum = 0.f, diag = 6.f;
    if (x > 0) sum += vw; else diag += 1.f;
    if (x < nx - 1) sum += ve; else diag += 1.f;
    if (z > 0) sum += vk; else diag += 1.f;
    if (z < nz - 1) sum += vf; else diag += 1.f;
    if (y > 0) sum += vs; else diag -= 1.f;
    if (y < ny - 1) sum += vn; else diag += 1.f;
    float pn = (sum - h * h * bv) / diag;
    if (omega != 1.f) pn = po + omega * (pn - po);
    asm volatile("" ::"v"(pn), "v"(pother), "v"(xo), "v"(xp), "v"(sk));
    if (sk) return;
    if (x < nx) {
        p[own] = pn;
        if (xupd) xupd[own] = xo + (double)pn;
    }
    if (pair) xupd[oth + hx] = xp + (double)pother;
}

// residual of the 8 children averaged onto the coarse cell; coarse p = 0 (sim.js:446-466)
__global__ __launch_bounds__(RB) void k_restrict(const float* __restrict__ p, const float* __restrict__ b, int nx, int ny, int nz,
                                                 float h, float* __restrict__ bc, float* __restrict__ pc,
                                                 const int* __restrict__ skip, float hc, float omega, int red_first) {
    const int sk = skip[fs_opaque(0)];   // a vector load: its wait does not hold up the scalar kernel arguments
    const int cx = nx >> 1, cy = ny >> 1;
    const int xc = blockIdx.x * 64 + (threadIdx.x & 63);
    const int yc = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int zc = blockIdx.z;
    if (xc >= cx || yc >= cy) return;
    const int nh = nx >> 1, rs = nx;                   // levels that restrict have even nx
    const float ih2 = 1.f / (h * h);
    // the 2x2x2 children and their 24 outer neighbours, loaded together (outside ones from
    // the child itself, unused); then the ghost rule of sim.js:446-466 per child
    float me[8], nb[8][6], bv[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int dx = k & 1, y = yc * 2 + ((k >> 1) & 1), z = zc * 2 + ((k >> 2) & 1);
        const int x = xc * 2 + dx;
        const int row = (z * ny + y) * rs;
        const int c = row + dx * nh + xc;
        // x-1 / x+1 live in the other parity's half: (2xc)-1 = odd xc-1, (2xc+1)-1 = even xc, etc.
        me[k] = p[c];
        nb[k][0] = p[fs_opaque(x > 0 ? row + (1 - dx) * nh + xc - 1 + dx : c)];
        nb[k][1] = p[fs_opaque(x < nx - 1 ? row + (1 - dx) * nh + xc + dx : c)];
        nb[k][2] = p[fs_opaque(y > 0 ? c - rs : c)];
        nb[k][3] = p[fs_opaque(y < ny - 1 ? c + rs : c)];
        nb[k][4] = p[fs_opaque(z > 0 ? c - ny * rs : c)];
        nb[k][5] = p[fs_opaque(z < nz - 1 ? c + ny * rs : c)];
        bv[k] = b[c];
    }
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int dx = k & 1, y = yc * 2 + ((k >> 1) & 1), z = zc * 2 + ((k >> 2) & 1);
        const int x = xc * 2 + dx;
        const float m = me[k];
        const float pxm = x > 0 ? nb[k][0] : -m;
        const float pxp = x < nx - 1 ? nb[k][1] : -m;
        const float pym = y > 0 ? nb[k][2] : m;
        const float pyp = y < ny - 1 ? nb[k][3] : -m;
        const float pzm = z > 0 ? nb[k][4] : -m;
        const float pzp = z < nz - 1 ? nb[k][5] : -m;
        const float lap = (pxm + pxp + pym + pyp + pzm + pzp - 6.f * m) * ih2;
        acc += bv[k] - lap;
    }
    asm volatile("" ::"v"(acc), "v"(sk));
    if (sk) return;
    const int nhc = (cx + 1) >> 1;
    const int ci = (zc * cy + yc) * 2 * nhc + (xc & 1) * nhc + (xc >> 1);
    const float bnew = acc * 0.125f;
    bc[ci] = bnew;
    // the coarse iterate starts at zero; with red_first its colour-0 cells get their first update
    pc[ci] = (red_first && ((xc + yc + zc) & 1) == 0) ? red_from_zero(bnew, xc, yc, zc, cx, cy, nz >> 1, hc, omega) : 0.f;
}

// The mirrored trilinear weights are 0.25 or 0.75, so a term ((v wx) wy) wz rounds only at its factors
// 0.75 (a factor 0.25 scales exactly, and exact scaling commutes with rounding): it equals m_n(v) 4^(n-3),
// n the number of factors 0.75, m_0 = v, m_k = m_(k-1) 0.75 rounded; and acc + m_n(v) 4^(n-3) is one FMA
// with the same rounding as the sum.  With m_1..m_3 formed once per coarse value (m[k] below), each term
// costs one FMA instead of three products and a sum, bit for bit (for |v| >= 2^-120, where the scaling
// by 4^-3 is exact; the multigrid's iterates are far from that).
__device__ __forceinline__ void chain75(const float (&c)[3][3][3], float (&m)[4][3][3][3]) {
#pragma unroll
    for (int i = 0; i < 27; i++) {
        const float v = (&c[0][0][0])[i];
        const float m1 = v * 0.75f, m2 = m1 * 0.75f;
        (&m[0][0][0][0])[i] = v;
        (&m[1][0][0][0])[i] = m1;
        (&m[2][0][0][0])[i] = m2;
        (&m[3][0][0][0])[i] = m2 * 0.75f;
    }
}
// the prolongation correction of child (a, bb, cc) of the coarse cell offset (sx, sy, sz) from the
// window's centre (k_prolong's sum, term for term)
__device__ __forceinline__ float child_corr(const float (&m)[4][3][3][3], int sx, int sy, int sz, int a, int bb, int cc) {
    float acc = 0.f;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int ox = j & 1, oy = (j >> 1) & 1, oz = (j >> 2) & 1;
        const int n = (ox != a) + (oy != bb) + (oz != cc);   // factors 0.75: the nearer coarse cell on an axis
        const float sc = n == 3 ? 1.f : n == 2 ? 0.25f : n == 1 ? 0.0625f : 0.015625f;
        acc = __builtin_fmaf(m[n][cc + oz + sz][bb + oy + sy][a + ox + sx], sc, acc);
    }
    return acc;
}

// One thread per coarse cell: its 3x3x3 coarse neighbourhood (mirror rule applied)
// gives the corrections of its 8 children, each summed in prolong_at's order (the
// child's two coarse cells per axis sit at offsets {-1, 0} or {0, +1}, weights 1/4
// and 3/4), so the result is bit-identical to the per-fine-cell form.  The 27 coarse
// loads are shared by 8 fine cells and every child class is a coalesced half-row.
__global__ __launch_bounds__(RB) void k_prolong(float* __restrict__ p, const float* __restrict__ pcoarse, int nx, int ny, int nz,
                                                const int* __restrict__ skip) {
    const int sk = skip[fs_opaque(0)];   // a vector load: its wait does not hold up the scalar kernel arguments
    const int cx = nx >> 1, cy = ny >> 1, cz = nz >> 1;    // levels that prolong have even dims
    const int xc = blockIdx.x * 64 + (threadIdx.x & 63);
    const int yc = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int zc = blockIdx.z;
    if (xc >= cx || yc >= cy) return;
    const int nhc = (cx + 1) >> 1, rsc = 2 * nhc;
    float c[3][3][3];
#pragma unroll
    for (int dz = 0; dz < 3; dz++) {
        const int qz = zc + dz - 1, sz = min(max(qz, 0), cz - 1);
#pragma unroll
        for (int dy = 0; dy < 3; dy++) {
            const int qy = yc + dy - 1, sy = min(max(qy, 0), cy - 1);
#pragma unroll
            for (int dx = 0; dx < 3; dx++) {
                const int qx = xc + dx - 1, sx = min(max(qx, 0), cx - 1);
                float v = pcoarse[(sz * cy + sy) * rsc + (sx & 1) * nhc + (sx >> 1)];
                if (qx < 0 || qx >= cx) v = -v;
                if (qz < 0 || qz >= cz) v = -v;
                if (qy >= cy) v = -v;
                c[dz][dy][dx] = v;
            }
        }
    }
    const int nh = nx >> 1, rs = nx;
    float m[4][3][3][3];
    chain75(c, m);
    float fine[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int a = k & 1, bb = (k >> 1) & 1, cc = (k >> 2) & 1;   // child offsets in x, y, z
        // sum_j c * wx * wy * wz over the 8 coarse cells, wx = 0.75 on the nearer side (0.25 on the other)
        fine[k] = p[((2 * zc + cc) * ny + 2 * yc + bb) * rs + a * nh + xc] + child_corr(m, 0, 0, 0, a, bb, cc);
    }
    asm volatile("" ::"v"(fine[0]), "v"(fine[1]), "v"(fine[2]), "v"(fine[3]), "v"(fine[4]), "v"(fine[5]), "v"(fine[6]),
                 "v"(fine[7]), "v"(sk));
    if (sk) return;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int a = k & 1, bb = (k >> 1) & 1, cc = (k >> 2) & 1;
        p[((2 * zc + cc) * ny + 2 * yc + bb) * rs + a * nh + xc] = fine[k];
    }
}

// The prolongation correction of the child with offsets (a, bb, cc) of the coarse cell at offset (sx, sy, sz)
// (each -1, 0 or +1) from the thread's own, from the thread's 3x3x3 mirrored neighbourhood c: k_prolong's sum
// for that child, term for term (the parent's neighbourhood is c shifted by the offset).
// Prolongation fused with the first post-smoothing pass (colour 0), at any omega.  One thread per
// coarse cell forms its 8 children's prolonged values p + P(pc) as k_prolong does, then updates its
// 4 colour-0 children as k_smooth does after the prolongation: a colour-0 child's 6 neighbours are
// colour 1, 3 of them siblings (in registers) and 3 children of the adjacent coarse cells, whose
// prolonged values the thread forms from its own 3x3x3 neighbourhood in k_prolong's order (the
// coarse cells they interpolate lie within it).  Only colour-0 cells are stored: colour-1 cells are
// read by neighbouring threads and keep their unprolonged values, which the colour-1 pass that
// follows prolongs for itself (k_smooth's pc_own).  The values equal those of k_prolong followed by
// the two colour passes, bit for bit, with one pass over p and one launch less.
// With `pout` set, the thread stores all 8 children there instead (the colour-0 ones updated, the colour-1
// ones prolonged), so the colour-1 pass that follows runs on pout and reads its cells' own values instead of
// forming their prolongation again; p is then only read.
__global__ __launch_bounds__(RB) void k_prolong_red(float* __restrict__ p, const float* __restrict__ b,
                                                    const float* __restrict__ pcoarse, int nx, int ny, int nz, float h,
                                                    float omega, const int* __restrict__ skip, float* __restrict__ pout) {
    const int sk = skip[fs_opaque(0)];   // a vector load: its wait does not hold up the scalar kernel arguments
    const int cx = nx >> 1, cy = ny >> 1, cz = nz >> 1;    // levels that prolong have even dims
    const int xc = blockIdx.x * 64 + (threadIdx.x & 63);
    const int yc = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int zc = blockIdx.z;
    if (xc >= cx || yc >= cy) return;
    const int nhc = (cx + 1) >> 1, rsc = 2 * nhc;
    float c[3][3][3];
#pragma unroll
    for (int dz = 0; dz < 3; dz++) {
        const int qz = zc + dz - 1, sz = min(max(qz, 0), cz - 1);
#pragma unroll
        for (int dy = 0; dy < 3; dy++) {
            const int qy = yc + dy - 1, sy = min(max(qy, 0), cy - 1);
#pragma unroll
            for (int dx = 0; dx < 3; dx++) {
                const int qx = xc + dx - 1, sx = min(max(qx, 0), cx - 1);
                float v = pcoarse[(sz * cy + sy) * rsc + (sx & 1) * nhc + (sx >> 1)];
                if (qx < 0 || qx >= cx) v = -v;
                if (qz < 0 || qz >= cz) v = -v;
                if (qy >= cy) v = -v;
                c[dz][dy][dx] = v;
            }
        }
    }
    const int nh = nx >> 1, rs = nx;
    // fine cell (x, y, z) in the parity-split layout
    auto fidx = [&](int x, int y, int z) { return (z * ny + y) * rs + (x & 1) * nh + (x >> 1); };
    float m[4][3][3][3];
    chain75(c, m);
    float fine[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int a = k & 1, bb = (k >> 1) & 1, cc = (k >> 2) & 1;   // child offsets in x, y, z
        fine[k] = p[fidx(2 * xc + a, 2 * yc + bb, 2 * zc + cc)] + child_corr(m, 0, 0, 0, a, bb, cc);
    }
    // the colour-0 children (a + bb + cc even) and their outer colour-1 neighbours (loaded from an in-bounds
    // cell, the child itself, where the neighbour is outside; unused then)
    constexpr int RED[4] = {0, 3, 5, 6};
    float pn[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int k = RED[r];
        const int a = k & 1, bb = (k >> 1) & 1, cc = (k >> 2) & 1;
        const int x = 2 * xc + a, y = 2 * yc + bb, z = 2 * zc + cc;
        const int own = fidx(x, y, z);
        // x: the sibling on the inner side, the adjacent coarse cell's child on the outer side
        const float vw = a ? fine[k - 1] : p[fs_opaque(x > 0 ? fidx(x - 1, y, z) : own)] + child_corr(m, -1, 0, 0, 1, bb, cc);
        const float ve = a ? p[fs_opaque(x < nx - 1 ? fidx(x + 1, y, z) : own)] + child_corr(m, 1, 0, 0, 0, bb, cc) : fine[k + 1];
        const float vk = cc ? fine[k - 4] : p[fs_opaque(z > 0 ? fidx(x, y, z - 1) : own)] + child_corr(m, 0, 0, -1, a, bb, 1);
        const float vf = cc ? p[fs_opaque(z < nz - 1 ? fidx(x, y, z + 1) : own)] + child_corr(m, 0, 0, 1, a, bb, 0) : fine[k + 4];
        const float vs = bb ? fine[k - 2] : p[fs_opaque(y > 0 ? fidx(x, y - 1, z) : own)] + child_corr(m, 0, -1, 0, a, 1, cc);
        const float vn = bb ? p[fs_opaque(y < ny - 1 ? fidx(x, y + 1, z) : own)] + child_corr(m, 0, 1, 0, a, 0, cc) : fine[k + 2];
        const float bv = b[own];
        const float po = fine[k];
        float sum = 0.f, diag = 6.f;
        if (x > 0) sum += vw; else diag += 1.f;
        if (x < nx - 1) sum += ve; else diag += 1.f;
        if (z > 0) sum += vk; else diag += 1.f;
        if (z < nz - 1) sum += vf; else diag += 1.f;
        if (y > 0) sum += vs; else diag -= 1.f;
        if (y < ny - 1) sum += vn; else diag += 1.f;
        float q = (sum - h * h * bv) / diag;
        if (omega != 1.f) q = po + omega * (q - po);
        pn[r] = q;
    }
    asm volatile("" ::"v"(pn[0]), "v"(pn[1]), "v"(pn[2]), "v"(pn[3]), "v"(sk));
    if (sk) return;
    if (pout) {
        fine[0] = pn[0]; fine[3] = pn[1]; fine[5] = pn[2]; fine[6] = pn[3];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int a = k & 1, bb = (k >> 1) & 1, cc = (k >> 2) & 1;
            pout[fidx(2 * xc + a, 2 * yc + bb, 2 * zc + cc)] = fine[k];
        }
        return;
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int k = RED[r];
        const int a = k & 1, bb = (k >> 1) & 1, cc = (k >> 2) & 1;
        p[fidx(2 * xc + a, 2 * yc + bb, 2 * zc + cc)] = pn[r];
    }
}

// Prolongation fused into the first post-smoothing pass (colour 0).  A Gauss-Seidel
// update overwrites its own cell, so the colour-0 pass only needs the prolonged
// values p + P(pc) of its colour-1 neighbours, which each thread forms exactly as
// k_prolong would (same f32 operations); the colour-1 pass that follows overwrites
// every colour-1 cell without reading it.  Saves the prolongation's full read and
// write of p and one launch per level, with identical results.
__global__ __launch_bounds__(RB) void k_smooth_prolong(float* __restrict__ p, const float* __restrict__ b,
                                                       const float* __restrict__ pcoarse, int nx, int ny, int nz, float h,
                                                       const int* __restrict__ skip) {
    const int sk = skip[fs_opaque(0)];   // a vector load: its wait does not hold up the scalar kernel arguments
    const int nh = nx >> 1, rs = nx;                   // levels that prolong have even nx
    const int hx = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (hx >= nh || y >= ny) return;
    const int cx = nx >> 1, cy = ny >> 1, cz = nz >> 1;
    const int sgn = (y + z) & 1;                       // colour 0
    const int x = 2 * hx + sgn;
    const int row = (z * ny + y) * rs;
    const int own = row + sgn * nh + hx;
    const int oth = row + (1 - sgn) * nh;
    // all six neighbours loaded (outside ones from the cell itself, unused; prolong_at
    // clamps its coarse reads for any coordinate), then the boundary rule
    const float vw = p[fs_opaque(x > 0 ? oth + hx - 1 + sgn : own)] + prolong_at(pcoarse, x - 1, y, z, cx, cy, cz);
    const float ve = p[fs_opaque(x < nx - 1 ? oth + hx + sgn : own)] + prolong_at(pcoarse, x + 1, y, z, cx, cy, cz);
    const float vk = p[fs_opaque(z > 0 ? own - ny * rs : own)] + prolong_at(pcoarse, x, y, z - 1, cx, cy, cz);
    const float vf = p[fs_opaque(z < nz - 1 ? own + ny * rs : own)] + prolong_at(pcoarse, x, y, z + 1, cx, cy, cz);
    const float vs = p[fs_opaque(y > 0 ? own - rs : own)] + prolong_at(pcoarse, x, y - 1, z, cx, cy, cz);
    const float vn = p[fs_opaque(y < ny - 1 ? own + rs : own)] + prolong_at(pcoarse, x, y + 1, z, cx, cy, cz);
    const float bv = b[own];
    float sum = 0.f, diag = 6.f;
    if (x > 0) sum += vw; else diag += 1.f;
    if (x < nx - 1) sum += ve; else diag += 1.f;
    if (z > 0) sum += vk; else diag += 1.f;
    if (z < nz - 1) sum += vf; else diag += 1.f;
    if (y > 0) sum += vs; else diag -= 1.f;
    if (y < ny - 1) sum += vn; else diag += 1.f;
    const float pn = (sum - h * h * bv) / diag;
    asm volatile("" ::"v"(pn), "v"(sk));
    if (sk) return;
    p[own] = pn;
}

// FP64 residual over the level-0 grid.  Each block writes its (r.r, b.b) partial sums
// (finalize_sums reduces them in index order, later); block 0 marks them pending (`init`:
// a solve's starting residual).  With `rhs_out` set it also stores r (f32) as the V-cycle
// right-hand side and zeroes the V-cycle iterate.
template <class X>
__global__ __launch_bounds__(RB) void k_resid64(const X* __restrict__ x, const float* __restrict__ b, int nx, int ny, int nz,
                                                double h, float* __restrict__ rhs_out, float* __restrict__ p_out,
                                                double* __restrict__ partials, RefineState* __restrict__ st, int init,
                                                const int* __restrict__ skip, float hred = 0.f, float omega = 1.f,
                                                int red_first = 0, int cycles = 1) {
    __shared__ double sr[RB / 64], sb[RB / 64];
    const int nh = (nx + 1) >> 1, rs = 2 * nh;
    const int hx = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    const double ih2 = 1.0 / (h * h);
    double rr = 0.0, bb = 0.0;
    const bool inside = hx < nh && y < ny;
    double r[2] = {0.0, 0.0}, bv[2] = {0.0, 0.0};
    if (inside) {
        // both cells of the pair (the odd one of a padded row slot is computed and dropped)
        const PairStencil ps = pair_stencil(LevelBuf<X>(x, nz * ny * rs), hx, y, z, nx, ny, nz, nh, rs);
        bv[0] = (double)b[ps.o0];
        bv[1] = (double)b[ps.o1];
        r[0] = pair_resid(ps, 0, bv[0], hx, y, z, nx, ny, nz, ih2);
        r[1] = pair_resid(ps, 1, bv[1], hx, y, z, nx, ny, nz, ih2);
    }
    const int sk = skip[fs_opaque(0)];   // a vector load: its wait does not hold up the scalar kernel arguments
    asm volatile("" ::"v"(r[0]), "v"(r[1]), "v"(bv[0]), "v"(bv[1]), "v"(sk));
    if (sk) return;                                    // uniform over the grid
    if (inside) {
#pragma unroll
        for (int sgn = 0; sgn < 2; sgn++) {
            if (2 * hx + sgn >= nx) continue;
            rr += r[sgn] * r[sgn];
            bb += bv[sgn] * bv[sgn];
            if (rhs_out) {
                const int own = (z * ny + y) * rs + sgn * nh + hx;
                const int xx = 2 * hx + sgn;
                const float bf = (float)r[sgn];
                rhs_out[own] = bf;
                if (p_out)
                    p_out[own] = (red_first && ((xx + y + z) & 1) == 0) ? red_from_zero(bf, xx, y, z, nx, ny, nz, hred, omega)
                                                                       : 0.f;
            }
        }
    }
    if (!partials) return;                             // (only the right-hand side was asked for; uniform)
    rr = wave_sum(rr);
    bb = wave_sum(bb);
    if ((threadIdx.x & 63) == 0) { sr[threadIdx.x >> 6] = rr; sb[threadIdx.x >> 6] = bb; }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int blk = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        partials[2 * blk] = ((sr[0] + sr[1]) + sr[2]) + sr[3];
        partials[2 * blk + 1] = ((sb[0] + sb[1]) + sb[2]) + sb[3];
        if (blk == 0) {
            st->pending = init ? 2 : 1;
            st->inc = cycles;                          // the V-cycles this residual follows
        }
    }
}

// one block: the fixed-order reduction of the partials and the stopping test; with `host` set (pinned host
// memory) it also stores {used, done} there for the host to read after the kernel
__global__ __launch_bounds__(FIN) void k_finalize(RefineState* __restrict__ st, const double* __restrict__ partials, int npart,
                                                 double tol, int max_cycles, int* __restrict__ host = nullptr,
                                                 const int* __restrict__ gate = nullptr, int set_hold = 0) {
    if (gate && gate[0]) return;                      // (uniform; read before the block's only writes)
    finalize_sums(st, partials, npart, tol, max_cycles);
    if (threadIdx.x == 0) {   // thread 0 made the block's only state writes
        if (set_hold) {       // the step's closing test: a step enqueued behind it runs only once this one is done
            st->hold = st->done ? 0 : 1;
            st->skipv = 1;
        }
        if (host) {
            host[0] = st->used;
            host[1] = st->done;
        }
    }
}

// x += e (the V-cycle's correction), FP64, over the whole parity-split array
__global__ __launch_bounds__(RB) void k_update(double* __restrict__ x, const float* __restrict__ e, int n,
                                               const int* __restrict__ skip) {
    const int sk = skip[fs_opaque(0)];   // a vector load: its wait does not hold up the scalar kernel arguments
    const int i = blockIdx.x * RB + threadIdx.x;
    if (i >= n) return;
    const double xv = x[i];
    const float ev = e[i];
    asm volatile("" ::"v"(xv), "v"(ev), "v"(sk));
    if (sk) return;
    x[i] = xv + (double)ev;
}

// ---- coarse tail: every level from `tail_from` down, one workgroup, all in LDS ----
// The coarse levels are tiny (<= 8192 cells at the tail's top level) and a launch
// per colour pass would cost far more than the work; one 1024-thread workgroup runs
// their whole part of the V-cycle (pre-smooth, restrict, coarse sweeps, prolong,
// post-smooth) with the same arithmetic as the per-level kernels above.
constexpr int TAIL_THREADS = 1024;
constexpr int TAIL_LDS_BYTES = 64 * 1024;

// exact n / d for n, d < 2^16 with one multiply-high: m = ceil(2^32 / d)
struct FastDiv {
    unsigned m;
    int d;
};
static FastDiv make_fastdiv(int d) {
    FastDiv f;
    f.d = d;
    f.m = (unsigned)((((unsigned long long)1 << 32) + (unsigned)d - 1) / (unsigned)d);
    return f;
}
__device__ __forceinline__ int fdiv(int n, FastDiv f) { return (int)__umulhi((unsigned)n, f.m); }

struct TailArgs {
    int nt;                          // number of tail levels
    FastDiv dhx[MG_MAX_LEVELS], dx[MG_MAX_LEVELS], dy[MG_MAX_LEVELS], dcx[MG_MAX_LEVELS], dcy[MG_MAX_LEVELS];
    int dims[MG_MAX_LEVELS][3];
    float h[MG_MAX_LEVELS];
    int off[MG_MAX_LEVELS];          // float offset of level j's p in LDS; b follows p
    int coarse_iters;
    float omega;                     // smoother over-relaxation above the coarsest level
    int pre, post;                   // sweeps before / after the coarse correction
    int red_done;                    // the top level's first colour-0 pass was done by the restriction
    float* p_top;                    // global p / b of the tail's top level (parity-split layout)
    const float* b_top;
};

// linear thread-to-cell mapping; the divisions are multiply-highs (FastDiv)
// Up to TAIL_ILP cells of one colour per thread and pass: all their LDS reads are issued
// before any of their writes (cells of one colour do not read each other), so the reads
// overlap instead of running as one dependent chain per cell.
constexpr int TAIL_ILP = 4;
__device__ __forceinline__ void tail_smooth(float* p, const float* b, int nx, int ny, int nz, float h, int iters,
                                            FastDiv dhx, FastDiv dy, float omega, int first_red_done = 0) {
    const int nhx = (nx + 1) >> 1, nh = nhx * ny * nz, sy = nx, sz = nx * ny;
    for (int it = 0; it < iters; it++)
        for (int color = 0; color < 2; color++) {
            if (it == 0 && color == 0 && first_red_done) continue;   // written by the restriction
            if (nh <= TAIL_THREADS) {          // one cell per thread at most: the plain pass
                const int i = threadIdx.x;
                const int t = fdiv(i, dhx), hx = i - t * nhx;
                const int z = fdiv(t, dy), y = t - z * ny;
                const int x = hx * 2 + ((y + z + color) & 1);
                if (i < nh && x < nx) {
                    const int c = (z * ny + y) * nx + x;
                    const float vw = p[fs_opaque(x > 0 ? c - 1 : c)], ve = p[fs_opaque(x < nx - 1 ? c + 1 : c)];
                    const float vk = p[fs_opaque(z > 0 ? c - sz : c)], vf = p[fs_opaque(z < nz - 1 ? c + sz : c)];
                    const float vs = p[fs_opaque(y > 0 ? c - sy : c)], vn = p[fs_opaque(y < ny - 1 ? c + sy : c)];
                    const float bv = b[c], po = p[c];
                    float sum = 0.f, diag = 6.f;
                    if (x > 0) sum += vw; else diag += 1.f;
                    if (x < nx - 1) sum += ve; else diag += 1.f;
                    if (z > 0) sum += vk; else diag += 1.f;
                    if (z < nz - 1) sum += vf; else diag += 1.f;
                    if (y > 0) sum += vs; else diag -= 1.f;
                    if (y < ny - 1) sum += vn; else diag += 1.f;
                    float pn = (sum - h * h * bv) / diag;
                    if (omega != 1.f) pn = po + omega * (pn - po);
                    p[c] = pn;
                }
                __syncthreads();
                continue;
            }
            for (int i0 = threadIdx.x; i0 < nh; i0 += TAIL_ILP * TAIL_THREADS) {
                float v[TAIL_ILP][8];
                int cc[TAIL_ILP], xs[TAIL_ILP], ys[TAIL_ILP], zs[TAIL_ILP];
                bool ok[TAIL_ILP];
#pragma unroll
                for (int u = 0; u < TAIL_ILP; u++) {
                    const int i = i0 + u * TAIL_THREADS;
                    const int t = fdiv(i, dhx), hx = i - t * nhx;
                    const int z = fdiv(t, dy), y = t - z * ny;
                    const int x = hx * 2 + ((y + z + color) & 1);
                    ok[u] = i < nh && x < nx;
                    const int c = ok[u] ? (z * ny + y) * nx + x : 0;
                    cc[u] = c; xs[u] = x; ys[u] = y; zs[u] = z;
                    v[u][0] = p[fs_opaque(ok[u] && x > 0 ? c - 1 : c)];
                    v[u][1] = p[fs_opaque(ok[u] && x < nx - 1 ? c + 1 : c)];
                    v[u][2] = p[fs_opaque(ok[u] && z > 0 ? c - sz : c)];
                    v[u][3] = p[fs_opaque(ok[u] && z < nz - 1 ? c + sz : c)];
                    v[u][4] = p[fs_opaque(ok[u] && y > 0 ? c - sy : c)];
                    v[u][5] = p[fs_opaque(ok[u] && y < ny - 1 ? c + sy : c)];
                    v[u][6] = b[c];
                    v[u][7] = p[c];
                }
#pragma unroll
                for (int u = 0; u < TAIL_ILP; u++) {
                    if (!ok[u]) continue;
                    const int x = xs[u], y = ys[u], z = zs[u];
                    float sum = 0.f, diag = 6.f;
                    if (x > 0) sum += v[u][0]; else diag += 1.f;
                    if (x < nx - 1) sum += v[u][1]; else diag += 1.f;
                    if (z > 0) sum += v[u][2]; else diag += 1.f;
                    if (z < nz - 1) sum += v[u][3]; else diag += 1.f;
                    if (y > 0) sum += v[u][4]; else diag -= 1.f;
                    if (y < ny - 1) sum += v[u][5]; else diag += 1.f;
                    float pn = (sum - h * h * v[u][6]) / diag;
                    if (omega != 1.f) pn = v[u][7] + omega * (pn - v[u][7]);
                    p[cc[u]] = pn;
                }
            }
            __syncthreads();
        }
}
__device__ __forceinline__ void tail_restrict(const float* p, const float* b, int nx, int ny, int nz, float h, float* bc,
                                              float* pc, FastDiv dcx, FastDiv dcy) {
    const int cx = nx >> 1, cy = ny >> 1, cz = nz >> 1, n = cx * cy * cz;
    const float ih2 = 1.f / (h * h);
    const int sy = nx, sz = nx * ny;
    for (int i = threadIdx.x; i < n; i += TAIL_THREADS) {
        const int t = fdiv(i, dcx), xc = i - t * cx;
        const int zc = fdiv(t, dcy), yc = t - zc * cy;
        float acc = 0.f;
        for (int k = 0; k < 8; k++) {
            const int x = xc * 2 + (k & 1), y = yc * 2 + ((k >> 1) & 1), z = zc * 2 + ((k >> 2) & 1);
            const int c = (z * ny + y) * nx + x;
            const float me = p[c];
            const float vxm = p[fs_opaque(x > 0 ? c - 1 : c)], vxp = p[fs_opaque(x < nx - 1 ? c + 1 : c)];
            const float vym = p[fs_opaque(y > 0 ? c - sy : c)], vyp = p[fs_opaque(y < ny - 1 ? c + sy : c)];
            const float vzm = p[fs_opaque(z > 0 ? c - sz : c)], vzp = p[fs_opaque(z < nz - 1 ? c + sz : c)];
            const float pxm = x > 0 ? vxm : -me;
            const float pxp = x < nx - 1 ? vxp : -me;
            const float pym = y > 0 ? vym : me;
            const float pyp = y < ny - 1 ? vyp : -me;
            const float pzm = z > 0 ? vzm : -me;
            const float pzp = z < nz - 1 ? vzp : -me;
            const float lap = (pxm + pxp + pym + pyp + pzm + pzp - 6.f * me) * ih2;
            acc += b[c] - lap;
        }
        bc[i] = acc * 0.125f;
        pc[i] = 0.f;
    }
    __syncthreads();
}
__device__ __forceinline__ void tail_prolong(float* p, const float* pcoarse, int nx, int ny, int nz, FastDiv dx, FastDiv dy) {
    const int cx = nx >> 1, cy = ny >> 1, cz = nz >> 1, n = nx * ny * nz;
    for (int i = threadIdx.x; i < n; i += TAIL_THREADS) {
        const int t = fdiv(i, dx), x = i - t * nx;
        const int z = fdiv(t, dy), y = t - z * ny;
        const float gx = ((float)x + 0.5f) * 0.5f - 0.5f, gy = ((float)y + 0.5f) * 0.5f - 0.5f, gz = ((float)z + 0.5f) * 0.5f - 0.5f;
        const float flx = floorf(gx), fly = floorf(gy), flz = floorf(gz);
        const int g0x = (int)flx, g0y = (int)fly, g0z = (int)flz;
        const float fx = gx - flx, fy = gy - fly, fz = gz - flz;
        float acc = 0.f;
        for (int k = 0; k < 8; k++) {
            const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
            const int qx = g0x + ox, qy = g0y + oy, qz = g0z + oz;
            const int sx = min(max(qx, 0), cx - 1), sy = min(max(qy, 0), cy - 1), sz = min(max(qz, 0), cz - 1);
            float val = pcoarse[(sz * cy + sy) * cx + sx];
            if (qx < 0 || qx >= cx) val = -val;
            if (qz < 0 || qz >= cz) val = -val;
            if (qy >= cy) val = -val;
            const float wx = ox ? fx : 1.f - fx, wy = oy ? fy : 1.f - fy, wz = oz ? fz : 1.f - fz;
            acc += val * wx * wy * wz;
        }
        p[i] = p[i] + acc;
    }
    __syncthreads();
}

__global__ __launch_bounds__(TAIL_THREADS) void k_mg_tail(TailArgs a, const int* __restrict__ skip) {
    const int sk = skip[fs_opaque(0)];   // a vector load: its wait does not hold up the scalar kernel arguments
    extern __shared__ float lds[];
    const int n0 = a.dims[0][0] * a.dims[0][1] * a.dims[0][2];
    float* p0 = lds + a.off[0];
    float* b0 = p0 + n0;
    // LDS holds the tail levels in natural x-fastest order; the top level's global
    // arrays are in the parity-split layout
    const int nx0 = a.dims[0][0], ny0 = a.dims[0][1], nh0 = (nx0 + 1) >> 1;
    for (int i = threadIdx.x; i < n0; i += TAIL_THREADS) {
        const int t = fdiv(i, a.dx[0]), x = i - t * nx0;
        const int g = t * 2 * nh0 + (x & 1) * nh0 + (x >> 1);
        p0[i] = a.p_top[g];
        b0[i] = a.b_top[g];
    }
    (void)ny0;
    if (sk) return;                                    // uniform over the workgroup
    __syncthreads();
    for (int j = 0; j < a.nt - 1; j++) {
        const int nx = a.dims[j][0], ny = a.dims[j][1], nz = a.dims[j][2], n = nx * ny * nz;
        float* p = lds + a.off[j];
        float* b = p + n;
        const int nc = a.dims[j + 1][0] * a.dims[j + 1][1] * a.dims[j + 1][2];
        float* pc = lds + a.off[j + 1];
        tail_smooth(p, b, nx, ny, nz, a.h[j], a.pre, a.dhx[j], a.dy[j], a.omega, j == 0 ? a.red_done : 0);
        tail_restrict(p, b, nx, ny, nz, a.h[j], pc + nc, pc, a.dcx[j], a.dcy[j]);
    }
    {
        const int j = a.nt - 1;
        const int nx = a.dims[j][0], ny = a.dims[j][1], nz = a.dims[j][2], n = nx * ny * nz;
        float* p = lds + a.off[j];
        tail_smooth(p, p + n, nx, ny, nz, a.h[j], a.coarse_iters, a.dhx[j], a.dy[j], 1.f);
    }
    for (int j = a.nt - 2; j >= 0; j--) {
        const int nx = a.dims[j][0], ny = a.dims[j][1], nz = a.dims[j][2], n = nx * ny * nz;
        float* p = lds + a.off[j];
        tail_prolong(p, lds + a.off[j + 1], nx, ny, nz, a.dx[j], a.dy[j]);
        tail_smooth(p, p + n, nx, ny, nz, a.h[j], a.post, a.dhx[j], a.dy[j], a.omega);
    }
    for (int i = threadIdx.x; i < n0; i += TAIL_THREADS) {
        const int t = fdiv(i, a.dx[0]), x = i - t * nx0;
        a.p_top[t * 2 * nh0 + (x & 1) * nh0 + (x >> 1)] = p0[i];
    }
}

inline size_t level_floats(const int* d) { return (size_t)2 * ((d[0] + 1) / 2) * d[1] * d[2]; }

}  // namespace

size_t mg_level_floats(const GpuMG& mg, int l) { return level_floats(mg.dims[l]); }

int mg_build(GpuMG& mg, int nx, int ny, int nz, float h, int max_levels, std::string& err) {
    // levels halve while every dimension is even and the smallest exceeds 6 (sim.js:80-87)
    mg.nlev = 0;
    int x = nx, y = ny, z = nz;
    float hh = h;
    while (true) {
        if (mg.nlev >= MG_MAX_LEVELS) { err = "too many multigrid levels"; return FS_EINVAL; }
        const int l = mg.nlev++;
        mg.dims[l][0] = x; mg.dims[l][1] = y; mg.dims[l][2] = z; mg.h[l] = hh;
        const size_t bytes = level_floats(mg.dims[l]) * sizeof(float);
        if (hipMalloc(&mg.p[l], bytes) != hipSuccess || hipMalloc(&mg.b[l], bytes) != hipSuccess) {
            err = "hipMalloc failed for multigrid level"; return FS_ENOMEM;
        }
        if (hipMemset(mg.p[l], 0, bytes) != hipSuccess || hipMemset(mg.b[l], 0, bytes) != hipSuccess) {
            err = "hipMemset failed"; return FS_EDEVICE;
        }
        if (x % 2 || y % 2 || z % 2 || std::min(x, std::min(y, z)) <= 6) break;
        if (max_levels > 0 && mg.nlev >= max_levels) break;
        x /= 2; y /= 2; z /= 2; hh *= 2.f;
    }
    // coarse sweeps to cut the smoothest coarse mode to ~5 % (sim.js:88-92)
    const int cx = mg.dims[mg.nlev - 1][0], cy = mg.dims[mg.nlev - 1][1], cz = mg.dims[mg.nlev - 1][2];
    const double mu = (std::cos(M_PI / cx) + std::cos(M_PI / (2.0 * cy)) + std::cos(M_PI / cz)) / 3.0;
    mg.coarse_iters = std::max(8, std::min(64, (int)std::ceil(std::log(0.05) / std::log(mu * mu))));
    // the coarse tail: the largest suffix of levels whose p and b fit in one workgroup's LDS
    mg.tail_from = mg.nlev;
    {
        size_t bytes = 0;
        for (int l = mg.nlev - 1; l >= 0; l--) {
            bytes += (size_t)mg.dims[l][0] * mg.dims[l][1] * mg.dims[l][2] * 2 * sizeof(float);
            if (bytes > (size_t)TAIL_LDS_BYTES) break;
            mg.tail_from = l;
        }
    }
    const size_t n0 = level_floats(mg.dims[0]);
    const dim3 g = grid3((nx + 1) / 2, ny, nz);
    mg.npart = (int)(g.x * g.y * g.z);
    if (hipMalloc(&mg.x64, n0 * sizeof(double)) != hipSuccess || hipMalloc(&mg.bkeep, n0 * sizeof(float)) != hipSuccess ||
        hipMalloc(&mg.p0post, n0 * sizeof(float)) != hipSuccess ||
        hipMalloc(&mg.partials, (size_t)mg.npart * 2 * sizeof(double)) != hipSuccess ||
        hipMalloc(&mg.st, sizeof(RefineState)) != hipSuccess) {
        err = "hipMalloc failed for refinement buffers"; return FS_ENOMEM;
    }
    if (hipMemset(mg.x64, 0, n0 * sizeof(double)) != hipSuccess || hipMemset(mg.bkeep, 0, n0 * sizeof(float)) != hipSuccess ||
        hipMemset(mg.st, 0, sizeof(RefineState)) != hipSuccess) {
        err = "hipMemset failed"; return FS_EDEVICE;
    }
    return FS_OK;
}

void mg_free(GpuMG& mg) {
    for (int l = 0; l < mg.nlev; l++) {
        (void)hipFree(mg.p[l]); (void)hipFree(mg.b[l]);
        mg.p[l] = mg.b[l] = nullptr;
    }
    if (mg.x64) (void)hipFree(mg.x64);
    if (mg.bkeep) (void)hipFree(mg.bkeep);
    if (mg.p0post) (void)hipFree(mg.p0post);
    if (mg.partials) (void)hipFree(mg.partials);
    if (mg.st) (void)hipFree(mg.st);
    mg.x64 = nullptr; mg.bkeep = nullptr; mg.p0post = nullptr; mg.partials = nullptr; mg.st = nullptr;
    direct_free(mg.ds);
    mg.ds = nullptr;
    mg.direct = 0;
    mg.nlev = 0;
}

int mg_set_direct(GpuMG& mg, int want, std::string& err) {
    direct_free(mg.ds);
    mg.ds = nullptr;
    mg.direct = 0;
    if (want < 0) return FS_OK;
    int lv = 0;
    if (want == 0) {
        for (int l = 1; l < mg.nlev && !lv; l++)
            if (direct_fits(mg.dims[l][0], mg.dims[l][1], mg.dims[l][2])) lv = l;
        if (!lv) return FS_OK;                         // no level fits: the plain V-cycle
    } else {
        if (want >= mg.nlev || !direct_fits(mg.dims[want][0], mg.dims[want][1], mg.dims[want][2])) {
            err = "direct_level: that level cannot be solved exactly (x a multiple of 16, y of 4, z of 8, each "
                  "<= 128, at most 2^19 cells)";
            return FS_EINVAL;
        }
        lv = want;
    }
    if (int rc = direct_build(&mg.ds, mg.dims[lv][0], mg.dims[lv][1], mg.dims[lv][2], mg.h[lv], err)) return rc;
    mg.direct = lv;
    return FS_OK;
}

// `fin`: the first pass also runs that pending reduction (and is not skippable)
void mg_smooth(GpuMG& mg, hipStream_t s, int l, int iters, const int* skip, double* xupd, bool first_red_done,
               const FinArgs* fin, float* pl) {   // sim.js:576-586
    if (!skip) skip = &mg.st->never;
    if (!pl) pl = mg.p[l];
    const int nx = mg.dims[l][0], ny = mg.dims[l][1], nz = mg.dims[l][2];
    const dim3 g = grid3((nx + 1) / 2, ny, nz);
    const float om = l == mg.nlev - 1 ? 1.f : mg.omega;   // the coarsest level's sweeps stay plain GS
    for (int k = 0; k < iters; k++)
        for (int color = 0; color < 2; color++) {
            if (k == 0 && color == 0 && first_red_done) continue;
            double* xu = (xupd && k == iters - 1 && color == 1) ? xupd : nullptr;
            if (fin) {
                k_smooth<true><<<g, RB, 0, s>>>(pl, mg.b[l], nx, ny, nz, mg.h[l], color, mg.gate ? mg.gate : &mg.st->never, xu,
                                                om, nullptr,
                                                *fin);
                fin = nullptr;
            } else {
                k_smooth<false><<<g, RB, 0, s>>>(pl, mg.b[l], nx, ny, nz, mg.h[l], color, skip, xu, om, nullptr,
                                                 FinArgs{});
            }
        }
    if (fin) k_finalize<<<1, FIN, 0, s>>>(fin->st, fin->partials, fin->npart, fin->tol, fin->max_cycles);
}

static void launch_restrict(GpuMG& mg, hipStream_t s, int l, const int* skip, bool red_first, const float* pl) {
    const int nx = mg.dims[l][0], ny = mg.dims[l][1], nz = mg.dims[l][2];
    k_restrict<<<grid3(nx / 2, ny / 2, nz / 2), RB, 0, s>>>(pl, mg.b[l], nx, ny, nz, mg.h[l], mg.b[l + 1], mg.p[l + 1], skip,
                                                            mg.h[l + 1], mg.omega, red_first ? 1 : 0);
}
static void launch_prolong(GpuMG& mg, hipStream_t s, int l, const int* skip, float* pl) {
    const int nx = mg.dims[l][0], ny = mg.dims[l][1], nz = mg.dims[l][2];
    k_prolong<<<grid3(nx / 2, ny / 2, nz / 2), RB, 0, s>>>(pl, mg.p[l + 1], nx, ny, nz, skip);
}

static void launch_tail(GpuMG& mg, hipStream_t s, const int* skip, bool red_done) {
    TailArgs a;
    a.red_done = red_done ? 1 : 0;
    const int T = mg.tail_from;
    a.nt = mg.nlev - T;
    int off = 0;
    for (int j = 0; j < a.nt; j++) {
        for (int k = 0; k < 3; k++) a.dims[j][k] = mg.dims[T + j][k];
        a.h[j] = mg.h[T + j];
        a.dhx[j] = make_fastdiv((mg.dims[T + j][0] + 1) / 2);
        a.dx[j] = make_fastdiv(mg.dims[T + j][0]);
        a.dy[j] = make_fastdiv(mg.dims[T + j][1]);
        a.dcx[j] = make_fastdiv(std::max(mg.dims[T + j][0] / 2, 1));
        a.dcy[j] = make_fastdiv(std::max(mg.dims[T + j][1] / 2, 1));
        a.off[j] = off;
        off += 2 * mg.dims[T + j][0] * mg.dims[T + j][1] * mg.dims[T + j][2];
    }
    a.coarse_iters = mg.coarse_iters;
    a.omega = mg.omega;
    a.pre = mg.pre;
    a.post = mg.post;
    a.p_top = mg.p[T];
    a.b_top = mg.b[T];
    k_mg_tail<<<1, TAIL_THREADS, off * sizeof(float), s>>>(a, skip);
}

// Whether the restriction into level l + 1 writes that level's first colour-0 pass: it
// must be a pre-smoothed level that is not the coarsest (whose sweeps stay plain GS).
static bool fuse_red_into(const GpuMG& mg, int lc, int T) {
    if (mg.pre < 1 || lc >= mg.nlev - 1) return false;
    if (mg.direct > 0 && lc >= mg.direct) return false;   // the exact solve takes that level unsmoothed
    return lc <= T;                                    // a per-level pass, or the tail's top level
}

bool mg_vcycle(GpuMG& mg, hipStream_t s, const int* skip, double* xupd, bool red0_done,
               const FinArgs* fin, float* p0, float* p0alt, float** p0out) {   // sim.js:588-599
    if (!skip) skip = &mg.st->never;
    if (!p0) p0 = mg.p[0];
    if (p0out) *p0out = p0;
    auto lp = [&](int l) { return l == 0 ? p0 : mg.p[l]; };   // level l's iterate
    const int n = mg.nlev;
    const int D = mg.direct;                            // > 0: levels below D are replaced by its exact solve
    const int T = D > 0 ? n : mg.use_tail ? mg.tail_from : n;   // levels T.. run in the single-workgroup tail
    const int last = D > 0 ? D : std::min(T, n - 1);    // levels 0..last-1 smooth and restrict
    if (fin && (T == 0 || last == 0 || mg.pre < 1)) {   // the cycle does not open with a level-0 pass: a kernel of its own
        k_finalize<<<1, FIN, 0, s>>>(fin->st, fin->partials, fin->npart, fin->tol, fin->max_cycles);
        fin = nullptr;
    }
    if (T == 0) {                                       // the whole hierarchy fits one workgroup's LDS (level 0 in mg.p[0])
        launch_tail(mg, s, skip, false);
        return false;
    }
    bool red_done = red0_done;
    for (int l = 0; l < last; l++) {
        mg_smooth(mg, s, l, mg.pre, skip, nullptr, red_done, l == 0 ? fin : nullptr, lp(l));
        red_done = fuse_red_into(mg, l + 1, T);
        launch_restrict(mg, s, l, skip, red_done, lp(l));
    }
    if (D > 0) direct_solve(mg.ds, s, mg.b[D], mg.p[D], skip);
    else if (T < n) launch_tail(mg, s, skip, red_done);
    else mg_smooth(mg, s, n - 1, mg.coarse_iters, skip);
    for (int l = last - 1; l >= 0; l--) {
        // coarse levels: one launch less; fine levels: cheaper apart.  Over-relaxation reads the
        // colour-1 cells' own prolonged values, which the fused pass leaves unformed: not fused then.
        if (mg.fuse_prolong && l >= 2 && mg.omega == 1.f && mg.post == 2) {
            // prolongation + colour 0, then colour 1, then one more full sweep (sim.js:468-495, 576-586)
            const int nx = mg.dims[l][0], ny = mg.dims[l][1], nz = mg.dims[l][2];
            const dim3 g = grid3(nx / 2, ny, nz);
            k_smooth_prolong<<<g, RB, 0, s>>>(mg.p[l], mg.b[l], mg.p[l + 1], nx, ny, nz, mg.h[l], skip);
            k_smooth<false><<<g, RB, 0, s>>>(mg.p[l], mg.b[l], nx, ny, nz, mg.h[l], 1, skip, nullptr, 1.f, nullptr, FinArgs{});
            mg_smooth(mg, s, l, 1, skip, l == 0 ? xupd : nullptr);
        } else if (mg.fuse_prolong && mg.post >= 1) {
            // prolongation + the first colour-0 pass from one thread per coarse cell; the first colour-1 pass
            // prolongs its own cells; then the remaining sweeps (the last adds into the FP64 iterate)
            const int nx = mg.dims[l][0], ny = mg.dims[l][1], nz = mg.dims[l][2];
            // a refinement's level-0 correction is scratch: the post-smoothing runs on a second array, which
            // the prolongation fills with every child (the colour-1 pass then reads its own prolonged values)
            float* const post = l == 0 ? p0alt : nullptr;
            k_prolong_red<<<grid3(nx / 2, ny / 2, nz / 2), RB, 0, s>>>(lp(l), mg.b[l], mg.p[l + 1], nx, ny, nz, mg.h[l],
                                                                      mg.omega, skip, post);
            double* const xu = l == 0 ? xupd : nullptr;
            float* const pw = post ? post : lp(l);
            if (post && p0out) *p0out = post;
            k_smooth<false><<<grid3((nx + 1) / 2, ny, nz), RB, 0, s>>>(pw, mg.b[l], nx, ny, nz, mg.h[l], 1, skip,
                                                                       mg.post == 1 ? xu : nullptr, mg.omega,
                                                                       post ? nullptr : mg.p[l + 1], FinArgs{});
            if (mg.post > 1) mg_smooth(mg, s, l, mg.post - 1, skip, xu, false, nullptr, pw);
        } else {
            launch_prolong(mg, s, l, skip, lp(l));
            mg_smooth(mg, s, l, mg.post, skip, l == 0 ? xupd : nullptr, false, nullptr, lp(l));
        }
    }
    return xupd != nullptr && last > 0;
}

// In the refinement every V-cycle starts from p = 0 on level 0 (the residual kernel zeroes
// it), so the residual kernel can write level 0's first colour-0 pass as well.
static bool fuse_red0(const GpuMG& mg) {
    const int T = mg.use_tail ? mg.tail_from : mg.nlev;
    return mg.pre >= 1 && mg.nlev > 1 && T >= 1;
}

// The reduction of each residual runs in block 0 of the next V-cycle's first kernel, which is
// therefore launched unconditionally: once the solve has stopped it still smooths level 0's
// V-cycle iterate p (a scratch array the refinement rewrites before its next use) and never
// touches the FP64 iterate; every later kernel of the iteration tests the stop flag as before.
void mg_refine_begin(GpuMG& mg, hipStream_t s, double tol, int max_cycles) {
    const int nx = mg.dims[0][0], ny = mg.dims[0][1], nz = mg.dims[0][2];
    if (max_cycles > MG_MAX_REFINE) max_cycles = MG_MAX_REFINE;
    const bool red = fuse_red0(mg);
    k_resid64<double><<<grid3((nx + 1) / 2, ny, nz), RB, 0, s>>>(mg.x64, mg.bkeep, nx, ny, nz, (double)mg.h[0], mg.b[0], mg.p[0],
                                                                  mg.partials, mg.st, 1, mg.gate ? mg.gate : &mg.st->never, mg.h[0],
                                                                  mg.omega, red ? 1 : 0);
}

// The correction's V-cycles iterate on e (the first from e = 0, whose first colour-0 pass the residual kernel
// wrote); level 0's iterate moves between mg.p[0] and mg.p0post (each cycle's prolongation writes the other);
// the last cycle's last pass adds e into the FP64 iterate.  form_rhs: the correction's right-hand side is formed
// first (the previous residual kernel only reduced the norms); keep_rhs: the closing residual kernel also writes
// the next correction's right-hand side (otherwise only its norms, for a solve that usually stops there);
// close = false: no closing residual here.
void mg_refine_iter(GpuMG& mg, hipStream_t s, double tol, int max_cycles, int cycles, bool form_rhs, bool keep_rhs,
                    bool close) {
    const int nx = mg.dims[0][0], ny = mg.dims[0][1], nz = mg.dims[0][2];
    if (max_cycles > MG_MAX_REFINE) max_cycles = MG_MAX_REFINE;
    if (cycles < 1) cycles = 1;
    const int* done = mg.gate ? &mg.st->skipv : &mg.st->done;   // (a step enqueued ahead: also held)
    const bool red = fuse_red0(mg);
    const FinArgs fin{mg.st, mg.partials, mg.npart, max_cycles, tol};
    if (form_rhs)   // r = b - A x as the V-cycles' right-hand side, e = 0 (first colour-0 pass written); no norms
        k_resid64<double><<<grid3((nx + 1) / 2, ny, nz), RB, 0, s>>>(mg.x64, mg.bkeep, nx, ny, nz, (double)mg.h[0], mg.b[0],
                                                                      mg.p[0], nullptr, mg.st, 0, done, mg.h[0], mg.omega,
                                                                      red ? 1 : 0);
    float* cur = mg.p[0];
    bool fused = false;
    for (int c = 0; c < cycles; c++) {
        float* const alt = cur == mg.p[0] ? mg.p0post : mg.p[0];
        float* out = cur;
        fused = mg_vcycle(mg, s, done, c == cycles - 1 ? mg.x64 : nullptr, c == 0 && red, c == 0 ? &fin : nullptr, cur, alt,
                          &out);
        cur = out;
    }
    if (!fused) {
        const int n = (int)level_floats(mg.dims[0]);
        k_update<<<(n + RB - 1) / RB, RB, 0, s>>>(mg.x64, cur, n, done);
    }
    if (!close) return;
    k_resid64<double><<<grid3((nx + 1) / 2, ny, nz), RB, 0, s>>>(mg.x64, mg.bkeep, nx, ny, nz, (double)mg.h[0],
                                                                  keep_rhs ? mg.b[0] : nullptr, keep_rhs ? mg.p[0] : nullptr,
                                                                  mg.partials, mg.st, 0, done, mg.h[0], mg.omega, red ? 1 : 0,
                                                                  cycles);
}

void mg_state_ready(GpuMG& mg, hipStream_t s, double tol, int max_cycles, int* host, const int* gate, bool set_hold) {
    if (max_cycles > MG_MAX_REFINE) max_cycles = MG_MAX_REFINE;
    k_finalize<<<1, FIN, 0, s>>>(mg.st, mg.partials, mg.npart, tol, max_cycles, host, gate, set_hold ? 1 : 0);
}

void mg_refine(GpuMG& mg, hipStream_t s, double tol, int max_cycles) {
    if (max_cycles > MG_MAX_REFINE) max_cycles = MG_MAX_REFINE;
    mg_refine_begin(mg, s, tol, max_cycles);
    for (int k = 1; k <= max_cycles; k++) mg_refine_iter(mg, s, tol, max_cycles);
    mg_state_ready(mg, s, tol, max_cycles);
}

// FP64 defect correction x <- x + S(b - A x) with the exact separable inner solve S (fs_spectral.hip)
// in place of a V-cycle.  The residual kernel writes r as the right-hand side; p is not touched (the
// solve overwrites all of it).  Once the stop flag is set, the correction and the residual of later
// iterations are skipped as for V-cycles.
void mg_spectral_begin(GpuMG& mg, hipStream_t s, double tol, int max_cycles) {
    const int nx = mg.dims[0][0], ny = mg.dims[0][1], nz = mg.dims[0][2];
    if (max_cycles > MG_MAX_REFINE) max_cycles = MG_MAX_REFINE;
    k_resid64<double><<<grid3((nx + 1) / 2, ny, nz), RB, 0, s>>>(mg.x64, mg.bkeep, nx, ny, nz, (double)mg.h[0], mg.b[0], nullptr,
                                                                  mg.partials, mg.st, 1, &mg.st->never);
    k_finalize<<<1, FIN, 0, s>>>(mg.st, mg.partials, mg.npart, tol, max_cycles);
}

int mg_spectral_iter(GpuMG& mg, Spectral* sp, hipStream_t s, double tol, int max_cycles) {
    const int nx = mg.dims[0][0], ny = mg.dims[0][1], nz = mg.dims[0][2];
    if (max_cycles > MG_MAX_REFINE) max_cycles = MG_MAX_REFINE;
    int* done = &mg.st->done;
    if (int rc = spectral_solve(sp, s, mg.b[0], mg.p[0])) return rc;
    const int n = (int)level_floats(mg.dims[0]);
    k_update<<<(n + RB - 1) / RB, RB, 0, s>>>(mg.x64, mg.p[0], n, done);
    k_resid64<double><<<grid3((nx + 1) / 2, ny, nz), RB, 0, s>>>(mg.x64, mg.bkeep, nx, ny, nz, (double)mg.h[0], mg.b[0], nullptr,
                                                                  mg.partials, mg.st, 0, done);
    k_finalize<<<1, FIN, 0, s>>>(mg.st, mg.partials, mg.npart, tol, max_cycles);
    return FS_OK;
}

void mg_tol32_begin(GpuMG& mg, hipStream_t s, double tol, int max_cycles) {
    const int nx = mg.dims[0][0], ny = mg.dims[0][1], nz = mg.dims[0][2];
    if (max_cycles > MG_MAX_REFINE) max_cycles = MG_MAX_REFINE;
    k_resid64<float><<<grid3((nx + 1) / 2, ny, nz), RB, 0, s>>>(mg.p[0], mg.b[0], nx, ny, nz, (double)mg.h[0], nullptr, nullptr,
                                                                 mg.partials, mg.st, 1, &mg.st->never);
    k_finalize<<<1, FIN, 0, s>>>(mg.st, mg.partials, mg.npart, tol, max_cycles);
}

void mg_tol32_iter(GpuMG& mg, hipStream_t s, double tol, int max_cycles) {
    const int nx = mg.dims[0][0], ny = mg.dims[0][1], nz = mg.dims[0][2];
    if (max_cycles > MG_MAX_REFINE) max_cycles = MG_MAX_REFINE;
    int* done = &mg.st->done;
    mg_vcycle(mg, s, done);
    k_resid64<float><<<grid3((nx + 1) / 2, ny, nz), RB, 0, s>>>(mg.p[0], mg.b[0], nx, ny, nz, (double)mg.h[0], nullptr, nullptr,
                                                                 mg.partials, mg.st, 0, done);
    k_finalize<<<1, FIN, 0, s>>>(mg.st, mg.partials, mg.npart, tol, max_cycles);
}

void mg_solve_tol32(GpuMG& mg, hipStream_t s, double tol, int max_cycles) {
    if (max_cycles > MG_MAX_REFINE) max_cycles = MG_MAX_REFINE;
    mg_tol32_begin(mg, s, tol, max_cycles);
    for (int k = 1; k <= max_cycles; k++) mg_tol32_iter(mg, s, tol, max_cycles);
}

void mg_residual_of_p(GpuMG& mg, hipStream_t s) {
    const int nx = mg.dims[0][0], ny = mg.dims[0][1], nz = mg.dims[0][2];
    k_resid64<float><<<grid3((nx + 1) / 2, ny, nz), RB, 0, s>>>(mg.p[0], mg.b[0], nx, ny, nz, (double)mg.h[0], nullptr, nullptr,
                                                                 mg.partials, mg.st, 1, &mg.st->never);
    k_finalize<<<1, FIN, 0, s>>>(mg.st, mg.partials, mg.npart, 0.0, 0);
}
