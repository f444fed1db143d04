// FP64 mass and energy budget of the step, reduced on the device.
//
// With the budget on, the step's source kernel runs in a tracing instantiation: the same cell
// update (forces_cell, bit-identical fields) records the fuel fraction and the temperature at
// the boundaries of its sub-models, and each cell contributes FP64 differences of those f32
// values: fuel added by the log bands / coal bed / bursts, injected by the spots and the floor
// flux, burned, dissipated, removed in solids and NaN resets, and the rounding of the stored
// value; heat added by the source pilots, released by burning, lost to radiation and to
// entrainment cooling, and rounding.  Each term is a difference of consecutive values of one
// cell, so the terms telescope: the stored totals after a step equal the totals after
// advection plus the terms, up to the FP64 summation.  The change by advection (boundary
// inflow and outflow, and the semi-Lagrangian scheme's non-conservation) is the difference of
// the totals before the step and after advection.
//
// Reductions are deterministic: a block reduces its 256 cells in a fixed tree and writes one
// partial per term; the finalize kernel sums partials in index order.
#include <hip/hip_runtime.h>

#include "fs_codec.h"
#include "fs_internal.h"
#include "fs_math.h"
#include "fs_forces_cell.h"

namespace {

constexpr int BB = 256;   // threads per block (the per-cell kernels' 64 x 4 tiles)
constexpr double CP_AIR = 1005.0;   // J / (kg K), with rho0 = FS_RHO0: the thermal energy of a cell

// per-cell terms of the tracing source kernel
enum {
    T_YIN, T_YSRC, T_YFLUX, T_YBURN, T_YDISS, T_YREM, T_YROUND, T_YST,
    T_UIN, T_USRC, T_UHEAT, T_URAD, T_UCOOL, T_UREM, T_UROUND, T_UST,
    T_PHI, T_BAD, T_COUNT
};

// Sums v[0..N) over the block (256 threads = 4 waves) in a fixed order; thread t < N of the
// block ends with the sum of term t in out[t] (LDS).
template <int N>
__device__ __forceinline__ void block_sum(double* v, double (*sh)[N]) {
#pragma unroll
    for (int k = 0; k < N; k++) {
        double x = v[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        v[k] = x;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < N; k++) sh[wave][k] = v[k];
    }
    __syncthreads();
}

template <class S>
__global__ __launch_bounds__(BB) void k_forces_budget(const DevParams* __restrict__ P, Dims g,
                                                      const typename S::raw* __restrict__ velY,
                                                      const typename S::raw* __restrict__ sY,
                                                      const typename S::raw* __restrict__ s2Y,
                                                      const typename S::raw* __restrict__ vort,
                                                      const typename S::raw* __restrict__ geo,
                                                      const half4_t* __restrict__ noise, const half4_t* __restrict__ curlT,
                                                      int nn, const fs_spot* __restrict__ spots,
                                                      const float* __restrict__ faceFlux, typename S::raw* __restrict__ velOut,
                                                      typename S::raw* __restrict__ sOut, typename S::raw* __restrict__ s2Out,
                                                      double* __restrict__ part) {
    __shared__ double sh[4][T_COUNT];
    double v[T_COUNT];
#pragma unroll
    for (int k = 0; k < T_COUNT; k++) v[k] = 0.0;
    Cell c;
    if (cell_of(g, c)) {
        const ForcesCell k = forces_setup<S>(P, g, c, velY, geo, noise, curlT, nn);
        const float4 v4 = Codec<S>::ld(velY, c.i);
        const float4 s4 = Codec<S>::ld(sY, c.i), t4 = Codec<S>::ld(s2Y, c.i);
        const float4 w = Codec<S>::ld(vort, c.i);
        const float gx = ldc<S>(vort, g, c.x + 1, c.y, c.z).w - ldc<S>(vort, g, c.x - 1, c.y, c.z).w;
        const float gy = ldc<S>(vort, g, c.x, c.y + 1, c.z).w - ldc<S>(vort, g, c.x, c.y - 1, c.z).w;
        const float gz = ldc<S>(vort, g, c.x, c.y, c.z + 1).w - ldc<S>(vort, g, c.x, c.y, c.z - 1).w;
        float vx = v4.x, vy = v4.y, vz = v4.z, phi;
        float s[4] = {s4.x, s4.y, s4.z, s4.w}, s2[4] = {t4.x, t4.y, t4.z, t4.w};
        const float wv[3] = {w.x, w.y, w.z};
        const bool face = P->has_face_flux != 0;
        const float fv = (face && c.y == 0) ? faceFlux[c.z * g.nx + c.x] : 0.f;
        CellTrace t;
        const float lap[3] = {k.lapx, k.lapy, k.lapz};
        forces_cell<float, CellTrace>(P->f, P->d, k, spots, P->n_spots, -1, nullptr, face, fv, vx, vy, vz, s, s2, wv, gx, gy,
                                      gz, lap, phi, &t);
        Codec<S>::st(velOut, c.i, make_float4(vx, vy, vz, phi));
        Codec<S>::st(sOut, c.i, make_float4(s[0], s[1], s[2], s[3]));
        Codec<S>::st(s2Out, c.i, make_float4(s2[0], s2[1], s2[2], s2[3]));
        const double yst = Codec<S>::rt(t.y_out), thst = Codec<S>::rt(t.th_out);
        const double ph = Codec<S>::rt(phi);
        double d[T_COUNT];
        d[T_YIN] = t.y_in;
        d[T_YSRC] = (double)t.y_src - t.y_in;
        d[T_YFLUX] = (double)t.y_flux - t.y_src;
        d[T_YBURN] = (double)t.y_flux - t.y_burn;
        d[T_YDISS] = (double)t.y_burn - t.y_diss;
        d[T_YREM] = (double)t.y_diss - t.y_out;
        d[T_YROUND] = yst - t.y_out;
        d[T_YST] = yst;
        d[T_UIN] = 1000.0 * t.th_in;
        d[T_USRC] = 1000.0 * ((double)t.th_flux - t.th_in);
        d[T_UHEAT] = (double)t.tk_burn - t.t0;
        d[T_URAD] = (double)t.tk_burn - t.tk_rad;
        d[T_UCOOL] = (double)t.tk_rad - t.tk_cool;
        d[T_UREM] = 1000.0 * ((double)t.th_cool - t.th_out);
        d[T_UROUND] = ((double)t.t0 - 293.0 - 1000.0 * (double)t.th_flux) +
                      (1000.0 * (double)t.th_cool - ((double)t.tk_cool - 293.0)) + 1000.0 * (thst - t.th_out);
        d[T_UST] = 1000.0 * thst;
        d[T_PHI] = ph;
        d[T_BAD] = 0.0;
        bool fin = true;
#pragma unroll
        for (int q = 0; q < T_BAD; q++) fin = fin && isfinite(d[q]);
        if (fin) {
#pragma unroll
            for (int q = 0; q < T_BAD; q++) v[q] = d[q];
        } else {
            v[T_BAD] = 1.0;          // a cell whose trace is not finite is counted, not summed
        }
    }
    block_sum<T_COUNT>(v, sh);
    if (threadIdx.x < T_COUNT) {
        const int k = threadIdx.x;
        part[(size_t)blockIdx.x * T_COUNT + k] = ((sh[0][k] + sh[1][k]) + sh[2][k]) + sh[3][k];
    }
}

// fuel and theta totals of a state (at enable): 2 terms per block
template <class S>
__global__ __launch_bounds__(BB) void k_budget_state(Dims g, const typename S::raw* __restrict__ Sf, double* __restrict__ part) {
    __shared__ double sh[4][2];
    double v[2] = {0.0, 0.0};
    Cell c;
    if (cell_of(g, c)) {
        const float4 s = Codec<S>::ld(Sf, c.i);
        v[0] = s.y;
        v[1] = 1000.0 * s.x;
    }
    block_sum<2>(v, sh);
    if (threadIdx.x < 2) {
        const int k = threadIdx.x;
        part[(size_t)blockIdx.x * 2 + k] = ((sh[0][k] + sh[1][k]) + sh[2][k]) + sh[3][k];
    }
}

// sum of the level-0 right-hand side (the divergence minus the expansion), parity-split rows
__global__ __launch_bounds__(BB) void k_budget_div(Dims g, const float* __restrict__ b, double* __restrict__ part) {
    __shared__ double sh[4][1];
    double v[1] = {0.0};
    Cell c;
    if (cell_of(g, c)) v[0] = b[rb_index(c.x, c.y, c.z, g.nx, g.ny)];
    block_sum<1>(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = ((sh[0][0] + sh[1][0]) + sh[2][0]) + sh[3][0];
}

// Sums partials part[b * stride + k], b = 0 .. nb-1, for k < nterms, into out[k]: thread t adds
// blocks t, t + 256, ... in order, then a fixed tree.
__device__ void sum_partials(const double* __restrict__ part, int nb, int stride, int nterms, double* out) {
    __shared__ double sh[BB];
    for (int k = 0; k < nterms; k++) {
        double x = 0.0;
        for (int b = threadIdx.x; b < nb; b += BB) x += part[(size_t)b * stride + k];
        sh[threadIdx.x] = x;
        __syncthreads();
        for (int s = BB / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[k] = sh[0];
        __syncthreads();
    }
}

// Record layout: the FS_BUDGET_* indices of include/firesim.h.
__global__ __launch_bounds__(BB) void k_budget_init(const DevParams* __restrict__ P, const double* __restrict__ part, int nb,
                                                    double* __restrict__ rec) {
    __shared__ double tot[2];
    sum_partials(part, nb, 2, 2, tot);
    if (threadIdx.x < FS_BUDGET_N) rec[threadIdx.x] = 0.0;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double dx = P->f[P_dx], V = dx * dx * dx;
        const double kg = FS_RHO0 * V, J = FS_RHO0 * CP_AIR * V;
        rec[FS_BUDGET_FUEL0] = rec[FS_BUDGET_FUEL] = kg * tot[0];
        rec[FS_BUDGET_HEAT0] = rec[FS_BUDGET_HEAT] = J * tot[1];
    }
}

__global__ __launch_bounds__(BB) void k_budget_step(const DevParams* __restrict__ P, Dims g, const double* __restrict__ part,
                                                    const double* __restrict__ dpart, const fs_spot* __restrict__ spots,
                                                    const float* __restrict__ faceFlux, double* __restrict__ rec) {
    __shared__ double tot[T_COUNT + 2];
    sum_partials(part, g.nblocks, T_COUNT, T_COUNT, tot);
    sum_partials(dpart, g.nblocks, 1, 1, tot + T_COUNT);
    // the floor flux over the face, in index order
    __shared__ double fsh[BB];
    double fx = 0.0;
    if (P->has_face_flux)
        for (int i = threadIdx.x; i < g.nx * g.nz; i += BB) fx += faceFlux[i];
    fsh[threadIdx.x] = fx;
    __syncthreads();
    for (int s = BB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) fsh[threadIdx.x] += fsh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double dx = P->f[P_dx], V = dx * dx * dx;
        const double kg = FS_RHO0 * V, J = FS_RHO0 * CP_AIR * V;
        const double dt = (float)P->dt;
        rec[FS_BUDGET_STEPS] += 1.0;
        rec[FS_BUDGET_FUEL_TRANSPORT] += kg * tot[T_YIN] - rec[FS_BUDGET_FUEL];
        rec[FS_BUDGET_FUEL_SOURCES] += kg * tot[T_YSRC];
        rec[FS_BUDGET_FUEL_INJECTED] += kg * tot[T_YFLUX];
        rec[FS_BUDGET_FUEL_BURNED] += kg * tot[T_YBURN];
        rec[FS_BUDGET_FUEL_DISSIPATED] += kg * tot[T_YDISS];
        rec[FS_BUDGET_FUEL_REMOVED] += kg * tot[T_YREM];
        rec[FS_BUDGET_FUEL_ROUNDING] += kg * tot[T_YROUND];
        rec[FS_BUDGET_FUEL] = kg * tot[T_YST];
        rec[FS_BUDGET_HEAT_TRANSPORT] += J * tot[T_UIN] - rec[FS_BUDGET_HEAT];
        rec[FS_BUDGET_HEAT_SOURCES] += J * tot[T_USRC];
        rec[FS_BUDGET_HEAT_RELEASED] += J * tot[T_UHEAT];
        rec[FS_BUDGET_HEAT_RADIATED] += J * tot[T_URAD];
        rec[FS_BUDGET_HEAT_ENTRAINED] += J * tot[T_UCOOL];
        rec[FS_BUDGET_HEAT_REMOVED] += J * tot[T_UREM];
        rec[FS_BUDGET_HEAT_ROUNDING] += J * tot[T_UROUND];
        rec[FS_BUDGET_HEAT] = J * tot[T_UST];
        double inj = 0.0;
        for (int q = 0; q < P->n_spots; q++) inj += (double)spots[q].flux * dt;
        inj += fsh[0] * dx * dx * dt;
        rec[FS_BUDGET_FUEL_EXPECTED] += inj;
        rec[FS_BUDGET_EXPANSION] = V * tot[T_PHI];
        rec[FS_BUDGET_DIVERGENCE] = V * tot[T_COUNT];
        rec[FS_BUDGET_BAD_CELLS] += tot[T_BAD];
    }
}

}  // namespace

#define DISPATCH_PREC(prec, BODY)                                  \
    do {                                                           \
        if ((prec) == FS_PREC_FP32) { typedef StF32 S; BODY; }     \
        else if ((prec) == FS_PREC_FP16) { typedef StF16 S; BODY; } \
        else { typedef StBF16 S; BODY; }                           \
    } while (0)
#define RAW(p) (typename S::raw*)(p)
#define CRAW(p) (const typename S::raw*)(p)

size_t budget_partials(const Dims& g) { return (size_t)g.nblocks * (T_COUNT + 1); }

template <class S>
static void forces_budget_t(const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F, double* part) {
    k_forces_budget<S><<<g.nblocks, BB, 0, s>>>(P, g, CRAW(F.velY), CRAW(F.SY), CRAW(F.S2Y), CRAW(F.vort), CRAW(F.geo),
                                                F.noise, F.curl, F.noise_n, F.spots, F.face_flux, RAW(F.velX), RAW(F.S),
                                                RAW(F.S2), part);
}
void launch_forces_budget(int prec, const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F, double* part) {
    DISPATCH_PREC(prec, forces_budget_t<S>(g, s, P, F, part));
}
void launch_budget_step(const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F, const float* b, double* part,
                        double* rec) {
    double* dpart = part + (size_t)g.nblocks * T_COUNT;
    k_budget_div<<<g.nblocks, BB, 0, s>>>(g, b, dpart);
    k_budget_step<<<1, BB, 0, s>>>(P, g, part, dpart, F.spots, F.face_flux, rec);
}
void launch_budget_init(int prec, const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F, double* part,
                        double* rec) {
    DISPATCH_PREC(prec, (k_budget_state<S><<<g.nblocks, BB, 0, s>>>(g, CRAW(F.S), part)));
    k_budget_init<<<1, BB, 0, s>>>(P, part, g.nblocks, rec);
}
