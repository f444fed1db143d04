// Field diagnostics on the device: the FIRE.stats and FIRE.profile reductions
// (src/main.js:268-294, 322-355) and a per-step time series (heat release, peak
// temperature, flame tip and the axis temperature profile) for the flame-height
// and puffing-frequency estimators.
//
// Reductions are deterministic: each block reduces a fixed range of cells in a
// fixed order and writes a partial; partials are combined in index order (on the
// device for the per-step record, on the host for stats/profile).  Sums are FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "fs_codec.h"
#include "fs_internal.h"

namespace {

constexpr int DB = 256;              // threads per block
constexpr int STATS_CELLS = 8192;    // cells per block of the stats reduction

// ---- FIRE.stats: per channel min / max / sum / non-finite count / first argmax -----
struct ChanAcc {
    float mn, mx;
    int arg;          // linear index of the first maximum, -1 if none
    int bad;
    double sum;
};
__device__ __forceinline__ void chan_merge(ChanAcc& a, const ChanAcc& b) {   // b covers later cells than a
    a.mn = fminf(a.mn, b.mn);
    if (b.arg >= 0 && (a.arg < 0 || b.mx > a.mx || (b.mx == a.mx && b.arg < a.arg))) { a.mx = b.mx; a.arg = b.arg; }
    a.bad += b.bad;
    a.sum += b.sum;
}

template <class S>
__global__ __launch_bounds__(DB) void k_stats(const typename S::raw* __restrict__ f, int n, double* __restrict__ part) {
    __shared__ ChanAcc sh[4][DB];
    ChanAcc acc[4];
#pragma unroll
    for (int c = 0; c < 4; c++) acc[c] = ChanAcc{INFINITY, -INFINITY, -1, 0, 0.0};
    const int lo = blockIdx.x * STATS_CELLS, hi = min(n, lo + STATS_CELLS);
    // thread t walks cells lo + t, lo + t + DB, ... in increasing order ("first" maximum)
    for (int i = lo + threadIdx.x; i < hi; i += DB) {
        const float4 v4 = Codec<S>::ld(f, i);
        const float v[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
        for (int c = 0; c < 4; c++) {
            if (!isfinite(v[c])) { acc[c].bad++; continue; }
            acc[c].mn = fminf(acc[c].mn, v[c]);
            if (acc[c].arg < 0 || v[c] > acc[c].mx) { acc[c].mx = v[c]; acc[c].arg = i; }
            acc[c].sum += (double)v[c];
        }
    }
#pragma unroll
    for (int c = 0; c < 4; c++) sh[c][threadIdx.x] = acc[c];
    __syncthreads();
    // fixed-order tree; the argmax rule (larger value, then smaller index) is order-free
    for (int s = DB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
#pragma unroll
            for (int c = 0; c < 4; c++) chan_merge(sh[c][threadIdx.x], sh[c][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 4) {
        const ChanAcc& a = sh[threadIdx.x][0];
        double* o = part + ((size_t)blockIdx.x * 4 + threadIdx.x) * 5;
        o[0] = a.mn; o[1] = a.mx; o[2] = a.sum; o[3] = a.bad; o[4] = a.arg;
    }
}

// ---- FIRE.profile per y-plane, and the per-step record ----------------------------
// Per plane j (one block): Tmax, sum of T over the 5x5 centre columns, their count,
// |soot| max, |fuel| max, cells with T > 1000 K, heat release Q (W), and the max T
// of the centre columns.  The reference decodes the stored halves without their
// sign in FIRE.profile; so does this (|value|), and it computes in double as JS does.
template <class S>
__global__ __launch_bounds__(DB) void k_planes(const typename S::raw* __restrict__ Sf, int nx, int ny, int nz, double V,
                                               double* __restrict__ out) {
    __shared__ double sh[8][DB];
    const int j = blockIdx.x;
    double Tmax = 0.0, Tc = 0.0, nC = 0.0, soot = 0.0, fuel = 0.0, hot = 0.0, Q = 0.0, axT = 0.0;
    for (int t = threadIdx.x; t < nx * nz; t += DB) {
        const int k = t / nx, i = t - k * nx;
        const float4 s = Codec<S>::ld(Sf, (k * ny + j) * nx + i);
        const double T = 293.0 + 1000.0 * fabs((double)s.x);
        Tmax = fmax(Tmax, T);
        soot = fmax(soot, fabs((double)s.z));
        fuel = fmax(fuel, fabs((double)s.y));
        if (T > 1000.0) hot += 1.0;
        if (fabs(i - nx / 2.0) < 3.0 && fabs(k - nz / 2.0) < 3.0) { Tc += T; nC += 1.0; axT = fmax(axT, T); }
        const double q = fabs((double)s.w) * 1e4;
        if (q > 0.0) Q += 354800.0 * q / T * V;
    }
    const double v[8] = {Tmax, Tc, nC, soot, fuel, hot, Q, axT};
#pragma unroll
    for (int c = 0; c < 8; c++) sh[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int s = DB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const int a = threadIdx.x, b = threadIdx.x + s;
            sh[0][a] = fmax(sh[0][a], sh[0][b]);
            sh[1][a] += sh[1][b];
            sh[2][a] += sh[2][b];
            sh[3][a] = fmax(sh[3][a], sh[3][b]);
            sh[4][a] = fmax(sh[4][a], sh[4][b]);
            sh[5][a] += sh[5][b];
            sh[6][a] += sh[6][b];
            sh[7][a] = fmax(sh[7][a], sh[7][b]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 8) out[(size_t)j * 8 + threadIdx.x] = sh[threadIdx.x][0];
}

// Combines the stats partials of nb blocks into the 20 results of fs_stats, exactly as the host
// combine does (blocks in index order: the earlier maximum wins ties; sums added in order).
__global__ __launch_bounds__(64) void k_stats_combine(const double* __restrict__ part, int nb, double ncell,
                                                      double* __restrict__ out) {
    const int c = threadIdx.x;
    if (c >= 4) return;
    double mn = INFINITY, mx = -INFINITY, sum = 0.0, bad = 0.0, arg = -1.0;
    for (int b = 0; b < nb; b++) {
        const double* q = part + ((size_t)b * 4 + c) * 5;
        mn = fmin(mn, q[0]);
        if (q[4] >= 0 && (arg < 0 || q[1] > mx)) { mx = q[1]; arg = q[4]; }
        sum += q[2];
        bad += q[3];
    }
    out[c * 5 + 0] = mn;
    out[c * 5 + 1] = mx;
    out[c * 5 + 2] = sum / ncell;
    out[c * 5 + 3] = bad;
    out[c * 5 + 4] = arg;
}

// One record per step into a ring of `cap` records of (4 + ny) doubles:
// time (s, after the step), Q (W), Tmax (K), flame tip (m: top of the highest plane
// whose max T exceeds t_flame, 0 if none), then the centre-column max T per plane.
__global__ __launch_bounds__(DB) void k_record(const DevParams* __restrict__ P, const double* __restrict__ planes, int ny,
                                               double dx, double t_flame, double* __restrict__ ring, int cap,
                                               int* __restrict__ count) {
    const int n = *count;
    double* rec = ring + (size_t)(n % cap) * (4 + ny);
    for (int j = threadIdx.x; j < ny; j += DB) rec[4 + j] = planes[(size_t)j * 8 + 7];
    if (threadIdx.x == 0) {
        double Q = 0.0, Tm = 0.0;
        int tip = -1;
        for (int j = 0; j < ny; j++) {
            Q += planes[(size_t)j * 8 + 6];
            Tm = fmax(Tm, planes[(size_t)j * 8]);
            if (planes[(size_t)j * 8] > t_flame) tip = j;
        }
        rec[0] = P->time;
        rec[1] = Q;
        rec[2] = Tm;
        rec[3] = tip >= 0 ? (tip + 1) * dx : 0.0;
    }
    __syncthreads();                 // every thread has read *count
    if (threadIdx.x == 0) *count = n + 1;
}

}  // namespace

// ---- fs_stats without allocation ----------------------------------------------------
// The partials and the combined result live in this file-scope device array, one copy per GPU for
// the life of the process.  Host code never uses the array's name as an address: the device
// address is looked up once per device (hipGetSymbolAddress, with that device current) and kept.
constexpr int STATS_NOALLOC_BLOCKS = (FS_STATS_NOALLOC_MAX_CELLS + STATS_CELLS - 1) / STATS_CELLS;
constexpr size_t STATS_SCRATCH_DOUBLES = (size_t)STATS_NOALLOC_BLOCKS * 20 + 32;
static __device__ __align__(256) double scratch[STATS_SCRATCH_DOUBLES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};
}  // namespace

// Device address of `scratch` on the current device (looked up on its first use there).
static int scratch_address(int device, double** out) {
    if (device < 0 || device >= MAX_DEVICES) {
        fs_set_error("device ordinal out of range for the allocation-free statistics");
        return FS_EINVAL;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fs_set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return FS_EDEVICE;
        }
        g_scratch_ptr[device] = (double*)p;
    }
    *out = g_scratch_ptr[device];
    return FS_OK;
}

int stats_noalloc(int prec, hipStream_t s, int device, const void* field, size_t n, double out[20]) {
    if (n > (size_t)FS_STATS_NOALLOC_MAX_CELLS) {
        fs_set_error("fs_stats_noalloc: the field has " + std::to_string(n) + " cells, more than the " +
                     std::to_string(FS_STATS_NOALLOC_MAX_CELLS) + " its fixed scratch holds; use fs_stats");
        return FS_EINVAL;
    }
    double* ptr = nullptr;
    if (int rc = scratch_address(device, &ptr)) return rc;
    const int nb = stats_blocks(n);
    double* res = ptr + (size_t)STATS_NOALLOC_BLOCKS * 20;     // the 20 combined values
    launch_stats(prec, s, field, n, ptr);
    k_stats_combine<<<1, 64, 0, s>>>(ptr, nb, (double)n, res);
    FS_HIP_CHECK(hipGetLastError());
    FS_HIP_CHECK(hipMemcpyAsync(out, res, 20 * sizeof(double), hipMemcpyDeviceToHost, s));
    FS_HIP_CHECK(hipStreamSynchronize(s));
    return FS_OK;
}

#define DISPATCH_PREC(prec, BODY)                                  \
    do {                                                           \
        if ((prec) == FS_PREC_FP32) { typedef StF32 S; BODY; }     \
        else if ((prec) == FS_PREC_FP16) { typedef StF16 S; BODY; } \
        else { typedef StBF16 S; BODY; }                           \
    } while (0)

int stats_blocks(size_t n) { return (int)((n + STATS_CELLS - 1) / STATS_CELLS); }

void launch_stats(int prec, hipStream_t s, const void* field, size_t n, double* part) {
    DISPATCH_PREC(prec, (k_stats<S><<<stats_blocks(n), DB, 0, s>>>((const typename S::raw*)field, (int)n, part)));
}

void launch_planes(int prec, hipStream_t s, const void* S_field, int nx, int ny, int nz, double V, double* out) {
    DISPATCH_PREC(prec, (k_planes<S><<<ny, DB, 0, s>>>((const typename S::raw*)S_field, nx, ny, nz, V, out)));
}

void launch_record(hipStream_t s, const DevParams* P, const double* planes, int ny, double dx, double t_flame,
                   double* ring, int cap, int* count) {
    k_record<<<1, DB, 0, s>>>(P, planes, ny, dx, t_flame, ring, cap, count);
}
