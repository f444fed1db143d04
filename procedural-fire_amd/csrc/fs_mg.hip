// Geometric multigrid for the pressure Poisson equation on gfx950.
//
// Operator (src/sim.js:416-495): cell-centred 7-point Laplacian, p = 0 on the
// x/z faces and the top (ghost = -p), Neumann floor (ghost = +p).  V(2,2) with
// red-black Gauss-Seidel (omega = 1), residual fused into the 8-child restriction,
// mirrored trilinear prolongation, `coarse_iters` sweeps on the coarsest level.
//
// Mixed precision: FP32 V-cycles on the correction inside FP64 defect correction
// (x <- x + V(b - A x)).  The residual and its norms are FP64; the stopping test
// runs on the device and later kernels of the same step early-exit on a device
// flag, so a whole step stays free of host round trips (and graph-capturable).
// Norm reductions are two-pass with a fixed order: bitwise reproducible.
#include <hip/hip_runtime.h>

#include <cmath>

#include "fs_internal.h"

namespace {

constexpr int RB = 256;   // threads per block of the reduction kernels

__device__ __forceinline__ bool skipped(const int* skip) { return skip != nullptr && *skip != 0; }

// one colour pass of red-black GS (sim.js:422-437); threads cover the half grid
__global__ __launch_bounds__(256) void k_smooth(float* __restrict__ p, const float* __restrict__ b, int nx, int ny, int nz,
                                                float h, int color, const int* __restrict__ skip) {
    if (skipped(skip)) return;
    const int nhx = (nx + 1) >> 1;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nhx * ny * nz) return;
    const int hx = i % nhx;
    const int t = i / nhx;
    const int y = t % ny, z = t / ny;
    const int x = hx * 2 + ((y + z + color) & 1);
    if (x >= nx) return;
    const int sy = nx, sz = nx * ny;
    const int c = (z * ny + y) * nx + x;
    float sum = 0.f, diag = 6.f;
    if (x > 0) sum += p[c - 1]; else diag += 1.f;
    if (x < nx - 1) sum += p[c + 1]; else diag += 1.f;
    if (z > 0) sum += p[c - sz]; else diag += 1.f;
    if (z < nz - 1) sum += p[c + sz]; else diag += 1.f;
    if (y > 0) sum += p[c - sy]; else diag -= 1.f;
    if (y < ny - 1) sum += p[c + sy]; else diag += 1.f;
    p[c] = (sum - h * h * b[c]) / diag;
}

// residual of the 8 children averaged onto the coarse cell; coarse p = 0 (sim.js:446-466)
__global__ __launch_bounds__(256) void k_restrict(const float* __restrict__ p, const float* __restrict__ b, int nx, int ny, int nz,
                                                  float h, float* __restrict__ bc, float* __restrict__ pc,
                                                  const int* __restrict__ skip) {
    if (skipped(skip)) return;
    const int cx = nx >> 1, cy = ny >> 1, cz = nz >> 1;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= cx * cy * cz) return;
    const int xc = i % cx;
    const int t = i / cx;
    const int yc = t % cy, zc = t / cy;
    const float ih2 = 1.f / (h * h);
    const int sy = nx, sz = nx * ny;
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int x = xc * 2 + (k & 1), y = yc * 2 + ((k >> 1) & 1), z = zc * 2 + ((k >> 2) & 1);
        const int c = (z * ny + y) * nx + x;
        const float me = p[c];
        const float pxm = x > 0 ? p[c - 1] : -me;
        const float pxp = x < nx - 1 ? p[c + 1] : -me;
        const float pym = y > 0 ? p[c - sy] : me;
        const float pyp = y < ny - 1 ? p[c + sy] : -me;
        const float pzm = z > 0 ? p[c - sz] : -me;
        const float pzp = z < nz - 1 ? p[c + sz] : -me;
        const float lap = (pxm + pxp + pym + pyp + pzm + pzp - 6.f * me) * ih2;
        acc += b[c] - lap;
    }
    bc[i] = acc * 0.125f;
    pc[i] = 0.f;
}

// p_fine += mirrored trilinear interpolation of the coarse p (sim.js:473-495)
__global__ __launch_bounds__(256) void k_prolong(float* __restrict__ p, const float* __restrict__ pcoarse, int nx, int ny, int nz,
                                                 const int* __restrict__ skip) {
    if (skipped(skip)) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nx * ny * nz) return;
    const int x = i % nx;
    const int t = i / nx;
    const int y = t % ny, z = t / ny;
    const int cx = nx >> 1, cy = ny >> 1, cz = nz >> 1;
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
        float val = pcoarse[(sz * cy + sy) * cx + sx];
        if (qx < 0 || qx >= cx) val = -val;
        if (qz < 0 || qz >= cz) val = -val;
        if (qy >= cy) val = -val;
        const float wx = ox ? fx : 1.f - fx, wy = oy ? fy : 1.f - fy, wz = oz ? fz : 1.f - fz;
        acc += val * wx * wy * wz;
    }
    p[i] = p[i] + acc;
}

// FP64 residual r = b - A x of the level-0 operator.  Writes the per-block
// (r.r, b.b) partial sums; with `rhs_out` set it also stores r (f32) as the
// V-cycle right-hand side and zeroes the V-cycle iterate.
template <class X>
__global__ __launch_bounds__(RB) void k_resid64(const X* __restrict__ x, const float* __restrict__ b, int nx, int ny, int nz,
                                                double h, float* __restrict__ rhs_out, float* __restrict__ p_out,
                                                double* __restrict__ partials, const int* __restrict__ skip) {
    if (skipped(skip)) return;
    __shared__ double sr[RB], sb[RB];
    const int n = nx * ny * nz;
    const int i = blockIdx.x * RB + threadIdx.x;
    double rr = 0.0, bb = 0.0;
    if (i < n) {
        const int xx = i % nx;
        const int t = i / nx;
        const int y = t % ny, z = t / ny;
        const int sy = nx, sz = nx * ny;
        const double me = (double)x[i];
        double sum = 0.0, diag = 6.0;
        if (xx > 0) sum += (double)x[i - 1]; else diag += 1.0;
        if (xx < nx - 1) sum += (double)x[i + 1]; else diag += 1.0;
        if (z > 0) sum += (double)x[i - sz]; else diag += 1.0;
        if (z < nz - 1) sum += (double)x[i + sz]; else diag += 1.0;
        if (y > 0) sum += (double)x[i - sy]; else diag -= 1.0;
        if (y < ny - 1) sum += (double)x[i + sy]; else diag += 1.0;
        const double bv = (double)b[i];
        const double r = bv - (sum - diag * me) * (1.0 / (h * h));
        rr = r * r; bb = bv * bv;
        if (rhs_out) { rhs_out[i] = (float)r; p_out[i] = 0.f; }
    }
    sr[threadIdx.x] = rr; sb[threadIdx.x] = bb;
    __syncthreads();
    for (int s = RB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { sr[threadIdx.x] += sr[threadIdx.x + s]; sb[threadIdx.x] += sb[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { partials[2 * blockIdx.x] = sr[0]; partials[2 * blockIdx.x + 1] = sb[0]; }
}

// Sums the partials in a fixed order.  phase 0: start of a solve; phase k >= 1:
// after V-cycle k.  Sets st->done once rel <= tol or k == max_cycles.
__global__ __launch_bounds__(RB) void k_finalize(RefineState* __restrict__ st, const double* __restrict__ partials, int npart,
                                                 double tol, int max_cycles, int phase) {
    if (phase > 0 && st->done) return;
    __shared__ double sr[RB], sb[RB];
    double rr = 0.0, bb = 0.0;
    for (int j = threadIdx.x; j < npart; j += RB) { rr += partials[2 * j]; bb += partials[2 * j + 1]; }
    sr[threadIdx.x] = rr; sb[threadIdx.x] = bb;
    __syncthreads();
    for (int s = RB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { sr[threadIdx.x] += sr[threadIdx.x + s]; sb[threadIdx.x] += sb[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double rel = sb[0] > 0.0 ? sqrt(sr[0] / sb[0]) : 0.0;
        st->rel[phase] = rel;
        st->used = phase;
        st->bnorm2 = sb[0];
        st->done = (rel <= tol || phase >= max_cycles) ? 1 : 0;
    }
}

// x += e (the V-cycle's correction), FP64
__global__ __launch_bounds__(256) void k_update(double* __restrict__ x, const float* __restrict__ e, int n,
                                                const int* __restrict__ skip) {
    if (skipped(skip)) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) x[i] = x[i] + (double)e[i];
}

inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

int mg_build(GpuMG& mg, int nx, int ny, int nz, float h, int max_levels, std::string& err) {
    // levels halve while every dimension is even and the smallest exceeds 6 (sim.js:80-87)
    mg.nlev = 0;
    int x = nx, y = ny, z = nz;
    float hh = h;
    while (true) {
        if (mg.nlev >= MG_MAX_LEVELS) { err = "too many multigrid levels"; return FS_EINVAL; }
        const int l = mg.nlev++;
        mg.dims[l][0] = x; mg.dims[l][1] = y; mg.dims[l][2] = z; mg.h[l] = hh;
        const size_t bytes = (size_t)x * y * z * sizeof(float);
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
    const long n0 = (long)nx * ny * nz;
    mg.npart = (int)((n0 + RB - 1) / RB);
    if (hipMalloc(&mg.x64, n0 * sizeof(double)) != hipSuccess || hipMalloc(&mg.bkeep, n0 * sizeof(float)) != hipSuccess ||
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
    for (int l = 0; l < mg.nlev; l++) { (void)hipFree(mg.p[l]); (void)hipFree(mg.b[l]); mg.p[l] = mg.b[l] = nullptr; }
    if (mg.x64) (void)hipFree(mg.x64);
    if (mg.bkeep) (void)hipFree(mg.bkeep);
    if (mg.partials) (void)hipFree(mg.partials);
    if (mg.st) (void)hipFree(mg.st);
    mg.x64 = nullptr; mg.bkeep = nullptr; mg.partials = nullptr; mg.st = nullptr;
    mg.nlev = 0;
}

void mg_smooth(GpuMG& mg, hipStream_t s, int l, int iters, const int* skip) {   // sim.js:576-586
    const int nx = mg.dims[l][0], ny = mg.dims[l][1], nz = mg.dims[l][2];
    const long half = (long)((nx + 1) / 2) * ny * nz;
    for (int k = 0; k < iters; k++)
        for (int color = 0; color < 2; color++)
            k_smooth<<<nblk(half), 256, 0, s>>>(mg.p[l], mg.b[l], nx, ny, nz, mg.h[l], color, skip);
}

void mg_vcycle(GpuMG& mg, hipStream_t s, const int* skip) {   // sim.js:588-599
    const int n = mg.nlev;
    for (int l = 0; l < n - 1; l++) {
        mg_smooth(mg, s, l, 2, skip);
        const int nx = mg.dims[l][0], ny = mg.dims[l][1], nz = mg.dims[l][2];
        const long nc = (long)(nx / 2) * (ny / 2) * (nz / 2);
        k_restrict<<<nblk(nc), 256, 0, s>>>(mg.p[l], mg.b[l], nx, ny, nz, mg.h[l], mg.b[l + 1], mg.p[l + 1], skip);
    }
    mg_smooth(mg, s, n - 1, mg.coarse_iters, skip);
    for (int l = n - 2; l >= 0; l--) {
        const int nx = mg.dims[l][0], ny = mg.dims[l][1], nz = mg.dims[l][2];
        k_prolong<<<nblk((long)nx * ny * nz), 256, 0, s>>>(mg.p[l], mg.p[l + 1], nx, ny, nz, skip);
        mg_smooth(mg, s, l, 2, skip);
    }
}

void mg_refine(GpuMG& mg, hipStream_t s, double tol, int max_cycles) {
    const int nx = mg.dims[0][0], ny = mg.dims[0][1], nz = mg.dims[0][2];
    const long n = (long)nx * ny * nz;
    const double h = (double)mg.h[0];
    if (max_cycles > MG_MAX_REFINE) max_cycles = MG_MAX_REFINE;
    int* done = &mg.st->done;
    k_resid64<double><<<mg.npart, RB, 0, s>>>(mg.x64, mg.bkeep, nx, ny, nz, h, mg.b[0], mg.p[0], mg.partials, nullptr);
    k_finalize<<<1, RB, 0, s>>>(mg.st, mg.partials, mg.npart, tol, max_cycles, 0);
    for (int k = 1; k <= max_cycles; k++) {
        mg_vcycle(mg, s, done);
        k_update<<<nblk(n), 256, 0, s>>>(mg.x64, mg.p[0], (int)n, done);
        k_resid64<double><<<mg.npart, RB, 0, s>>>(mg.x64, mg.bkeep, nx, ny, nz, h, mg.b[0], mg.p[0], mg.partials, done);
        k_finalize<<<1, RB, 0, s>>>(mg.st, mg.partials, mg.npart, tol, max_cycles, k);
    }
}

void mg_residual_of_p(GpuMG& mg, hipStream_t s) {
    const int nx = mg.dims[0][0], ny = mg.dims[0][1], nz = mg.dims[0][2];
    k_resid64<float><<<mg.npart, RB, 0, s>>>(mg.p[0], mg.b[0], nx, ny, nz, (double)mg.h[0], nullptr, nullptr, mg.partials, nullptr);
    k_finalize<<<1, RB, 0, s>>>(mg.st, mg.partials, mg.npart, 0.0, 0, 0);
}
