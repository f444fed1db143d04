// TEST INFRASTRUCTURE ONLY — CPU oracle of one FireSim.step() (src/sim.js:627-648).
//
// A restatement, kernel by kernel, of the WGSL in src/sim.js (plus src/noise.js,
// src/shaders/lib.js, src/scene.js:13-42,92-98), templated on
//   R  math type        (float = reference semantics, double = FP64 error oracle)
//   S  field storage    (half_t = rgba16float as in the reference, float, bf16_t, double)
//   M  multigrid storage(float = r32float as in the reference, double)
// It is the checker for the HIP library and the timed CPU baseline; only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The product
// path never links or calls it.
//
// Conventions fixed here where the reference leaves them to the WebGPU stack
// (all "parity unpinned" against real WGSL: no WebGPU implementation runs here to compare with):
//   * trilinear sampling uses exact f32 weights (hardware samplers use ~8-bit
//     fixed-point weights; research/webgpu.md:103);
//   * clamp/min/max return the non-NaN operand (IEEE maxNum/minNum);
//   * stores to rgba16float round to nearest even.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include <omp.h>
#include <array>

#include "../include/firesim.h"
#include "oracle_common.h"

using namespace orc;

namespace {

// ---- SIM_FIELDS / SIM_DEFAULTS (src/sim.js:14-44) ---------------------------
const char* kNames[FS_NUM_PARAMS] = {
    "nx", "ny", "nz", "dx", "dt", "time", "frame", "seed",
    "ox", "oy", "oz", "nLogs",
    "fuelSrc", "bedFuel", "srcTemp", "bedTemp", "srcVel", "srcFreq", "srcSpeed", "srcThresh",
    "ignT", "kMix", "rs", "heatOut", "expansion", "sootForm", "sootOx", "radCool",
    "coolRate", "sootDissip", "smokeDissip", "vitDecay", "fuelDissip", "buoyancy", "vortEps", "turbAmp",
    "turbFreq", "turbRise", "turbEvolve", "windX", "windZ", "gust", "intensity", "velDissip",
    "bedRadius", "logSrc", "burstX", "burstY", "burstZ", "burstAmt", "smolder", "sponge", "hotR0", "hotR1", "hotY1",
    "visc", "lamAge", "anchorFreq", "srcFloor", "anchorGain",
};
enum {
    P_nx, P_ny, P_nz, P_dx, P_dt, P_time, P_frame, P_seed, P_ox, P_oy, P_oz, P_nLogs,
    P_fuelSrc, P_bedFuel, P_srcTemp, P_bedTemp, P_srcVel, P_srcFreq, P_srcSpeed, P_srcThresh,
    P_ignT, P_kMix, P_rs, P_heatOut, P_expansion, P_sootForm, P_sootOx, P_radCool,
    P_coolRate, P_sootDissip, P_smokeDissip, P_vitDecay, P_fuelDissip, P_buoyancy, P_vortEps, P_turbAmp,
    P_turbFreq, P_turbRise, P_turbEvolve, P_windX, P_windZ, P_gust, P_intensity, P_velDissip,
    P_bedRadius, P_logSrc, P_burstX, P_burstY, P_burstZ, P_burstAmt, P_smolder, P_sponge, P_hotR0, P_hotR1, P_hotY1,
    P_visc, P_lamAge, P_anchorFreq, P_srcFloor, P_anchorGain,
};
void set_defaults(float* p) {
    for (int i = 0; i < FS_NUM_PARAMS; i++) p[i] = 0.f;
    p[P_fuelSrc] = 0.13f; p[P_bedFuel] = 0.16f; p[P_srcTemp] = 0.75f; p[P_bedTemp] = 0.82f;
    p[P_srcVel] = 0.25f; p[P_srcFreq] = 35.0f; p[P_srcSpeed] = 0.06f; p[P_srcThresh] = 0.0f;
    p[P_ignT] = 550.f; p[P_kMix] = 30.f; p[P_rs] = 4.7f; p[P_heatOut] = 7700.f; p[P_expansion] = 0.7f;
    p[P_sootForm] = 25.f; p[P_sootOx] = 30.f; p[P_radCool] = 0.25f;
    p[P_coolRate] = 3.0f; p[P_sootDissip] = 0.4f; p[P_smokeDissip] = 0.3f; p[P_vitDecay] = 1.5f; p[P_fuelDissip] = 0.3f;
    p[P_buoyancy] = 26.8f;
    p[P_vortEps] = 8.f; p[P_turbAmp] = 0.8f; p[P_turbFreq] = 1.5f; p[P_turbRise] = 1.5f; p[P_turbEvolve] = 0.6f;
    p[P_windX] = 0.f; p[P_windZ] = 0.f; p[P_gust] = 0.f; p[P_intensity] = 1.f; p[P_velDissip] = 0.05f;
    p[P_bedRadius] = 0.22f; p[P_hotR0] = 0.12f; p[P_hotR1] = 0.45f; p[P_hotY1] = 0.5f; p[P_logSrc] = 1.f;
    p[P_burstX] = 0.f; p[P_burstY] = 0.f; p[P_burstZ] = 0.f; p[P_burstAmt] = 0.f; p[P_smolder] = 0.02f; p[P_sponge] = 1.f;
    p[P_visc] = 0.f; p[P_lamAge] = 0.f; p[P_anchorFreq] = 1.6f; p[P_srcFloor] = 0.25f; p[P_anchorGain] = 0.95f;
}

// ---- noise textures (src/noise.js:8-54, 69-80), always stored as rgba16float -
struct NoiseTex {
    int n = 0;
    std::vector<float> v;   // 4 channels, values exactly representable in binary16
};
inline int wrapi(int c, int per) { return ((c % per) + per) % per; }   // noise.js:11
void grad(int cx, int cy, int cz, int per, int ch, int seed, float g[3]) {  // noise.js:12-15
    float h[3];
    hash3f(wrapi(cx, per) + ch * 1013 + seed, wrapi(cy, per) + ch * 7919, wrapi(cz, per) + ch * 104729, h);
    float a = h[0] * 2.f - 1.f + 1e-5f, b = h[1] * 2.f - 1.f + 1e-5f, c = h[2] * 2.f - 1.f + 1e-5f;
    float l = std::sqrt(a * a + b * b + c * c);
    g[0] = a / l; g[1] = b / l; g[2] = c / l;
}
float tnoise(float px, float py, float pz, int per, int ch, int seed) {   // noise.js:16-28
    float flx = std::floor(px), fly = std::floor(py), flz = std::floor(pz);
    int ix = (int)flx, iy = (int)fly, iz = (int)flz;
    float fx = px - flx, fy = py - fly, fz = pz - flz;
    float ux = fx * fx * fx * (fx * (fx * 6.f - 15.f) + 10.f);
    float uy = fy * fy * fy * (fy * (fy * 6.f - 15.f) + 10.f);
    float uz = fz * fz * fz * (fz * (fz * 6.f - 15.f) + 10.f);
    float acc[8];
    for (int k = 0; k < 8; k++) {
        int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
        float g[3]; grad(ix + ox, iy + oy, iz + oz, per, ch, seed, g);
        acc[k] = g[0] * (fx - (float)ox) + g[1] * (fy - (float)oy) + g[2] * (fz - (float)oz);
    }
    float x0 = mix(acc[0], acc[1], ux), x1 = mix(acc[2], acc[3], ux);
    float x2 = mix(acc[4], acc[5], ux), x3 = mix(acc[6], acc[7], ux);
    return mix(mix(x0, x1, uy), mix(x2, x3, uy), uz);
}
float tfbm(float px, float py, float pz, int ch, int seed) {   // noise.js:29-36
    float s = 0.f, a = 0.5f; int per = 4;
    for (int o = 0; o < 5; o++) {
        float fp = (float)per;
        s += a * tnoise(px * fp, py * fp, pz * fp, per, ch, seed);
        per *= 2; a *= 0.5f;
    }
    return s * 1.5f;
}
float worley(float px, float py, float pz, int per, int seed) {   // noise.js:37-47
    float fp = (float)per;
    float qx = px * fp, qy = py * fp, qz = pz * fp;
    float flx = std::floor(qx), fly = std::floor(qy), flz = std::floor(qz);
    int ix = (int)flx, iy = (int)fly, iz = (int)flz;
    float fx = qx - flx, fy = qy - fly, fz = qz - flz;
    float d = 9.f;
    for (int k = 0; k < 27; k++) {
        int ox = k % 3 - 1, oy = (k / 3) % 3 - 1, oz = k / 9 - 1;
        float h[3];
        hash3f(wrapi(ix + ox, per) + 77, wrapi(iy + oy, per) + 33, wrapi(iz + oz, per) + seed, h);
        float a = (float)ox + h[0] - fx, b = (float)oy + h[1] - fy, c = (float)oz + h[2] - fz;
        d = std::fmin(d, std::sqrt(a * a + b * b + c * c));
    }
    return d;
}
void bake_noise(NoiseTex& t, int n, int seed) {   // noise.js:49-54
    t.n = n; t.v.assign((size_t)n * n * n * 4, 0.f);
#pragma omp parallel for schedule(dynamic)
    for (int z = 0; z < n; z++)
        for (int y = 0; y < n; y++)
            for (int x = 0; x < n; x++) {
                float px = ((float)x + 0.5f) / (float)n, py = ((float)y + 0.5f) / (float)n, pz = ((float)z + 0.5f) / (float)n;
                float a = tfbm(px, py, pz, 0, seed), b = tfbm(px, py, pz, 1, seed), c = tfbm(px, py, pz, 2, seed);
                float w = 0.65f * worley(px, py, pz, 8, seed) + 0.35f * worley(px, py, pz, 16, seed);
                float* o = &t.v[(((size_t)z * n + y) * n + x) * 4];
                o[0] = f16_to_f32(f32_to_f16(a * 0.5f + 0.5f));
                o[1] = f16_to_f32(f32_to_f16(b * 0.5f + 0.5f));
                o[2] = f16_to_f32(f32_to_f16(c * 0.5f + 0.5f));
                o[3] = f16_to_f32(f32_to_f16(wclamp(w, 0.f, 1.f)));
            }
}
void bake_curl(NoiseTex& out, const NoiseTex& src) {   // noise.js:69-80
    int n = src.n; out.n = n; out.v.assign(src.v.size(), 0.f);
    auto L = [&](int x, int y, int z, float o[3]) {
        x = (x + n) % n; y = (y + n) % n; z = (z + n) % n;
        const float* s = &src.v[(((size_t)z * n + y) * n + x) * 4];
        o[0] = s[0] * 2.f - 1.f; o[1] = s[1] * 2.f - 1.f; o[2] = s[2] * 2.f - 1.f;
    };
#pragma omp parallel for
    for (int z = 0; z < n; z++)
        for (int y = 0; y < n; y++)
            for (int x = 0; x < n; x++) {
                float a[3], b[3], dx[3], dy[3], dz[3];
                L(x + 1, y, z, a); L(x - 1, y, z, b); for (int i = 0; i < 3; i++) dx[i] = a[i] - b[i];
                L(x, y + 1, z, a); L(x, y - 1, z, b); for (int i = 0; i < 3; i++) dy[i] = a[i] - b[i];
                L(x, y, z + 1, a); L(x, y, z - 1, b); for (int i = 0; i < 3; i++) dz[i] = a[i] - b[i];
                float k = (float)n / 2.0f;
                float cx = (dy[2] - dz[1]) * k, cy = (dz[0] - dx[2]) * k, cz = (dx[1] - dy[0]) * k;
                float* o = &out.v[(((size_t)z * n + y) * n + x) * 4];
                o[0] = f16_to_f32(f32_to_f16(cx)); o[1] = f16_to_f32(f32_to_f16(cy));
                o[2] = f16_to_f32(f32_to_f16(cz));
                o[3] = f16_to_f32(f32_to_f16(std::sqrt(cx * cx + cy * cy + cz * cz)));
            }
}

// ---- campfire layout (src/scene.js:13-42, packLogs :92-98) -------------------
int build_campfire(int seed, float* out, int max_logs) {
    uint32_t s = (uint32_t)seed;
    auto R = [&]() { s = s * 1664525u + 1013904223u; return (double)s / 4294967296.0; };
    std::vector<std::vector<double>> logs;   // a.xyz, r, b.xyz, burn
    auto add = [&](double ax, double ay, double az, double bx, double by, double bz, double r, double burn) {
        logs.push_back({ax, ay, az, r, bx, by, bz, burn});
    };
    add(-0.40, 0.055, -0.17, 0.38, 0.065, -0.13, 0.062, 0.8);
    add(-0.36, 0.060, 0.15, 0.41, 0.050, 0.19, 0.058, 0.8);
    const int n = 5;
    for (int i = 0; i < n; i++) {
        double ang = ((double)i / n) * M_PI * 2 + 0.35 + (R() - 0.5) * 0.4;
        double rb = 0.30 + R() * 0.06;
        double bx = std::cos(ang) * rb, by = 0.03, bz = std::sin(ang) * rb;
        double topR = 0.035 + R() * 0.03;
        double ta = ang + M_PI + (R() - 0.5) * 0.8;
        double tx = std::cos(ta) * topR, ty = 0.36 + R() * 0.1, tz = std::sin(ta) * topR;
        double rr = 0.032 + R() * 0.012;
        add(bx, by, bz, tx, ty, tz, rr, 1.0);
    }
    add(-0.12, 0.035, 0.02, 0.10, 0.03, -0.05, 0.035, 1.0);
    for (int i = 0; i < max_logs * 8; i++) out[i] = 0.f;
    int cnt = std::min((int)logs.size(), max_logs);
    for (int i = 0; i < cnt; i++) for (int k = 0; k < 8; k++) out[i * 8 + k] = (float)logs[i][k];
    return cnt;
}

// ---- multigrid (src/sim.js:78-92, 416-495, 576-599) --------------------------
template <class M>
struct Level { int nx, ny, nz; double h; std::vector<M> p, b; };

template <class M>
struct Multigrid {
    std::vector<Level<M>> L;
    int coarse_iters = 8;
    M omega = 1;   // over-relaxation above the coarsest level (tolerance solves; the reference uses 1)
    int pre = 2, post = 2;   // sweeps per level (tolerance solves may set them; the reference uses V(2,2))

    void build(int x, int y, int z, double h, int max_levels) {   // sim.js:79-92
        L.clear();
        while (true) {
            Level<M> l; l.nx = x; l.ny = y; l.nz = z; l.h = h;
            l.p.assign((size_t)x * y * z, M(0)); l.b.assign((size_t)x * y * z, M(0));
            L.push_back(std::move(l));
            if (x % 2 || y % 2 || z % 2 || std::min(x, std::min(y, z)) <= 6) break;
            if (max_levels > 0 && (int)L.size() >= max_levels) break;
            x /= 2; y /= 2; z /= 2; h *= 2;
        }
        const Level<M>& c = L.back();
        double mu = (std::cos(M_PI / c.nx) + std::cos(M_PI / (2.0 * c.ny)) + std::cos(M_PI / c.nz)) / 3.0;
        coarse_iters = std::max(8, std::min(64, (int)std::ceil(std::log(0.05) / std::log(mu * mu))));
    }
    // one colour pass of red-black Gauss-Seidel (sim.js:422-437); with omega != 1 on a
    // level above the coarsest, p <- p + omega (p_GS - p)
    void smooth_color(int li, int color) {
        Level<M>& l = L[li];
        const M om = li + 1 < (int)L.size() ? omega : M(1);
        const int nx = l.nx, ny = l.ny, nz = l.nz;
        const M h2 = (M)l.h * (M)l.h;   // L.h is an f32 uniform in the reference
        M* p = l.p.data(); const M* b = l.b.data();
#pragma omp parallel for if (nx * ny * nz > 32768)
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = (y + z + color) & 1; x < nx; x += 2) {
                    size_t c = ((size_t)z * ny + y) * nx + x;
                    M sum = 0, diag = 6;
                    if (x > 0) sum += p[c - 1]; else diag += 1;
                    if (x < nx - 1) sum += p[c + 1]; else diag += 1;
                    if (z > 0) sum += p[c - (size_t)nx * ny]; else diag += 1;
                    if (z < nz - 1) sum += p[c + (size_t)nx * ny]; else diag += 1;
                    if (y > 0) sum += p[c - nx]; else diag -= 1;
                    if (y < ny - 1) sum += p[c + nx]; else diag += 1;
                    M pn = (sum - h2 * b[c]) / diag;
                    if (om != M(1)) pn = p[c] + om * (pn - p[c]);
                    p[c] = pn;
                }
    }
    void smooth(int li, int iters) {   // sim.js:576-586
        for (int k = 0; k < iters; k++) { smooth_color(li, 0); smooth_color(li, 1); }
    }
    // residual of 8 children averaged onto the coarse cell, coarse p = 0 (sim.js:446-466)
    void restrict_(int li) {
        Level<M>& f = L[li]; Level<M>& c = L[li + 1];
        const int nx = f.nx, ny = f.ny, nz = f.nz;
        const M ih2 = M(1) / ((M)f.h * (M)f.h);
        auto P = [&](int x, int y, int z, M me) -> M {
            if (x < 0 || z < 0 || x >= nx || z >= nz || y >= ny) return -me;
            if (y < 0) return me;
            return f.p[((size_t)z * ny + y) * nx + x];
        };
#pragma omp parallel for if (c.nx * c.ny * c.nz > 4096)
        for (int zc = 0; zc < c.nz; zc++)
            for (int yc = 0; yc < c.ny; yc++)
                for (int xc = 0; xc < c.nx; xc++) {
                    M acc = 0;
                    for (int k = 0; k < 8; k++) {
                        int x = xc * 2 + (k & 1), y = yc * 2 + ((k >> 1) & 1), z = zc * 2 + ((k >> 2) & 1);
                        size_t i = ((size_t)z * ny + y) * nx + x;
                        M me = f.p[i];
                        M lap = (P(x - 1, y, z, me) + P(x + 1, y, z, me) + P(x, y - 1, z, me) +
                                 P(x, y + 1, z, me) + P(x, y, z - 1, me) + P(x, y, z + 1, me) - M(6) * me) * ih2;
                        acc += f.b[i] - lap;
                    }
                    size_t ci = ((size_t)zc * c.ny + yc) * c.nx + xc;
                    c.b[ci] = acc * M(0.125);
                    c.p[ci] = 0;
                }
    }
    // cell-centred trilinear prolongation with odd (Dirichlet) / even (floor)
    // mirroring of out-of-range coarse nodes (sim.js:473-495)
    void prolong(int li) {
        Level<M>& f = L[li]; const Level<M>& c = L[li + 1];
        const int nx = f.nx, ny = f.ny, nz = f.nz;
        const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
#pragma omp parallel for if (nx * ny * nz > 32768)
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    M gx = ((M)x + M(0.5)) * M(0.5) - M(0.5), gy = ((M)y + M(0.5)) * M(0.5) - M(0.5),
                      gz = ((M)z + M(0.5)) * M(0.5) - M(0.5);
                    M flx = std::floor(gx), fly = std::floor(gy), flz = std::floor(gz);
                    int g0x = (int)flx, g0y = (int)fly, g0z = (int)flz;
                    M fx = gx - flx, fy = gy - fly, fz = gz - flz;
                    M acc = 0;
                    for (int k = 0; k < 8; k++) {
                        int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
                        int qx = g0x + ox, qy = g0y + oy, qz = g0z + oz;
                        int sx = std::min(std::max(qx, 0), cx - 1), sy = std::min(std::max(qy, 0), cy - 1),
                            sz = std::min(std::max(qz, 0), cz - 1);
                        M val = c.p[((size_t)sz * cy + sy) * cx + sx];
                        if (qx < 0 || qx >= cx) val = -val;
                        if (qz < 0 || qz >= cz) val = -val;
                        if (qy >= cy) val = -val;
                        M wx = ox ? fx : M(1) - fx, wy = oy ? fy : M(1) - fy, wz = oz ? fz : M(1) - fz;
                        acc += val * wx * wy * wz;
                    }
                    size_t i = ((size_t)z * ny + y) * nx + x;
                    f.p[i] = f.p[i] + acc;
                }
    }
    // ---- exact solve of one coarse level (the library's direct coarse solve, fs_direct.hip) ----
    // The level's operator is diagonalised by DST-II in x and z and DCT-IV in y; p = A^-1 b is
    // W = Tx b Tz^T per y-plane, W <- Ty^T diag(1/lambda) Ty W per (kx, kz) column, p = Tx^T W Tz.
    // Every value is one FMA chain over its contraction index in increasing order, from 0, and the
    // division uses ((lx + ly) + lz) / h^2, as the library's kernels do.
    int direct = 0;
    std::vector<M> dTx, dTy, dTz, dlx, dly, dlz;
    M dih2 = 1;
    static void basis(int n, int kind, std::vector<M>& T, std::vector<M>& lam) {
        const double pi = 3.14159265358979323846;
        T.assign((size_t)n * n, M(0));
        lam.assign(n, M(0));
        for (int k = 0; k < n; k++) {
            const double th = kind == 0 ? pi * (k + 1) / n : pi * (k + 0.5) / n;
            lam[k] = (M)(-(2.0 - 2.0 * std::cos(th)));
            const double nrm = (kind == 0 && k == n - 1) ? std::sqrt(1.0 / n) : std::sqrt(2.0 / n);
            for (int i = 0; i < n; i++)
                T[(size_t)k * n + i] = (M)(nrm * (kind == 0 ? std::sin(th * (i + 0.5)) : std::cos(th * (i + 0.5))));
        }
    }
    // the shapes the library's kernels take (x in chunks of 16, y in fours, z in eights, each <= 128,
    // at most 2^19 cells, the kernels' padded LDS rows within 160 KiB)
    static bool direct_fits(int X, int Y, int Z) {
        if (X < 16 || Y < 4 || Z < 8 || X % 16 || Y % 4 || Z % 8 || X > 128 || Z > 128 || Y > 128) return false;
        if ((size_t)X * Y * Z > ((size_t)1 << 19)) return false;
        if (X * Z > 8192 || X * X > 8192 || Z * Z > 8192 || Y * Y > 12288) return false;   // one register batch
        const size_t lim = 160 * 1024, m = std::max(X, Z) + 4;
        const size_t fwd = 4 * ((size_t)(X / 2) * (X + 4) + (size_t)Z * m + (size_t)(X / 2) * (Z + 4));
        const size_t yl = 4 * ((size_t)Y * (Y + 4) + 2 * (size_t)16 * (Y + 4));
        const size_t inv = 4 * ((size_t)X * m + (size_t)(Z / 2) * (Z + 4) + (size_t)(Z / 2) * (X + 4));
        return fwd <= lim && yl <= lim && inv <= lim;
    }
    // want: 0 automatic (the first level below the finest that fits), -1 none, k >= 1 level k
    void set_direct(int want) {
        direct = 0;
        if (want < 0) return;
        int lv = 0;
        if (want == 0) {
            for (int l = 1; l < (int)L.size() && !lv; l++)
                if (direct_fits(L[l].nx, L[l].ny, L[l].nz)) lv = l;
        } else if (want < (int)L.size() && direct_fits(L[want].nx, L[want].ny, L[want].nz)) {
            lv = want;
        }
        if (!lv) return;
        setup_direct(lv);
    }
    void setup_direct(int lv) {
        direct = lv;
        basis(L[lv].nx, 0, dTx, dlx);
        basis(L[lv].ny, 1, dTy, dly);
        basis(L[lv].nz, 0, dTz, dlz);
        dih2 = M(1) / ((M)L[lv].h * (M)L[lv].h);
    }
    void direct_solve() {
        Level<M>& l = L[direct];
        const int X = l.nx, Y = l.ny, Z = l.nz;
        std::vector<M> W((size_t)X * Y * Z);
        auto bi = [&](int x, int y, int z) { return ((size_t)z * Y + y) * X + x; };
#pragma omp parallel for
        for (int y = 0; y < Y; y++) {
            std::vector<M> C((size_t)X * Z);
            for (int kx = 0; kx < X; kx++)
                for (int z = 0; z < Z; z++) {
                    M acc = 0;
                    for (int x = 0; x < X; x++) acc = std::fma(dTx[(size_t)kx * X + x], l.b[bi(x, y, z)], acc);
                    C[(size_t)kx * Z + z] = acc;
                }
            for (int kx = 0; kx < X; kx++)
                for (int kz = 0; kz < Z; kz++) {
                    M acc = 0;
                    for (int z = 0; z < Z; z++) acc = std::fma(C[(size_t)kx * Z + z], dTz[(size_t)kz * Z + z], acc);
                    W[bi(kx, y, kz)] = acc;
                }
        }
#pragma omp parallel for
        for (int kz = 0; kz < Z; kz++) {
            std::vector<M> G(Y);
            for (int kx = 0; kx < X; kx++) {
                for (int ky = 0; ky < Y; ky++) {
                    M acc = 0;
                    for (int y = 0; y < Y; y++) acc = std::fma(dTy[(size_t)ky * Y + y], W[bi(kx, y, kz)], acc);
                    const M d = ((dlx[kx] + dly[ky]) + dlz[kz]) * dih2;
                    G[ky] = acc / d;
                }
                for (int y = 0; y < Y; y++) {
                    M acc = 0;
                    for (int ky = 0; ky < Y; ky++) acc = std::fma(dTy[(size_t)ky * Y + y], G[ky], acc);
                    W[bi(kx, y, kz)] = acc;
                }
            }
        }
#pragma omp parallel for
        for (int y = 0; y < Y; y++) {
            std::vector<M> Q((size_t)X * Z);
            for (int kx = 0; kx < X; kx++)
                for (int z = 0; z < Z; z++) {
                    M acc = 0;
                    for (int kz = 0; kz < Z; kz++) acc = std::fma(W[bi(kx, y, kz)], dTz[(size_t)kz * Z + z], acc);
                    Q[(size_t)kx * Z + z] = acc;
                }
            for (int x = 0; x < X; x++)
                for (int z = 0; z < Z; z++) {
                    M acc = 0;
                    for (int kx = 0; kx < X; kx++) acc = std::fma(dTx[(size_t)kx * X + x], Q[(size_t)kx * Z + z], acc);
                    l.p[bi(x, y, z)] = acc;
                }
        }
    }
    void vcycle() {   // sim.js:588-599
        const int n = (int)L.size();
        if (direct > 0) {   // the levels below `direct` replaced by its exact solve
            for (int i = 0; i < direct; i++) { smooth(i, pre); restrict_(i); }
            direct_solve();
            for (int i = direct - 1; i >= 0; i--) { prolong(i); smooth(i, post); }
            return;
        }
        for (int i = 0; i < n - 1; i++) { smooth(i, pre); restrict_(i); }
        smooth(n - 1, coarse_iters);
        for (int i = n - 2; i >= 0; i--) { prolong(i); smooth(i, post); }
    }
    // FP64 residual r = b - A x of level 0 for an FP64 iterate x; returns ||r||^2, ||b||^2
    template <class X>
    void residual64(const X* x, std::vector<double>* r, double* rr, double* bb) const {
        const Level<M>& l = L[0];
        const int nx = l.nx, ny = l.ny, nz = l.nz;
        const double ih2 = 1.0 / (l.h * l.h);
        std::vector<double> prr(nz, 0.0), pbb(nz, 0.0);
#pragma omp parallel for
        for (int z = 0; z < nz; z++) {
            double sr = 0, sb = 0;
            for (int y = 0; y < ny; y++)
                for (int x0 = 0; x0 < nx; x0++) {
                    size_t c = ((size_t)z * ny + y) * nx + x0;
                    double me = (double)x[c], sum = 0, diag = 6;
                    if (x0 > 0) sum += (double)x[c - 1]; else diag += 1;
                    if (x0 < nx - 1) sum += (double)x[c + 1]; else diag += 1;
                    if (z > 0) sum += (double)x[c - (size_t)nx * ny]; else diag += 1;
                    if (z < nz - 1) sum += (double)x[c + (size_t)nx * ny]; else diag += 1;
                    if (y > 0) sum += (double)x[c - nx]; else diag -= 1;
                    if (y < ny - 1) sum += (double)x[c + nx]; else diag += 1;
                    double bv = (double)l.b[c];
                    double rv = bv - (sum - diag * me) * ih2;
                    if (r) (*r)[c] = rv;
                    sr += rv * rv; sb += bv * bv;
                }
            prr[z] = sr; pbb[z] = sb;
        }
        double a = 0, b2 = 0;
        for (int z = 0; z < nz; z++) { a += prr[z]; b2 += pbb[z]; }
        *rr = a; *bb = b2;
    }
};

// FP32 V-cycles inside FP64 defect correction: x <- x + e, e from V-cycles on A e = b - A x (from e = 0), until
// ||b - A x|| / ||b|| <= tol (build's mixed-precision solve).  The first correction takes `first` V-cycles
// (capped at max_cycles), each later one a single V-cycle.  b_keep holds the level-0 right-hand side; x is the
// FP64 iterate.  Returns V-cycles used; hist[k] is the residual after k + 1 V-cycles (set after each correction).
template <class M>
int refine64(Multigrid<M>& mg, std::vector<double>& x, const std::vector<M>& b_keep, double tol,
             int max_cycles, double* rel_out, double* hist, int first = 1) {
    Level<M>& l0 = mg.L[0];
    std::vector<double> r(x.size());
    int used = 0;
    double rr, bb;
    l0.b = b_keep;
    mg.residual64(x.data(), &r, &rr, &bb);
    double rel = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    while (rel > tol && used < max_cycles) {
        const int k = used == 0 ? std::max(1, std::min(first, max_cycles)) : 1;
        for (size_t i = 0; i < x.size(); i++) { l0.b[i] = (M)r[i]; l0.p[i] = 0; }
        for (int c = 0; c < k; c++) mg.vcycle();
        for (size_t i = 0; i < x.size(); i++) x[i] += (double)l0.p[i];
        l0.b = b_keep;
        mg.residual64(x.data(), &r, &rr, &bb);
        rel = bb > 0 ? std::sqrt(rr / bb) : 0.0;
        used += k;
        if (hist) hist[used - 1] = rel;
    }
    l0.b = b_keep;
    if (rel_out) *rel_out = rel;
    return used;
}

// ---- spark particles (src/sparks.js) ------------------------------------------
// The update kernel's uniforms, SPARK_FIELDS order (src/sparks.js:10-14), f32 as uploaded.
struct SparkU {
    float dt, time, frame, count, nBase, nBurst, burstX, burstY, burstZ, burstSpeed, burstRadius;
    float ox, oy, oz, sx, sy, sz, turb, lifeK, bedRadius, seed, sizeK, upK, burstFan, burstSize, aimX, aimZ;
};

// ---- the simulation ---------------------------------------------------------
struct OracleBase {
    virtual ~OracleBase() {}
    virtual void set_params(const float* p) = 0;
    virtual void get_params(float* p) const = 0;
    virtual void set_logs(const float* logs, int n) = 0;
    virtual void reset() = 0;
    virtual void step(double dt) = 0;
    virtual void stage(int st, double dt) = 0;
    virtual int read_field(int f, double* out) const = 0;
    virtual int write_field(int f, const double* in) = 0;
    virtual void field_dims(int f, int d[3], int* ch) const = 0;
    virtual double time() const = 0;
    virtual void set_clock(double t, long f) = 0;
    virtual long frame() const = 0;
    virtual int mg_info(int* dims) const = 0;
    virtual int coarse_iters() const = 0;
    virtual double residual(int* used) const = 0;
    virtual int set_spots(const fs_spot* sp, int n, fs_spot* out) = 0;
    virtual void set_face(const float* f) = 0;
    virtual void set_flux64(const double* spot, int n, const double* face) = 0;
    virtual void spark_update(const struct SparkU& U, float* sp, int count) const = 0;
    virtual void box(double org[3], float size[3]) const = 0;
    virtual void budget_enable(int on) = 0;
    virtual void budget_read(double* out) const = 0;
};

template <class R, class S, class M>
struct Oracle : OracleBase {
    using St = Store<S>;
    int nx, ny, nz;
    size_t ncell;
    double dx, ox, oy, oz;
    float size_[3];
    int vcycles, solve_mode, max_vcycles;
    int first_cycles = 1;      // V-cycles of a step's first FP64 correction (FS_SOLVE_REFINE64)
    double div_tol;
    float prm[FS_NUM_PARAMS];
    std::vector<float> logs;
    int nLogs = 0;
    bool geoReady = false;
    double time_ = 0; long frame_ = 0;
    std::vector<S> vel, velX, velY, Sf, SX, SY, S2, S2X, S2Y, vort, geo;
    NoiseTex noise, curl;
    Multigrid<M> mg;
    std::vector<double> x64;   // FP64 pressure iterate (FS_SOLVE_REFINE64)
    std::vector<fs_spot> spots;          // pyrolysis spots (build extension)
    std::vector<float> face;             // floor fuel flux, nx*nz (empty = none)
    // double-precision overrides of the spot and floor fluxes (finite differences of the FP64
    // oracle with steps below float resolution); empty = the float values above
    std::vector<double> spot64, face64;
    double last_rel = -1; int last_used = 0;
    // FP64 mass and energy budget (the library's fs_budget_*): per-z partial sums of the per-cell
    // terms of the source kernel, combined in z order, and the record
    bool bud_on = false;
    double bud[FS_BUDGET_N] = {};
    std::vector<std::array<double, 18>> bud_z;

    Oracle(const fs_config* c, const float* lg, int nl) {
        nx = c->res[0]; ny = c->res[1]; nz = c->res[2];
        ncell = (size_t)nx * ny * nz;
        dx = (double)c->size[0] / nx;                                 // sim.js:53
        ox = -(double)c->size[0] / 2; oy = 0; oz = -(double)c->size[2] / 2;   // sim.js:54
        for (int k = 0; k < 3; k++) size_[k] = c->size[k];
        vcycles = c->vcycles; solve_mode = c->solve; max_vcycles = c->max_vcycles; div_tol = c->div_tol;
        set_defaults(prm);
        for (auto* f : {&vel, &velX, &velY, &Sf, &SX, &SY, &S2, &S2X, &S2Y, &vort, &geo})
            f->assign(ncell * 4, St::enc(0.0));
        mg.build(nx, ny, nz, (double)(float)dx, c->max_levels);   // L.h is an f32 uniform (sim.js:96)
        if (solve_mode != FS_SOLVE_VCYCLES) {
            mg.omega = (M)(c->omega > 0.f ? c->omega : 1.25f);
            mg.pre = c->pre_sweeps > 0 ? c->pre_sweeps : c->pre_sweeps < 0 ? 0 : 1;
            mg.post = c->post_sweeps > 0 ? c->post_sweeps : 2;
            if (solve_mode == FS_SOLVE_REFINE64) first_cycles = c->first_cycles > 0 ? c->first_cycles : 3;
        }
        if (solve_mode == FS_SOLVE_REFINE64 || solve_mode == FS_SOLVE_TOL32) mg.set_direct(c->direct_level);
        x64.assign(ncell, 0.0);
        set_logs(lg, nl);
        bake_noise(noise, c->noise_n, c->noise_seed);
        bake_curl(curl, noise);
    }
    void set_params(const float* p) override {
        // hotR0/hotR1/hotY1/bedRadius changes invalidate the geometry (sim.js:602)
        if (p[P_hotR0] != prm[P_hotR0] || p[P_hotR1] != prm[P_hotR1] || p[P_hotY1] != prm[P_hotY1] ||
            p[P_bedRadius] != prm[P_bedRadius]) geoReady = false;
        for (int i = 12; i < FS_NUM_PARAMS; i++) prm[i] = p[i];
        prm[P_seed] = p[P_seed];
    }
    void get_params(float* p) const override { for (int i = 0; i < FS_NUM_PARAMS; i++) p[i] = prm[i]; }
    void set_logs(const float* lg, int n) override {
        logs.assign(lg, lg + (size_t)n * 8); nLogs = n; geoReady = false;
    }
    void reset() override {   // sim.js:606-618
        for (auto* f : {&vel, &velX, &velY, &Sf, &SX, &SY, &S2, &S2X, &S2Y, &vort})
            std::fill(f->begin(), f->end(), St::enc(0.0));
        for (auto& l : mg.L) std::fill(l.p.begin(), l.p.end(), M(0));
        std::fill(x64.begin(), x64.end(), 0.0);
        time_ = 0; frame_ = 0;
    }

    // uniform values as the kernels see them (f32 uniform buffer, sim.js:620-624)
    R u(int i) const { return (R)prm[i]; }
    R uDt(double dt) const { return (R)(float)dt; }
    R uDx() const { return (R)(float)dx; }

    size_t idx(int x, int y, int z) const { return ((size_t)z * ny + y) * nx + x; }
    V4<R> ld(const std::vector<S>& f, int x, int y, int z) const {   // clamped textureLoad (sim.js:114)
        x = std::min(std::max(x, 0), nx - 1); y = std::min(std::max(y, 0), ny - 1); z = std::min(std::max(z, 0), nz - 1);
        const S* s = &f[idx(x, y, z) * 4];
        return V4<R>((R)St::dec(s[0]), (R)St::dec(s[1]), (R)St::dec(s[2]), (R)St::dec(s[3]));
    }
    void st(std::vector<S>& f, size_t i, const V4<R>& v) {
        S* s = &f[i * 4];
        s[0] = St::enc((double)v.x); s[1] = St::enc((double)v.y); s[2] = St::enc((double)v.z); s[3] = St::enc((double)v.w);
    }
    // trilinear sample, clamp-to-edge, at cell-space position b (texel centres at i + 0.5)
    V4<R> sample(const std::vector<S>& f, V3<R> b) const {
        R tx = b.x - R(0.5), ty = b.y - R(0.5), tz = b.z - R(0.5);
        tx = wclamp<R>(tx, R(-1), (R)nx); ty = wclamp<R>(ty, R(-1), (R)ny); tz = wclamp<R>(tz, R(-1), (R)nz);
        R fx, fy, fz;
        const int ix = tri_axis<R>(tx, nx, fx), iy = tri_axis<R>(ty, ny, fy), iz = tri_axis<R>(tz, nz, fz);
        V4<R> c[8];
        for (int k = 0; k < 8; k++) c[k] = ld(f, ix + (k & 1), iy + ((k >> 1) & 1), iz + (k >> 2));
        R w[8];
        tri_weights<R>(fx, fy, fz, w);
        return tri_blend<R>(c, w);
    }
    // repeat-addressed sample of a noise texture at normalised coordinate q
    V4<R> sample_rep(const NoiseTex& t, V3<R> q) const {
        const int n = t.n;
        R tx = q.x * (R)n - R(0.5), ty = q.y * (R)n - R(0.5), tz = q.z * (R)n - R(0.5);
        R flx = std::floor(tx), fly = std::floor(ty), flz = std::floor(tz);
        long ix = (long)flx, iy = (long)fly, iz = (long)flz;
        R fx = tx - flx, fy = ty - fly, fz = tz - flz;
        auto L = [&](long x, long y, long z) {
            x = ((x % n) + n) % n; y = ((y % n) + n) % n; z = ((z % n) + n) % n;
            const float* s = &t.v[(((size_t)z * n + y) * n + x) * 4];
            return V4<R>((R)s[0], (R)s[1], (R)s[2], (R)s[3]);
        };
        V4<R> c[8];
        for (int k = 0; k < 8; k++) c[k] = L(ix + (k & 1), iy + ((k >> 1) & 1), iz + (k >> 2));
        R w[8];
        tri_weights<R>(fx, fy, fz, w);
        return tri_blend<R>(c, w);
    }

    // ---- geo (sim.js:498-523) ----
    static R sdCapsule(V3<R> p, V3<R> a, V3<R> b, R r) {   // lib.js:58-62
        V3<R> pa = p - a, ba = b - a;
        R h = wclamp<R>(dot(pa, ba) / dot(ba, ba), R(0), R(1));
        return length(pa - ba * h) - r;
    }
    void k_geo() {
        const R DX = uDx();
#pragma omp parallel for
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    V3<R> p = V3<R>((R)(float)ox, (R)(float)oy, (R)(float)oz) + V3<R>((R)x + R(0.5), (R)y + R(0.5), (R)z + R(0.5)) * DX;
                    R dmin = R(1e9), band = 0;
                    R pxz = std::sqrt(p.x * p.x + p.z * p.z);
                    for (int i = 0; i < nLogs; i++) {
                        const float* A = &logs[(size_t)i * 8]; const float* B = A + 4;
                        R dd = sdCapsule(p, V3<R>(A[0], A[1], A[2]), V3<R>(B[0], B[1], B[2]), (R)A[3]);
                        dd += (R)A[3] * R(0.3) * (R(1) - smoothstep<R>(R(0), R(0.16), pxz)) * (R)B[3];
                        dmin = wmin(dmin, dd);
                        R bnd = R(1) - smoothstep<R>(R(0), R(1.6) * DX, std::fabs(dd - R(0.6) * DX));
                        R hot = (R)B[3] * (R(1) - smoothstep<R>(u(P_hotR0), u(P_hotR1), pxz)) *
                                (R(1) - smoothstep<R>(R(0.08), u(P_hotY1), p.y));
                        band = wmax(band, bnd * hot);
                    }
                    R r = pxz;
                    R bedH = R(0.05) * (R(1) - smoothstep<R>(R(0), u(P_bedRadius), r)) + R(0.02);
                    R bed = (R(1) - smoothstep<R>(R(0), R(2) * DX, p.y - bedH)) *
                            (R(1) - smoothstep<R>(u(P_bedRadius) * R(0.6), u(P_bedRadius) * R(1.15), r));
                    R solid = dmin < R(-0.5) * DX ? R(1) : R(0);
                    st(geo, idx(x, y, z), V4<R>(band, bed, solid, dmin));
                }
    }

    // ---- advection (sim.js:123-197) ----
    V3<R> backtrace(V3<R> c, R dt) const {   // sim.js:130-135
        const R DX = uDx();
        V3<R> v0 = sample(vel, c).xyz();
        V3<R> mid = c - v0 * (R(0.5) * dt) / DX;
        V3<R> v1 = sample(vel, mid).xyz();
        return c - v1 * dt / DX;
    }
    R outsideFade(V3<R> c) const {   // sim.js:137-140
        R lo = wmin(c.x, c.z);
        R hi = wmax(c.x - (R)nx, wmax(c.y - (R)ny, c.z - (R)nz));
        return wclamp<R>(R(1) - wmax(-lo, hi), R(0), R(1));
    }
    void k_adv1(double dtd) {   // sim.js:146-157
        const R dt = uDt(dtd), DX = uDx();
#pragma omp parallel for
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    V3<R> c((R)x + R(0.5), (R)y + R(0.5), (R)z + R(0.5));
                    V3<R> b = backtrace(c, dt);
                    R f = outsideFade(b);
                    R ws = smoothstep<R>(R(0.03), R(0.4), b.y * DX);
                    V3<R> wind = V3<R>(u(P_windX), R(0), u(P_windZ)) * ws;
                    V3<R> sv = sample(vel, b).xyz();
                    size_t i = idx(x, y, z);
                    st(velX, i, V4<R>(mix(wind.x, sv.x, f), mix(wind.y, sv.y, f), mix(wind.z, sv.z, f), R(0)));
                    st(SX, i, sample(Sf, b) * f);
                    st(S2X, i, sample(S2, b) * f);
                }
    }
    void k_adv2(double dtd) {   // sim.js:166-197
        const R dt = uDt(dtd), DX = uDx();
#pragma omp parallel for
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    size_t i = idx(x, y, z);
                    V3<R> c((R)x + R(0.5), (R)y + R(0.5), (R)z + R(0.5));
                    V3<R> v = ld(vel, x, y, z).xyz();
                    V3<R> fw = c + v * dt / DX;
                    V3<R> vH = ld(velX, x, y, z).xyz();
                    V4<R> sH = ld(SX, x, y, z), s2H = ld(S2X, x, y, z);
                    V3<R> vN = vH + (v - sample(velX, fw).xyz()) * R(0.5);
                    V4<R> sN = sH + (ld(Sf, x, y, z) - sample(SX, fw)) * R(0.5);
                    V4<R> s2N = s2H + (ld(S2, x, y, z) - sample(S2X, fw)) * R(0.5);
                    V3<R> bt = backtrace(c, dt);
                    V3<R> b(bt.x - R(0.5), bt.y - R(0.5), bt.z - R(0.5));
                    int b0x = (int)std::floor(wclamp<R>(b.x, R(-2), (R)nx + 1)),
                        b0y = (int)std::floor(wclamp<R>(b.y, R(-2), (R)ny + 1)),
                        b0z = (int)std::floor(wclamp<R>(b.z, R(-2), (R)nz + 1));
                    V3<R> vmn(R(1e9), R(1e9), R(1e9)), vmx(R(-1e9), R(-1e9), R(-1e9));
                    V4<R> smn(R(1e9), R(1e9), R(1e9), R(1e9)), smx(R(-1e9), R(-1e9), R(-1e9), R(-1e9));
                    V4<R> tmn = smn, tmx = smx;
                    for (int k = 0; k < 8; k++) {
                        int qx = b0x + (k & 1), qy = b0y + ((k >> 1) & 1), qz = b0z + ((k >> 2) & 1);
                        V3<R> vv = ld(vel, qx, qy, qz).xyz(); V4<R> ss = ld(Sf, qx, qy, qz), tt = ld(S2, qx, qy, qz);
                        vmn = V3<R>(wmin(vmn.x, vv.x), wmin(vmn.y, vv.y), wmin(vmn.z, vv.z));
                        vmx = V3<R>(wmax(vmx.x, vv.x), wmax(vmx.y, vv.y), wmax(vmx.z, vv.z));
                        smn = V4<R>(wmin(smn.x, ss.x), wmin(smn.y, ss.y), wmin(smn.z, ss.z), wmin(smn.w, ss.w));
                        smx = V4<R>(wmax(smx.x, ss.x), wmax(smx.y, ss.y), wmax(smx.z, ss.z), wmax(smx.w, ss.w));
                        tmn = V4<R>(wmin(tmn.x, tt.x), wmin(tmn.y, tt.y), wmin(tmn.z, tt.z), wmin(tmn.w, tt.w));
                        tmx = V4<R>(wmax(tmx.x, tt.x), wmax(tmx.y, tt.y), wmax(tmx.z, tt.z), wmax(tmx.w, tt.w));
                    }
                    V3<R> fb(b.x + R(0.5), b.y + R(0.5), b.z + R(0.5));
                    if (outsideFade(fb) < R(1)) { vN = vH; sN = sH; s2N = s2H; }
                    auto cl = [](R v, R a, R b2, R h) { return wclamp<R>(v, wmin(a, h), wmax(b2, h)); };
                    vN = V3<R>(cl(vN.x, vmn.x, vmx.x, vH.x), cl(vN.y, vmn.y, vmx.y, vH.y), cl(vN.z, vmn.z, vmx.z, vH.z));
                    sN = V4<R>(cl(sN.x, smn.x, smx.x, sH.x), cl(sN.y, smn.y, smx.y, sH.y), cl(sN.z, smn.z, smx.z, sH.z), cl(sN.w, smn.w, smx.w, sH.w));
                    s2N = V4<R>(cl(s2N.x, tmn.x, tmx.x, s2H.x), cl(s2N.y, tmn.y, tmx.y, s2H.y), cl(s2N.z, tmn.z, tmx.z, s2H.z), cl(s2N.w, tmn.w, tmx.w, s2H.w));
                    st(velY, i, V4<R>(vN.x, vN.y, vN.z, R(0)));
                    st(SY, i, V4<R>(wmax(sN.x, R(0)), wmax(sN.y, R(0)), wmax(sN.z, R(0)), wmax(sN.w, R(0))));
                    st(S2Y, i, V4<R>(wmax(s2N.x, R(0)), wmax(s2N.y, R(0)), wmax(s2N.z, R(0)), wmax(s2N.w, R(0))));
                }
    }
    // ---- curl (sim.js:203-211) ----
    void k_curl() {
#pragma omp parallel for
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    V3<R> L = ld(velY, x - 1, y, z).xyz(), Rr = ld(velY, x + 1, y, z).xyz();
                    V3<R> B = ld(velY, x, y - 1, z).xyz(), T = ld(velY, x, y + 1, z).xyz();
                    V3<R> K = ld(velY, x, y, z - 1).xyz(), F = ld(velY, x, y, z + 1).xyz();
                    V3<R> w = V3<R>((T.z - B.z) - (F.y - K.y), (F.x - K.x) - (Rr.z - L.z), (Rr.y - L.y) - (T.x - B.x)) * R(0.5);
                    st(vort, idx(x, y, z), V4<R>(w.x, w.y, w.z, length(w)));
                }
    }
    // ---- sources, combustion, forces (sim.js:230-396) ----
    V4<R> noise4(V3<R> q) const { return sample_rep(noise, q) * R(2) - V4<R>(1, 1, 1, 1); }   // sim.js:228
    void k_forces(double dtd) {
        const R dt = uDt(dtd), DX = uDx(), tm = (R)(float)time_;
        if (bud_on) bud_z.assign(nz, std::array<double, 18>{});
        const R I = u(P_intensity);
#pragma omp parallel for
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    size_t ci = idx(x, y, z);
                    V3<R> p = V3<R>((R)(float)ox, (R)(float)oy, (R)(float)oz) + V3<R>((R)x + R(0.5), (R)y + R(0.5), (R)z + R(0.5)) * DX;
                    V3<R> v = ld(velY, x, y, z).xyz();
                    V4<R> s = ld(SY, x, y, z), s2 = ld(S2Y, x, y, z);
                    V4<R> g4 = ld(geo, x, y, z);
                    // sources (sim.js:246-297)
                    R src = 0, anchor = 0;
                    if (g4.x > R(0.001)) {
                        V3<R> qa = p * u(P_anchorFreq) + V3<R>(u(P_seed) * R(0.37), R(0), R(0.11));
                        anchor = smoothstep<R>(R(-0.05), R(0.25), noise4(qa).z);
                        V3<R> qn = p * (u(P_srcFreq) * R(0.06)) + V3<R>(u(P_seed) * R(0.131), R(0), tm * u(P_srcSpeed));
                        R pulse = smoothstep<R>(u(P_srcThresh) - R(0.25), u(P_srcThresh) + R(0.25), noise4(qn).x);
                        src = wclamp<R>(g4.x * u(P_logSrc) * (u(P_srcFloor) + u(P_anchorGain) * anchor) * mix(R(0.55), R(1), pulse) * R(1.2), R(0), R(1)) * I;
                        src *= R(1) - g4.z;
                    }
                    R srcW = wmin(src, R(1));
                    R bedHot = g4.y;
                    R bed = g4.y * wclamp<R>(I * R(1.4), R(0.3), R(1));
                    if (bed > R(0.001)) {
                        V3<R> qb = p * R(2.4) + V3<R>(R(0.17), -tm * R(0.12), tm * R(0.05));
                        bed *= smoothstep<R>(R(0), R(0.3), noise4(qb).y);
                    }
                    const R y_in = s.y, th_in = s.x;
                    s.y = wmax(s.y, src * u(P_fuelSrc));
                    s.y = wmax(s.y, bed * u(P_bedFuel));
                    s.x = wmax(s.x, src * u(P_srcTemp));
                    s.x = wmax(s.x, bedHot * u(P_bedTemp));
                    R srcAll = wmax(srcW, bed);   // raised by the spots below
                    s2.w = mix(s2.w, R(1), bed * (R(1) - srcW));
                    s2.w = mix(s2.w, R(0), srcW);
                    s2.y = s2.y * (R(1) - srcAll);
                    s2.x = wmax(s2.x, (R(1) - s.y) * srcW);
                    s2.x = wmax(s2.x, (R(1) - s.y) * bed * R(0.6));
                    R wsrc = R(1) - wpow<R>(R(1) - wclamp<R>(srcW + bed * R(0.4), R(0), R(1)) * R(0.5), dt * R(90));
                    v.y = mix(v.y, wmax(v.y, u(P_srcVel)), wsrc);
                    s2.z += g4.x * (R(1) - anchor) * u(P_smolder) * dt * I;
                    if (u(P_burstAmt) > R(0)) {
                        V3<R> dp = p - V3<R>(u(P_burstX), u(P_burstY), u(P_burstZ));
                        R bd = length(dp);
                        R bw = wclamp<R>(wexp<R>(-bd * bd / R(0.006)) * u(P_burstAmt), R(0), R(1));
                        s.y = wmax(s.y, R(0.17) * bw);
                        s.x = wmax(s.x, R(0.9) * bw);
                        s2.x *= R(1) - bw;
                        s2.w *= R(1) - bw;
                        s2.y = wmax(s2.y, R(0.05) * bw * wclamp<R>(s.y + s2.x, R(0), R(1)));
                    }
                    const R y_src = s.y;
                    // pyrolysis spots and floor fuel flux (build extension, mass-flux sources)
                    for (size_t q = 0; q < spots.size(); q++) {
                        const fs_spot& sp = spots[q];
                        V3<R> dp = p - V3<R>((R)sp.pos[0], (R)sp.pos[1], (R)sp.pos[2]);
                        R d = length(dp);
                        if (d < (R)sp.radius) {
                            R w = R(1) - smoothstep<R>(R(0.5) * (R)sp.radius, (R)sp.radius, d);
                            const R flux = q < spot64.size() ? (R)spot64[q] : (R)sp.flux;
                            s.y = wmin(s.y + flux * dt * (R)sp.norm * w, R(1));
                            s.x = wmax(s.x, w * (R)sp.temp);
                            s2.w = mix(s2.w, R(0), w);
                            s2.y = s2.y * (R(1) - w);
                            srcAll = wmax(srcAll, w);
                        }
                    }
                    if (!face.empty() && y == 0) {
                        const size_t fi = (size_t)z * nx + x;
                        const R fv = face64.empty() ? (R)face[fi] : (R)face64[fi];
                        s.y = wmin(s.y + fv * dt / ((R)FS_RHO0 * DX), R(1));
                    }
                    const R y_flux = s.y, th_flux = s.x;
                    R ageF = s2.y / wmax(s.y + s2.x, R(1e-4));
                    // mixing-limited combustion (sim.js:300-309)
                    R T0 = R(293) + R(1000) * s.x;
                    R o = wclamp<R>(R(1) - s.y - s2.x, R(0), R(1));
                    R ign = smoothstep<R>(u(P_ignT) - R(50), u(P_ignT) + R(50), T0);
                    R b = wmin(s.y, o / u(P_rs)) * (R(1) - wexp<R>(-u(P_kMix) * ign * dt));
                    s.y -= b;
                    s2.x = wmin(s2.x + (R(1) + u(P_rs)) * b, R(1));
                    R dT = wmin(u(P_heatOut) * b, wmax(R(2100) - T0, R(0)));
                    R T = T0 + dT;
                    const R y_burn = s.y, tk_burn = T;
                    R q = dT / dt;
                    R phi = g4.z > R(0.5) ? R(0) : wmin(u(P_expansion) * q / T, R(60));
                    // soot (sim.js:312-318)
                    ageF = wmin(ageF + dt, R(4));
                    s.z += u(P_sootForm) * b * smoothstep<R>(R(0.015), R(0.05), ageF) * smoothstep<R>(R(1000), R(1400), T) * (R(1) - s2.w);
                    s.z *= wexp<R>(-u(P_sootOx) * (R(0.25) + o) * smoothstep<R>(R(1100), R(1400), T) * dt);
                    s.z *= wexp<R>(-u(P_sootOx) * R(0.5) * o * smoothstep<R>(R(850), R(1100), T) * (R(1) - smoothstep<R>(R(1100), R(1300), T)) * dt);
                    s.z *= wexp<R>(-u(P_sootDissip) * dt);
                    // cooling (sim.js:321-330)
                    R th = T * R(0.001);
                    T -= wmin(u(P_radCool) * R(970) * s.z * th * th * (th * th * th * th - R(0.00737)), R(6000)) * dt;
                    const R tk_rad = T;
                    R notBurning = R(1) - smoothstep<R>(R(500), R(5000), q);
                    T = R(293) + wmax(T - R(293), R(0)) * wexp<R>(-u(P_coolRate) * notBurning * dt);
                    s.x = (T - R(293)) * R(0.001);
                    const R tk_cool = T, th_cool = s.x;
                    s.w = q * R(1e-4);
                    s.y *= wexp<R>(-u(P_fuelDissip) * dt);
                    const R y_diss = s.y;
                    s2.x *= wexp<R>(-u(P_vitDecay) * dt);
                    s2.z *= wexp<R>(-u(P_smokeDissip) * dt);
                    s2.y = ageF * wclamp<R>(s.y + s2.x, R(0), R(1));
                    // buoyancy (sim.js:333)
                    v.y += u(P_buoyancy) * s.x * dt;
                    // hot-gas viscosity (sim.js:339-346)
                    if (u(P_visc) > R(0)) {
                        V3<R> lapV = ld(velY, x + 1, y, z).xyz() + ld(velY, x - 1, y, z).xyz() + ld(velY, x, y + 1, z).xyz() +
                                     ld(velY, x, y - 1, z).xyz() + ld(velY, x, y, z + 1).xyz() + ld(velY, x, y, z - 1).xyz() -
                                     ld(velY, x, y, z).xyz() * R(6);
                        R nu = u(P_visc) * R(1.5e-5) * wpow<R>(wmax(T, R(293)) / R(293), R(1.7));
                        v = v + lapV * wmin(nu * dt / (DX * DX), R(0.15));
                    }
                    // laminar gate (sim.js:350)
                    R lam = u(P_lamAge) > R(0) ? smoothstep<R>(R(0.5) * u(P_lamAge), wmax(u(P_lamAge), R(1e-3)), ageF) : R(1);
                    // vorticity confinement (sim.js:353-361)
                    V3<R> w = ld(vort, x, y, z).xyz();
                    V3<R> g(ld(vort, x + 1, y, z).w - ld(vort, x - 1, y, z).w,
                            ld(vort, x, y + 1, z).w - ld(vort, x, y - 1, z).w,
                            ld(vort, x, y, z + 1).w - ld(vort, x, y, z - 1).w);
                    R gl = length(g);
                    if (gl > R(1e-6)) {
                        R gate = smoothstep<R>(R(450), R(900), T) * (R(1) - srcAll) * smoothstep<R>(R(0.02), R(0.08), p.y) * lam;
                        v = v + cross(unit_dir<R>(g, gl), w) * (u(P_vortEps) * gate * dt);
                    }
                    // curl noise (sim.js:364-368)
                    V3<R> tq = (p - V3<R>(R(0), tm * u(P_turbRise), R(0))) * u(P_turbFreq) + V3<R>(R(0), R(0), tm * u(P_turbEvolve));
                    V3<R> cn = sample_rep(curl, tq).xyz();
                    R tgate = wmax(smoothstep<R>(R(500), R(1100), T), smoothstep<R>(R(0), R(3), s.w)) * (R(1) - srcAll) *
                              mix(R(0.6), R(1.8), smoothstep<R>(R(0.2), R(0.75), p.y)) * lam;
                    v = v + cn * ((u(P_turbAmp) / R(12)) * tgate * dt * R(10));
                    // wind (sim.js:372-375)
                    R windK = smoothstep<R>(R(0.05), R(0.5), p.y) * (R(1) + u(P_gust));
                    v.x += (u(P_windX) * windK - v.x) * dt * R(0.35) * windK;
                    v.z += (u(P_windZ) * windK - v.z) * dt * R(0.35) * windK;
                    v = v * wexp<R>(-u(P_velDissip) * dt);
                    // solids and sponges (sim.js:378-388)
                    if (g4.z > R(0.5)) { v = V3<R>(); s.y = 0; s2.y = 0; }
                    V3<R> fp((R)x + R(0.5), (R)y + R(0.5), (R)z + R(0.5));
                    R side = wmin(wmin(fp.x, (R)nx - fp.x), wmin(fp.z, (R)nz - fp.z));
                    R sp = (R(1) - smoothstep<R>(R(0), R(5), side)) * u(P_sponge);
                    R wsg = smoothstep<R>(R(0.03), R(0.4), p.y);
                    R wSx = u(P_windX) * wsg, wSz = u(P_windZ) * wsg;
                    R ks = R(1) - wpow<R>(R(1) - R(0.15) * wclamp<R>(sp, R(0), R(1)), dt * R(90));
                    v.x = mix(v.x, wSx, ks); v.z = mix(v.z, wSz, ks);
                    R top = smoothstep<R>((R)ny - R(10), (R)ny, fp.y);
                    R kt = wpow<R>(R(1) - R(0.1) * top, dt * R(90));
                    s.z *= kt; s2.z *= kt;
                    // clamps and NaN reset (sim.js:390-392)
                    R vl = length(v);
                    if (vl > R(15)) v = v * (R(15) / vl);
                    bool bad = std::isnan(s.x) || std::isnan(s.y) || std::isnan(s.z) || std::isnan(s.w) ||
                               std::isnan(s2.x) || std::isnan(s2.y) || std::isnan(s2.z) || std::isnan(s2.w) ||
                               std::isnan(v.x) || std::isnan(v.y) || std::isnan(v.z);
                    if (bad) { s = V4<R>(); s2 = V4<R>(); v = V3<R>(); }
                    st(velX, ci, V4<R>(v.x, v.y, v.z, phi));
                    st(Sf, ci, s);
                    st(S2, ci, s2);
                    if (bud_on) {   // the budget terms: differences of consecutive values of this cell
                        const double yst = St::dec(St::enc((double)s.y)), thst = St::dec(St::enc((double)s.x));
                        const double d[17] = {
                            (double)y_in, (double)y_src - y_in, (double)y_flux - y_src, (double)y_flux - y_burn,
                            (double)y_burn - y_diss, (double)y_diss - s.y, yst - s.y, yst,
                            1000.0 * th_in, 1000.0 * ((double)th_flux - th_in), (double)tk_burn - T0,
                            (double)tk_burn - tk_rad, (double)tk_rad - tk_cool, 1000.0 * ((double)th_cool - s.x),
                            ((double)T0 - 293.0 - 1000.0 * (double)th_flux) + (1000.0 * (double)th_cool - ((double)tk_cool - 293.0)) +
                                1000.0 * (thst - s.x),
                            1000.0 * thst, St::dec(St::enc((double)phi))};
                        bool fin = true;
                        for (double e : d) fin = fin && std::isfinite(e);
                        auto& a = bud_z[z];
                        if (fin) for (int q2 = 0; q2 < 17; q2++) a[q2] += d[q2];
                        else a[17] += 1.0;
                    }
                }
    }
    // sums of the per-z partials in z order, into the record (the library's k_budget_step)
    void budget_step(double dtd) {
        double t[18] = {};
        for (int z = 0; z < nz; z++) for (int k = 0; k < 18; k++) t[k] += bud_z[z][k];
        double bsum = 0.0;
        for (size_t i = 0; i < ncell; i++) bsum += (double)mg.L[0].b[i];
        const double dxf = (double)(float)dx, V = dxf * dxf * dxf, kg = FS_RHO0 * V, J = FS_RHO0 * 1005.0 * V;
        double* r = bud;
        r[FS_BUDGET_STEPS] += 1.0;
        r[FS_BUDGET_FUEL_TRANSPORT] += kg * t[0] - r[FS_BUDGET_FUEL];
        r[FS_BUDGET_FUEL_SOURCES] += kg * t[1];
        r[FS_BUDGET_FUEL_INJECTED] += kg * t[2];
        r[FS_BUDGET_FUEL_BURNED] += kg * t[3];
        r[FS_BUDGET_FUEL_DISSIPATED] += kg * t[4];
        r[FS_BUDGET_FUEL_REMOVED] += kg * t[5];
        r[FS_BUDGET_FUEL_ROUNDING] += kg * t[6];
        r[FS_BUDGET_FUEL] = kg * t[7];
        r[FS_BUDGET_HEAT_TRANSPORT] += J * t[8] - r[FS_BUDGET_HEAT];
        r[FS_BUDGET_HEAT_SOURCES] += J * t[9];
        r[FS_BUDGET_HEAT_RELEASED] += J * t[10];
        r[FS_BUDGET_HEAT_RADIATED] += J * t[11];
        r[FS_BUDGET_HEAT_ENTRAINED] += J * t[12];
        r[FS_BUDGET_HEAT_REMOVED] += J * t[13];
        r[FS_BUDGET_HEAT_ROUNDING] += J * t[14];
        r[FS_BUDGET_HEAT] = J * t[15];
        const double dt = (double)(float)dtd;
        double inj = 0.0;
        for (const fs_spot& sp : spots) inj += (double)sp.flux * dt;
        double fx = 0.0;
        for (float f : face) fx += f;
        r[FS_BUDGET_FUEL_EXPECTED] += inj + fx * dxf * dxf * dt;
        r[FS_BUDGET_EXPANSION] = V * t[16];
        r[FS_BUDGET_DIVERGENCE] = V * bsum;
        r[FS_BUDGET_BAD_CELLS] += t[17];
    }
    void budget_enable(int on) override {
        bud_on = on != 0;
        if (!bud_on) return;
        for (double& e : bud) e = 0.0;
        double y = 0.0, th = 0.0;
        for (size_t i = 0; i < ncell; i++) { y += St::dec(Sf[i * 4 + 1]); th += 1000.0 * St::dec(Sf[i * 4]); }
        const double dxf = (double)(float)dx, V = dxf * dxf * dxf;
        bud[FS_BUDGET_FUEL0] = bud[FS_BUDGET_FUEL] = FS_RHO0 * V * y;
        bud[FS_BUDGET_HEAT0] = bud[FS_BUDGET_HEAT] = FS_RHO0 * 1005.0 * V * th;
    }
    void budget_read(double* out) const override { for (int i = 0; i < FS_BUDGET_N; i++) out[i] = bud[i]; }
    // ---- divergence minus expansion (sim.js:403-413) ----
    void k_div() {
        const R DX = uDx();
        std::vector<M>& b = mg.L[0].b;
#pragma omp parallel for
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    V4<R> me = ld(velX, x, y, z);
                    R Lx = ld(velX, x - 1, y, z).x, Rx = ld(velX, x + 1, y, z).x;
                    R B = ld(velX, x, y - 1, z).y, T = ld(velX, x, y + 1, z).y;
                    R K = ld(velX, x, y, z - 1).z, F = ld(velX, x, y, z + 1).z;
                    if (y == 0) B = -me.y;
                    R dv = ((Rx - Lx) + (T - B) + (F - K)) / (R(2) * DX) - me.w;
                    b[idx(x, y, z)] = (M)dv;
                }
    }
    // ---- pressure solve (sim.js:642) ----
    void k_mg() {
        if (solve_mode == FS_SOLVE_REFINE64) {
            std::vector<M> bk = mg.L[0].b;
            last_used = refine64(mg, x64, bk, div_tol, max_vcycles, &last_rel, nullptr, first_cycles);
            for (size_t i = 0; i < ncell; i++) mg.L[0].p[i] = (M)x64[i];
        } else if (solve_mode == FS_SOLVE_TOL32) {
            double rr, bb;
            mg.residual64(mg.L[0].p.data(), nullptr, &rr, &bb);
            double rel = bb > 0 ? std::sqrt(rr / bb) : 0.0;
            int used = 0;
            while (rel > div_tol && used < max_vcycles) {
                mg.vcycle();
                mg.residual64(mg.L[0].p.data(), nullptr, &rr, &bb);
                rel = bb > 0 ? std::sqrt(rr / bb) : 0.0;
                used++;
            }
            last_used = used; last_rel = rel;
        } else {
            for (int i = 0; i < vcycles; i++) mg.vcycle();
            last_used = vcycles;
            double rr, bb;
            mg.residual64(mg.L[0].p.data(), nullptr, &rr, &bb);
            last_rel = bb > 0 ? std::sqrt(rr / bb) : 0.0;
        }
    }
    // ---- projection (sim.js:532-550) ----
    void k_project() {
        const R DX = uDx();
        const std::vector<M>& pp = mg.L[0].p;
        auto pr = [&](int x, int y, int z, R me) -> R {
            if (x < 0 || z < 0 || x >= nx || z >= nz || y >= ny) return -me;
            if (y < 0) return me;
            return (R)pp[idx(x, y, z)];
        };
#pragma omp parallel for
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    R me = (R)pp[idx(x, y, z)];
                    V3<R> g = V3<R>(pr(x + 1, y, z, me) - pr(x - 1, y, z, me), pr(x, y + 1, z, me) - pr(x, y - 1, z, me),
                                    pr(x, y, z + 1, me) - pr(x, y, z - 1, me)) / (R(2) * DX);
                    V3<R> v = ld(velX, x, y, z).xyz() - g;
                    if (ld(geo, x, y, z).z > R(0.5)) v = V3<R>();
                    if (y == 0) v.y = wmax(v.y, R(0));
                    R vl = length(v);
                    if (vl > R(20) || vl != vl) v = (vl != vl) ? V3<R>() : v * R(20) / vl;
                    st(vel, idx(x, y, z), V4<R>(v.x, v.y, v.z, R(0)));
                }
    }

    void stage(int s, double dt) override {
        switch (s) {
            case FS_STAGE_GEO: k_geo(); geoReady = true; break;
            case FS_STAGE_ADV1: k_adv1(dt); break;
            case FS_STAGE_ADV2: k_adv2(dt); break;
            case FS_STAGE_CURL: k_curl(); break;
            case FS_STAGE_FORCES: k_forces(dt); break;
            case FS_STAGE_DIV: k_div(); break;
            case FS_STAGE_MG: k_mg(); break;
            case FS_STAGE_PROJECT: k_project(); break;
        }
    }
    void step(double dt) override {   // sim.js:627-648
        if (!geoReady) { k_geo(); geoReady = true; }
        k_adv1(dt); k_adv2(dt); k_curl(); k_forces(dt); k_div();
        if (bud_on) budget_step(dt);
        k_mg(); k_project();
        time_ += dt; frame_++;
    }

    std::vector<S>* field4(int f) {
        switch (f) {
            case FS_VEL: return &vel; case FS_S: return &Sf; case FS_S2: return &S2; case FS_VORT: return &vort;
            case FS_GEO: return &geo; case FS_VELX: return &velX; case FS_SX: return &SX; case FS_S2X: return &S2X;
            case FS_VELY: return &velY; case FS_SY: return &SY; case FS_S2Y: return &S2Y;
        }
        return nullptr;
    }
    void field_dims(int f, int d[3], int* ch) const override {
        if (f == FS_NOISE || f == FS_CURL) { d[0] = d[1] = d[2] = noise.n; *ch = 4; return; }
        d[0] = nx; d[1] = ny; d[2] = nz; *ch = (f == FS_P || f == FS_B) ? 1 : 4;
    }
    int read_field(int f, double* out) const override {
        auto* self = const_cast<Oracle*>(this);
        if (f == FS_NOISE || f == FS_CURL) {
            const NoiseTex& t = f == FS_NOISE ? noise : curl;
            for (size_t i = 0; i < t.v.size(); i++) out[i] = t.v[i];
            return 0;
        }
        if (f == FS_P) {
            if (solve_mode == FS_SOLVE_REFINE64) { for (size_t i = 0; i < ncell; i++) out[i] = x64[i]; }
            else for (size_t i = 0; i < ncell; i++) out[i] = (double)mg.L[0].p[i];
            return 0;
        }
        if (f == FS_B) { for (size_t i = 0; i < ncell; i++) out[i] = (double)mg.L[0].b[i]; return 0; }
        std::vector<S>* v = self->field4(f);
        if (!v) return FS_EINVAL;
        for (size_t i = 0; i < v->size(); i++) out[i] = St::dec((*v)[i]);
        return 0;
    }
    int write_field(int f, const double* in) override {
        if (f == FS_NOISE || f == FS_CURL) {
            NoiseTex& t = f == FS_NOISE ? noise : curl;
            for (size_t i = 0; i < t.v.size(); i++) t.v[i] = f16_to_f32(f32_to_f16((float)in[i]));
            return 0;
        }
        if (f == FS_P) {
            for (size_t i = 0; i < ncell; i++) { mg.L[0].p[i] = (M)in[i]; x64[i] = in[i]; }
            return 0;
        }
        if (f == FS_B) { for (size_t i = 0; i < ncell; i++) mg.L[0].b[i] = (M)in[i]; return 0; }
        std::vector<S>* v = field4(f);
        if (!v) return FS_EINVAL;
        for (size_t i = 0; i < v->size(); i++) (*v)[i] = St::enc(in[i]);
        if (f == FS_GEO) geoReady = true;
        return 0;
    }
    double time() const override { return time_; }
    void set_clock(double t, long f) override { time_ = t; frame_ = f; }
    long frame() const override { return frame_; }
    int mg_info(int* dims) const override {
        for (size_t i = 0; i < mg.L.size(); i++) { dims[i * 3] = mg.L[i].nx; dims[i * 3 + 1] = mg.L[i].ny; dims[i * 3 + 2] = mg.L[i].nz; }
        return (int)mg.L.size();
    }
    int coarse_iters() const override { return mg.coarse_iters; }
    // normalisation of a spot's mass injection (double; cell centres origin + (i + 1/2) dx)
    double spot_weight_sum(const fs_spot& sp) const {
        const double d_x = (double)(float)dx;
        const double o_x = (double)(float)ox, o_y = (double)(float)oy, o_z = (double)(float)oz;
        const double r = sp.radius;
        auto cell = [&](double o, double pc) { return (int)std::floor((pc - o) / d_x - 0.5); };
        const int x0 = std::max(0, cell(o_x, sp.pos[0] - r) - 1), x1 = std::min(nx - 1, cell(o_x, sp.pos[0] + r) + 2);
        const int y0 = std::max(0, cell(o_y, sp.pos[1] - r) - 1), y1 = std::min(ny - 1, cell(o_y, sp.pos[1] + r) + 2);
        const int z0 = std::max(0, cell(o_z, sp.pos[2] - r) - 1), z1 = std::min(nz - 1, cell(o_z, sp.pos[2] + r) + 2);
        double sum = 0.0;
        for (int z = z0; z <= z1; z++)
            for (int y = y0; y <= y1; y++)
                for (int x = x0; x <= x1; x++) {
                    const double px = o_x + (x + 0.5) * d_x, py = o_y + (y + 0.5) * d_x, pz = o_z + (z + 0.5) * d_x;
                    const double d = std::sqrt((px - sp.pos[0]) * (px - sp.pos[0]) + (py - sp.pos[1]) * (py - sp.pos[1]) +
                                               (pz - sp.pos[2]) * (pz - sp.pos[2]));
                    if (d >= r) continue;
                    double t = (d - 0.5 * r) / (0.5 * r);
                    t = std::min(std::max(t, 0.0), 1.0);
                    sum += 1.0 - t * t * (3.0 - 2.0 * t);
                }
        return sum;
    }
    int set_spots(const fs_spot* sp, int n, fs_spot* out) override {
        spots.assign(sp, sp + n);
        const double dx3 = std::pow((double)(float)dx, 3.0);
        for (auto& q : spots) {
            const double ws = spot_weight_sum(q);
            q.norm = ws > 0.0 ? (float)(1.0 / ((double)FS_RHO0 * ws * dx3)) : 0.f;
        }
        if (out) for (int i = 0; i < n; i++) out[i] = spots[i];
        return 0;
    }
    void set_face(const float* f) override {
        if (f) face.assign(f, f + (size_t)nx * nz); else face.clear();
    }
    void box(double org[3], float size[3]) const override {
        org[0] = ox; org[1] = oy; org[2] = oz;
        for (int k = 0; k < 3; k++) size[k] = size_[k];
    }
    // One launch of the spark update (src/sparks.js:80-135) over the particle array, in f32.
    // Dead particles take the spawn slots in index order.
    void spark_update(const SparkU& U, float* sp, int count) const override {
        const uint32_t want = (uint32_t)(U.nBase + U.nBurst);
        uint32_t next = 0;
        const float dt = U.dt;
        for (int i = 0; i < count; i++) {
            float* P = sp + 16 * i; float* Vv = P + 4; float* Q = P + 8; float* Mm = P + 12;
            if (Mm[2] < 0.5f) {
                const uint32_t slot = next++;
                if (want == 0u || slot >= want) continue;
                const uint32_t seed = (uint32_t)U.seed, frame = (uint32_t)U.frame;
                auto r = [&](uint32_t k) { return rand2(slot * 7919u + k * 104729u + seed, frame); };
                auto gauss = [&](uint32_t k) {
                    const float u1 = std::fmax(r(k), 1e-6f);
                    const float u2 = r(k + 1u);
                    return std::sqrt(-2.0f * flog(u1)) * fcos(6.2831853f * u2);
                };
                const bool burst = slot >= (uint32_t)U.nBase;
                V3<float> p, v;
                if (burst) {
                    p = V3<float>(U.burstX, U.burstY, U.burstZ) +
                        (V3<float>(r(7u), r(8u), r(9u)) - V3<float>(0.5f, 0.5f, 0.5f)) * U.burstRadius * 2.0f;
                    V3<float> radial(p.x, 0.0f, p.z);
                    radial = length(radial) > 1e-3f ? radial / length(radial) : V3<float>();
                    const V3<float> aim(U.aimX, 0.0f, U.aimZ);
                    V3<float> d = V3<float>(gauss(1u), std::fabs(gauss(3u)) * 1.2f + 0.7f, gauss(5u)) + (radial + aim) * U.burstFan;
                    d = d / length(d);
                    v = d * U.burstSpeed * (0.5f + 0.8f * r(10u));
                } else {
                    const float a = r(1u) * 6.2831853f;
                    const float rr = std::sqrt(r(2u)) * U.bedRadius;
                    p = V3<float>(fcos(a) * rr, 0.05f + r(3u) * 0.12f, fsin(a) * rr);
                    v = V3<float>(gauss(4u) * 0.4f, 0.6f + r(6u) * 1.8f, gauss(11u) * 0.4f);
                }
                const float dmm = wclamp<float>(0.5f * fexp(0.55f * gauss(13u)) * U.sizeK * (burst ? U.burstSize : 1.0f), 0.15f, 2.5f);
                const float tb = (0.25f + 2.0f * dmm * dmm) * (0.6f + 0.8f * r(15u)) * U.lifeK;
                const float T0 = 1230.0f + 70.0f * gauss(16u) + (burst ? 80.0f : 0.0f);
                const float fl = r(20u) < 0.45f ? 5.0f + 15.0f * r(18u) : 0.0f;
                P[0] = p.x; P[1] = p.y; P[2] = p.z; P[3] = 0.0f;
                Vv[0] = v.x; Vv[1] = v.y; Vv[2] = v.z; Vv[3] = tb;
                Q[0] = p.x; Q[1] = p.y; Q[2] = p.z; Q[3] = T0;
                Mm[0] = dmm; Mm[1] = r(19u); Mm[2] = 1.0f; Mm[3] = fl;
                continue;
            }
            const float dmm = Mm[0];
            V3<float> p(P[0], P[1], P[2]), v(Vv[0], Vv[1], Vv[2]);
            const float tau = std::fmax(0.4f * dmm * dmm, 0.01f);
            V3<float> u;
            const float gx = (p.x - U.ox) / U.sx, gy = (p.y - U.oy) / U.sy, gz = (p.z - U.oz) / U.sz;
            if (gx < 0.f || gy < 0.f || gz < 0.f || gx > 1.f || gy > 1.f || gz > 1.f) {
                if (gy > 1.0f && std::sqrt(p.x * p.x + p.z * p.z) < 0.5f) u = V3<float>(0.0f, 1.6f / (1.0f + (gy - 1.0f) * 1.5f), 0.0f);
            } else {
                const V4<R> s4 = sample(vel, V3<R>((R)(gx * (float)nx), (R)(gy * (float)ny), (R)(gz * (float)nz)));
                u = V3<float>((float)s4.x, (float)s4.y, (float)s4.z);
            }
            u = u * U.upK;
            const float tqx = p.x * 0.9f + 0.0f, tqy = p.y * 0.9f + -U.time * 0.5f, tqz = p.z * 0.9f + (U.time * 0.13f + Mm[1] * 3.0f);
            const float turbAmp = U.turb * (0.3f + 3.0f * smoothstep<float>(0.4f, 1.2f, p.y)) / 12.0f;
            const V4<R> cn = sample_rep(curl, V3<R>((R)tqx, (R)tqy, (R)tqz));
            u = u + V3<float>((float)cn.x, (float)cn.y, (float)cn.z) * turbAmp;
            const float k = 1.0f - fexp(-dt / tau);
            v = V3<float>(mix(v.x, u.x, k), mix(v.y, u.y, k), mix(v.z, u.z, k));
            v.y = v.y - 9.77076f * dt * (dmm < 0.3f ? 0.2f : 1.0f);
            const V3<float> prev = p;
            p = p + v * dt;
            float T = Q[3];
            const float tb = Vv[3] - dt;
            if (tb > 0.0f) {
                const float rel = length(u - v);
                T = mix(T, 1180.0f + 60.0f * std::fmin(rel, 3.0f) + 60.0f * (Mm[1] - 0.5f), 1.0f - fexp(-dt / 0.3f));
            } else {
                T = 293.0f + (T - 293.0f) * fexp(-dt / (0.06f + 0.08f * dmm));
            }
            const float age = P[3] + dt;
            const bool kill = T < 760.0f || p.y < 0.0f || p.y > 4.5f || std::sqrt(p.x * p.x + p.z * p.z) > 3.0f || age > 12.0f;
            P[0] = p.x; P[1] = p.y; P[2] = p.z; P[3] = age;
            Vv[0] = v.x; Vv[1] = v.y; Vv[2] = v.z; Vv[3] = tb;
            Q[0] = prev.x; Q[1] = prev.y; Q[2] = prev.z; Q[3] = T;
            Mm[2] = kill ? 0.0f : 1.0f;
        }
    }
    void set_flux64(const double* spot, int n, const double* f) override {
        if (spot) spot64.assign(spot, spot + n); else spot64.clear();
        if (f) face64.assign(f, f + (size_t)nx * nz); else face64.clear();
    }
    double residual(int* used) const override { if (used) *used = last_used; return last_rel; }
};

// ---- the poisson.mjs variant (research/webgpu-bench/poisson.mjs:8-124) -------
// Kept apart from the sim.js multigrid: separate residual pass, restriction of r,
// prolongation with clamped (not mirrored) coarse nodes, 30 coarse sweeps, N^3
// cubes, omega-relaxed smoothers.  Only used to pin the oracle against the
// reference's committed convergence history (results/poisson-128-face.txt).
struct MjsLevel { int n; std::vector<float> p, b, r; };
struct MjsPoisson {
    bool face = true;
    double nbSumD(const MjsLevel& L, int x, int y, int z, double pc) const {
        const int n = L.n;
        auto g = [&](int xx, int yy, int zz) -> double {
            if (yy < 0) return pc;
            if (xx < 0 || yy >= n || zz < 0 || xx >= n || zz >= n) return face ? -pc : 0.0;
            return (double)L.p[(size_t)xx + (size_t)n * (yy + (size_t)n * zz)];
        };
        return g(x - 1, y, z) + g(x + 1, y, z) + g(x, y - 1, z) + g(x, y + 1, z) + g(x, y, z - 1) + g(x, y, z + 1);
    }
    void jacobi(MjsLevel& L, double h2, double omega, std::vector<float>& tmp) {   // poisson.mjs:20-27
        const int n = L.n;
#pragma omp parallel for
        for (int z = 0; z < n; z++) for (int y = 0; y < n; y++) for (int x = 0; x < n; x++) {
            size_t i = (size_t)x + (size_t)n * (y + (size_t)n * z); double pc = L.p[i];
            tmp[i] = (float)(pc + omega * ((nbSumD(L, x, y, z, pc) - h2 * (double)L.b[i]) / 6.0 - pc));
        }
        L.p = tmp;
    }
    void rbgs(MjsLevel& L, double h2, double omega) {   // poisson.mjs:28-35
        const int n = L.n;
        for (int color = 0; color < 2; color++) {
#pragma omp parallel for
            for (int z = 0; z < n; z++) for (int y = 0; y < n; y++) for (int x = (y + z + color) & 1; x < n; x += 2) {
                size_t i = (size_t)x + (size_t)n * (y + (size_t)n * z); double pc = L.p[i];
                L.p[i] = (float)(pc + omega * ((nbSumD(L, x, y, z, pc) - h2 * (double)L.b[i]) / 6.0 - pc));
            }
        }
    }
    double residual(MjsLevel& L, double h2) {   // poisson.mjs:36-43
        const int n = L.n;
        std::vector<double> part(n, 0.0);
#pragma omp parallel for
        for (int z = 0; z < n; z++) {
            double s = 0;
            for (int y = 0; y < n; y++) for (int x = 0; x < n; x++) {
                size_t i = (size_t)x + (size_t)n * (y + (size_t)n * z); double pc = L.p[i];
                L.r[i] = (float)((double)L.b[i] - (nbSumD(L, x, y, z, pc) - 6.0 * pc) / h2);
                s += (double)L.r[i] * (double)L.r[i];
            }
            part[z] = s;
        }
        double s = 0; for (int z = 0; z < n; z++) s += part[z];
        return std::sqrt(s / ((double)n * n * n));
    }
    void restrict_(const MjsLevel& F, MjsLevel& C) {   // poisson.mjs:44-51
        const int n = C.n;
        for (int z = 0; z < n; z++) for (int y = 0; y < n; y++) for (int x = 0; x < n; x++) {
            double s = 0;
            for (int k = 0; k < 8; k++)
                s += F.r[(size_t)(2 * x + (k & 1)) + (size_t)F.n * ((2 * y + ((k >> 1) & 1)) + (size_t)F.n * (2 * z + ((k >> 2) & 1)))];
            size_t i = (size_t)x + (size_t)n * (y + (size_t)n * z);
            C.b[i] = (float)(s / 8); C.p[i] = 0;
        }
    }
    void prolongAdd(const MjsLevel& C, MjsLevel& F) {   // poisson.mjs:52-65
        const int n = C.n;
        auto cl = [&](int v) { return std::max(0, std::min(n - 1, v)); };
#pragma omp parallel for
        for (int z = 0; z < F.n; z++) for (int y = 0; y < F.n; y++) for (int x = 0; x < F.n; x++) {
            int cx = x >> 1, cy = y >> 1, cz = z >> 1;
            int ox = (x & 1) ? 1 : -1, oy = (y & 1) ? 1 : -1, oz = (z & 1) ? 1 : -1;
            double v = 0;
            for (int k = 0; k < 8; k++) {
                int dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
                double w = (dx ? 0.25 : 0.75) * (dy ? 0.25 : 0.75) * (dz ? 0.25 : 0.75);
                v += w * (double)C.p[(size_t)cl(cx + dx * ox) + (size_t)n * (cl(cy + dy * oy) + (size_t)n * cl(cz + dz * oz))];
            }
            size_t i = (size_t)x + (size_t)F.n * (y + (size_t)F.n * z);
            F.p[i] = (float)((double)F.p[i] + v);
        }
    }
    void vcycle(std::vector<MjsLevel>& levels, int l, int pre, int post) {   // poisson.mjs:66-75
        MjsLevel& L = levels[l]; double h2 = std::pow(4.0, l);
        if (l == (int)levels.size() - 1) { for (int i = 0; i < 30; i++) rbgs(L, h2, 1.0); return; }
        for (int i = 0; i < pre; i++) rbgs(L, h2, 1.0);
        residual(L, h2);
        restrict_(L, levels[l + 1]);
        vcycle(levels, l + 1, pre, post);
        prolongAdd(levels[l + 1], L);
        for (int i = 0; i < post; i++) rbgs(L, h2, 1.0);
    }
    static std::vector<float> rhs(int N) {   // poisson.mjs:78-92
        std::vector<float> b((size_t)N * N * N);
        auto vel = [&](int x, int y, int z, double o[3]) {
            double uu = (x + 0.5) / N, v = (y + 0.5) / N, w = (z + 0.5) / N;
            double r = std::hypot(uu - 0.5, w - 0.5);
            double plume = std::exp(-(r * r) / 0.01) * std::min(1.0, v * 4);
            o[0] = std::sin(v * 12) * 0.3 + std::sin(w * 31 + uu * 7) * 0.2;
            o[1] = 1.5 * plume + std::cos(uu * 9) * 0.2;
            o[2] = std::sin(w * 7 + uu * 5) * 0.3;
        };
        auto c = [&](int a, int b2, int cc, double o[3]) {
            vel(std::max(0, std::min(N - 1, a)), std::max(0, std::min(N - 1, b2)), std::max(0, std::min(N - 1, cc)), o);
        };
#pragma omp parallel for
        for (int z = 0; z < N; z++) for (int y = 0; y < N; y++) for (int x = 0; x < N; x++) {
            double t[3], a[3], bb[3];
            double vD;
            if (y == 0) { vel(x, y, z, t); vD = -t[1]; } else { c(x, y - 1, z, t); vD = t[1]; }
            c(x + 1, y, z, a); c(x - 1, y, z, bb); double d1 = a[0] - bb[0];
            c(x, y + 1, z, a); double d2 = a[1] - vD;
            c(x, y, z + 1, a); c(x, y, z - 1, bb); double d3 = a[2] - bb[2];
            b[(size_t)x + (size_t)N * (y + (size_t)N * z)] = (float)(0.5 * (d1 + d2 + d3));
        }
        return b;
    }
};

thread_local std::string g_err;

}  // namespace

// ============================== C ABI ========================================
extern "C" {

struct orc_sim { OracleBase* o; };

// precision: FS_PREC_FP32 -> <float, float, float>; FS_PREC_FP16 -> <float, half, float>;
// FS_PREC_BF16 -> <float, bf16, float>; 3 -> FP64 error oracle <double, double, double>.
orc_sim* orc_create(const fs_config* c, const float* logs, int n_logs) {
    OracleBase* o = nullptr;
    switch (c->precision) {
        case FS_PREC_FP32: o = new Oracle<float, float, float>(c, logs, n_logs); break;
        case FS_PREC_FP16: o = new Oracle<float, half_t, float>(c, logs, n_logs); break;
        case FS_PREC_BF16: o = new Oracle<float, bf16_t, float>(c, logs, n_logs); break;
        case 3: o = new Oracle<double, double, double>(c, logs, n_logs); break;
        default: g_err = "bad precision"; return nullptr;
    }
    return new orc_sim{o};
}
void orc_destroy(orc_sim* s) { if (s) { delete s->o; delete s; } }
void orc_set_params(orc_sim* s, const float* p) { s->o->set_params(p); }
void orc_get_params(orc_sim* s, float* p) { s->o->get_params(p); }
void orc_param_defaults(float* p) { set_defaults(p); }
const char* orc_param_name(int i) { return (i >= 0 && i < FS_NUM_PARAMS) ? kNames[i] : nullptr; }
void orc_set_logs(orc_sim* s, const float* logs, int n) { s->o->set_logs(logs, n); }
void orc_reset(orc_sim* s) { s->o->reset(); }
void orc_step(orc_sim* s, double dt, int n) { for (int i = 0; i < n; i++) s->o->step(dt); }
void orc_stage(orc_sim* s, int stage, double dt) { s->o->stage(stage, dt); }
int orc_read_field(orc_sim* s, int f, double* out) { return s->o->read_field(f, out); }
int orc_write_field(orc_sim* s, int f, const double* in) { return s->o->write_field(f, in); }
void orc_field_dims(orc_sim* s, int f, int* d, int* ch) { s->o->field_dims(f, d, ch); }
double orc_time(orc_sim* s) { return s->o->time(); }
void orc_set_clock(orc_sim* s, double t, long f) { s->o->set_clock(t, f); }
long orc_frame(orc_sim* s) { return s->o->frame(); }
int orc_mg_info(orc_sim* s, int* dims) { return s->o->mg_info(dims); }
int orc_coarse_iters(orc_sim* s) { return s->o->coarse_iters(); }
double orc_residual(orc_sim* s, int* used) { return s->o->residual(used); }
int orc_build_campfire(int seed, float* out, int max_logs) { return build_campfire(seed, out, max_logs); }
int orc_set_spots(orc_sim* s, const fs_spot* sp, int n, fs_spot* out) { return s->o->set_spots(sp, n, out); }
void orc_set_face_flux(orc_sim* s, const float* f) { s->o->set_face(f); }
void orc_budget_enable(orc_sim* s, int on) { s->o->budget_enable(on); }
// The exact level solve alone on an nx x ny x nz level of cell size h (natural x-fastest order):
// p = A^-1 b in f32 (fp64 = 0) or f64.  Returns 0, or -1 if the library's kernels do not take the shape.
int orc_direct_solve(int nx, int ny, int nz, double h, int fp64, const double* b, double* p) {
    auto run = [&](auto tag) {
        using T = decltype(tag);
        if (!Multigrid<T>::direct_fits(nx, ny, nz)) return -1;
        Multigrid<T> mg;
        mg.build(nx, ny, nz, h, 1);
        mg.setup_direct(0);
        for (size_t i = 0; i < mg.L[0].b.size(); i++) mg.L[0].b[i] = (T)b[i];
        mg.direct_solve();
        for (size_t i = 0; i < mg.L[0].p.size(); i++) p[i] = (double)mg.L[0].p[i];
        return 0;
    };
    return fp64 ? run(0.0) : run(0.0f);
}
void orc_budget_read(orc_sim* s, double* out) { s->o->budget_read(out); }
// Double-precision spot / floor fluxes replacing the float ones (NULL: back to the float values).
void orc_set_flux64(orc_sim* s, const double* spot, int n, const double* face) { s->o->set_flux64(spot, n, face); }

// Noise bake alone (for the noise-texture golden fixtures).
void orc_bake_noise(int n, int seed, float* noise_out, float* curl_out) {
    NoiseTex t, c; bake_noise(t, n, seed); bake_curl(c, t);
    std::copy(t.v.begin(), t.v.end(), noise_out);
    if (curl_out) std::copy(c.v.begin(), c.v.end(), curl_out);
}
void orc_hash3f(int x, int y, int z, float* out) { hash3f(x, y, z, out); }
void orc_f16_round(const float* in, float* out, long n) { for (long i = 0; i < n; i++) out[i] = f16_to_f32(f32_to_f16(in[i])); }
void orc_bf16_round(const float* in, float* out, long n) { for (long i = 0; i < n; i++) out[i] = bf16_to_f32(f32_to_bf16(in[i])); }
// The f32 transcendental functions of the oracle (and of the kernels): which = 0 exp(a), 1 log(a),
// 2 pow(a, b), 3 sin(a), 4 cos(a).
void orc_fmath(int which, const float* a, const float* b, float* out, long n) {
    for (long i = 0; i < n; i++) {
        switch (which) {
            case 0: out[i] = fexp(a[i]); break;
            case 1: out[i] = flog(a[i]); break;
            case 2: out[i] = fpow(a[i], b[i]); break;
            case 3: out[i] = fsin(a[i]); break;
            default: out[i] = fcos(a[i]); break;
        }
    }
}

// ---- stand-alone sim.js multigrid (isolated Poisson), FP32 or FP64 storage ----
struct orc_mg { Multigrid<float>* f; Multigrid<double>* d; std::vector<double> x; };
orc_mg* orc_mg_create(int nx, int ny, int nz, double h, int max_levels, int fp64) {
    orc_mg* m = new orc_mg{nullptr, nullptr, {}};
    h = (double)(float)h;   // the level spacing is an f32 uniform (sim.js:96)
    if (fp64) { m->d = new Multigrid<double>(); m->d->build(nx, ny, nz, h, max_levels); }
    else { m->f = new Multigrid<float>(); m->f->build(nx, ny, nz, h, max_levels); }
    m->x.assign((size_t)nx * ny * nz, 0.0);
    return m;
}
void orc_mg_destroy(orc_mg* m) { delete m->f; delete m->d; delete m; }
void orc_mg_set_omega(orc_mg* m, float omega) { if (m->f) m->f->omega = omega; else m->d->omega = omega; }
int orc_mg_levels(orc_mg* m, int* dims, int* coarse) {
    int n = m->f ? (int)m->f->L.size() : (int)m->d->L.size();
    for (int i = 0; i < n; i++) {
        dims[i * 3] = m->f ? m->f->L[i].nx : m->d->L[i].nx;
        dims[i * 3 + 1] = m->f ? m->f->L[i].ny : m->d->L[i].ny;
        dims[i * 3 + 2] = m->f ? m->f->L[i].nz : m->d->L[i].nz;
    }
    *coarse = m->f ? m->f->coarse_iters : m->d->coarse_iters;
    return n;
}
void orc_mg_set(orc_mg* m, const double* b, const double* p) {
    size_t n = m->x.size();
    for (size_t i = 0; i < n; i++) {
        if (m->f) { if (b) m->f->L[0].b[i] = (float)b[i]; if (p) m->f->L[0].p[i] = (float)p[i]; }
        else { if (b) m->d->L[0].b[i] = b[i]; if (p) m->d->L[0].p[i] = p[i]; }
        if (p) m->x[i] = p[i];
    }
}
void orc_mg_get(orc_mg* m, double* p) {
    size_t n = m->x.size();
    for (size_t i = 0; i < n; i++) p[i] = m->f ? (double)m->f->L[0].p[i] : m->d->L[0].p[i];
}
// Plain V-cycles on p; hist[k] = relative FP64 residual after cycle k.
void orc_mg_vcycles(orc_mg* m, int cycles, double* hist) {
    for (int k = 0; k < cycles; k++) {
        double rr, bb;
        if (m->f) { m->f->vcycle(); m->f->residual64(m->f->L[0].p.data(), nullptr, &rr, &bb); }
        else { m->d->vcycle(); m->d->residual64(m->d->L[0].p.data(), nullptr, &rr, &bb); }
        if (hist) hist[k] = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    }
}
// single colour pass / restriction / prolongation on a level (per-operator tests)
void orc_mg_op(orc_mg* m, int op, int level, int arg) {
    if (m->f) { if (op == 0) m->f->smooth_color(level, arg); else if (op == 1) m->f->restrict_(level); else m->f->prolong(level); }
    else { if (op == 0) m->d->smooth_color(level, arg); else if (op == 1) m->d->restrict_(level); else m->d->prolong(level); }
}
void orc_mg_level_get(orc_mg* m, int level, double* p, double* b) {
    if (m->f) { auto& l = m->f->L[level]; for (size_t i = 0; i < l.p.size(); i++) { if (p) p[i] = l.p[i]; if (b) b[i] = l.b[i]; } }
    else { auto& l = m->d->L[level]; for (size_t i = 0; i < l.p.size(); i++) { if (p) p[i] = l.p[i]; if (b) b[i] = l.b[i]; } }
}
void orc_mg_level_set(orc_mg* m, int level, const double* p, const double* b) {
    if (m->f) { auto& l = m->f->L[level]; for (size_t i = 0; i < l.p.size(); i++) { if (p) l.p[i] = (float)p[i]; if (b) l.b[i] = (float)b[i]; } }
    else { auto& l = m->d->L[level]; for (size_t i = 0; i < l.p.size(); i++) { if (p) l.p[i] = p[i]; if (b) l.b[i] = b[i]; } }
}
// FP32 V-cycles inside FP64 defect correction (FP32 storage only); returns cycles used
int orc_mg_refine(orc_mg* m, double tol, int max_cycles, double* hist) {
    if (!m->f) return -1;
    std::vector<float> bk = m->f->L[0].b;
    double rel;
    int used = refine64(*m->f, m->x, bk, tol, max_cycles, &rel, hist);
    for (size_t i = 0; i < m->x.size(); i++) m->f->L[0].p[i] = (float)m->x[i];
    return used;
}
void orc_mg_get_x(orc_mg* m, double* x) { std::copy(m->x.begin(), m->x.end(), x); }
double orc_mg_residual(orc_mg* m) {
    double rr, bb;
    if (m->f) m->f->residual64(m->f->L[0].p.data(), nullptr, &rr, &bb);
    else m->d->residual64(m->d->L[0].p.data(), nullptr, &rr, &bb);
    return bb > 0 ? std::sqrt(rr / bb) : 0.0;
}

// ---- poisson.mjs study (results/poisson-128-face.txt) ----
// mode 0: Jacobi(omega) 160 its, hist at checkpoints 5,10,20,40,80,160
// mode 1: RB-SOR(omega), same checkpoints;  mode 2: MG V(pre,post) 8 cycles
// Returns r0 (initial rms residual); hist gets residual / r0.
double orc_mjs_study(int N, int face, int mode, double omega, int pre, int post, double* hist) {
    MjsPoisson P; P.face = face != 0;
    std::vector<float> B = MjsPoisson::rhs(N);
    auto fresh = [&](int n) { MjsLevel L; L.n = n; size_t c = (size_t)n * n * n; L.p.assign(c, 0.f); L.b.assign(c, 0.f); L.r.assign(c, 0.f); return L; };
    MjsLevel L0 = fresh(N); L0.b = B;
    double r0 = P.residual(L0, 1);
    const int cps[6] = {5, 10, 20, 40, 80, 160};
    if (mode == 0 || mode == 1) {
        MjsLevel L = fresh(N); L.b = B;
        std::vector<float> tmp((size_t)N * N * N);
        int j = 0;
        for (int k = 1; k <= 160; k++) {
            if (mode == 0) P.jacobi(L, 1, omega, tmp); else P.rbgs(L, 1, omega);
            if (j < 6 && k == cps[j]) { hist[j++] = P.residual(L, 1) / r0; }
        }
    } else {
        std::vector<MjsLevel> levels;
        for (int n = N; n >= 4; n >>= 1) levels.push_back(fresh(n));
        levels[0].b = B;
        for (int k = 1; k <= 8; k++) { P.vcycle(levels, 0, pre, post); hist[k - 1] = P.residual(levels[0], 1) / r0; }
    }
    return r0;
}

const char* orc_last_error(void) { return g_err.c_str(); }
int orc_num_threads(void) { return omp_get_max_threads(); }
// Thread count of the oracle's parallel loops (the CPU baseline runs once on every core, once on one).
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }


// ---- spark particles: the host side of src/sparks.js:17-46, 241-271 ------------------
struct orc_sparks {
    orc_sim* sim;
    int count;
    std::vector<float> sp;
    double rate = 3.0, carry = 0.0;
    float turb = 1.f, lifeK = 1.f, sizeK = 1.f, upK = 1.f;
    struct Burst { double t0; int n; float pos[3]; float speed, radius; double dur; float fan, size; int has_aim; float aim[3]; int emitted; };
    std::vector<Burst> bursts;
    double time = 0.0;
    long frame = 0;
};
orc_sparks* orc_sparks_create(orc_sim* s, int count) {
    orc_sparks* p = new orc_sparks();
    p->sim = s; p->count = count; p->sp.assign((size_t)count * 16, 0.f);
    return p;
}
void orc_sparks_destroy(orc_sparks* p) { delete p; }
void orc_sparks_reset(orc_sparks* p) {
    std::fill(p->sp.begin(), p->sp.end(), 0.f);
    p->carry = 0.0; p->bursts.clear(); p->time = 0.0; p->frame = 0;
}
int orc_sparks_set(orc_sparks* p, const char* name, float v) {
    const std::string n(name);
    if (n == "rate") p->rate = v; else if (n == "turb") p->turb = v; else if (n == "lifeK") p->lifeK = v;
    else if (n == "sizeK") p->sizeK = v; else if (n == "upK") p->upK = v; else if (n != "brightness") return -1;
    return 0;
}
void orc_sparks_burst(orc_sparks* p, int n, const float* pos, float speed, float radius, double dur, float fan, float size,
                      int has_aim, const float* aim) {
    orc_sparks::Burst b;
    b.t0 = p->time; b.n = n; b.speed = speed; b.radius = radius; b.dur = dur; b.fan = fan; b.size = size;
    b.has_aim = has_aim; b.emitted = 0;
    for (int k = 0; k < 3; k++) { b.pos[k] = pos[k]; b.aim[k] = aim ? aim[k] : 0.f; }
    p->bursts.push_back(b);
}
void orc_sparks_step(orc_sparks* p, double dt) {
    p->carry += p->rate * dt;
    const double nBase = std::floor(p->carry);
    p->carry -= nBase;
    int nBurst = 0;
    orc_sparks::Burst B{};
    bool have = false;
    for (auto& b : p->bursts) {
        const double frac = std::min((p->time + dt - b.t0) / b.dur, 1.0);
        const int due = (int)std::min((double)b.n, std::floor((double)b.n * frac + 0.5));
        const int k = due - b.emitted;
        if (k > 0 && !have) { have = true; nBurst = k; b.emitted += k; B = b; }
    }
    p->bursts.erase(std::remove_if(p->bursts.begin(), p->bursts.end(), [](const orc_sparks::Burst& b) { return b.emitted >= b.n; }),
                    p->bursts.end());
    double org[3]; float size[3];
    p->sim->o->box(org, size);
    SparkU U;
    U.dt = (float)dt; U.time = (float)p->time; U.frame = (float)p->frame; U.count = (float)p->count;
    U.nBase = (float)nBase; U.nBurst = (float)nBurst;
    U.burstX = have ? B.pos[0] : 0.f; U.burstY = have ? B.pos[1] : 0.f; U.burstZ = have ? B.pos[2] : 0.f;
    U.burstSpeed = have ? B.speed : 0.f; U.burstRadius = have ? B.radius : 0.f;
    U.burstFan = have ? B.fan : 0.f; U.burstSize = have ? B.size : 1.f;
    U.aimX = have && B.has_aim ? B.aim[0] : 0.f; U.aimZ = have && B.has_aim ? B.aim[2] : 0.f;
    U.ox = (float)org[0]; U.oy = (float)org[1]; U.oz = (float)org[2];
    U.sx = size[0]; U.sy = size[1]; U.sz = size[2];
    U.turb = p->turb; U.lifeK = p->lifeK; U.bedRadius = 0.2f; U.seed = 1234.f; U.sizeK = p->sizeK; U.upK = p->upK;
    p->sim->o->spark_update(U, p->sp.data(), p->count);
    p->time += dt;
    p->frame += 1;
}
void orc_sparks_read(orc_sparks* p, float* out) { std::copy(p->sp.begin(), p->sp.end(), out); }
void orc_sparks_write(orc_sparks* p, const float* in) { std::copy(in, in + p->sp.size(), p->sp.begin()); }
double orc_sparks_time(orc_sparks* p) { return p->time; }

}  // extern "C"
