// Per-cell kernels of one fire-simulation step on gfx950: emitter geometry,
// MacCormack advection (two passes), curl, the fused source/combustion/force
// update, divergence and projection, plus the init-time noise bake.
//
// Every kernel is an HBM-streaming 3D stencil.  Launch shape: 256-thread blocks
// covering 64 (x) x 4 (y) cells, one wave per 64-cell x-row so each field load
// is one 512-B (fp16) or 1-KiB (fp32) coalesced transaction per wave.  Block ids
// are remapped so each of the 8 XCDs walks one contiguous z-slab of tiles and
// the y/z stencil neighbours of a tile stay in that XCD's L2.
#include <hip/hip_runtime.h>

#include "fs_codec.h"
#include "fs_internal.h"



// ---- WGSL built-in semantics (clamp = min(max()), smoothstep, mix) -------------
__device__ __forceinline__ float wclamp(float e, float lo, float hi) { return fminf(fmaxf(e, lo), hi); }
__device__ __forceinline__ float sstep(float e0, float e1, float x) {
    float t = wclamp((x - e0) / (e1 - e0), 0.f, 1.f);
    return t * t * (3.f - 2.f * t);
}
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.f - t) + b * t; }
__device__ __forceinline__ float4 mix4(float4 a, float4 b, float t) {
    return make_float4(mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t), mixf(a.w, b.w, t));
}
__device__ __forceinline__ float4 scale4(float4 a, float s) { return make_float4(a.x * s, a.y * s, a.z * s, a.w * s); }
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// ---- exp / log / pow ----------------------------------------------------------
// WGSL leaves exp and pow to the implementation; these are a fixed textbook
// construction (identical to the oracle's, so results agree bit for bit):
//   exp: Cody-Waite reduction x = k ln2 + r (hi part of ln2 has 9 trailing zero
//        bits, so k*hi is exact for |k| < 512), degree-7 Taylor polynomial on
//        |r| <= ln2/2 in Horner form, then ldexp(., k): <= 1.2 ulp.
//   log: frexp to m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
//        log m = 2s(1 + s^2/3 + s^4/5 + s^6/7 + s^8/9) + e ln2: <= 1.8 ulp.
//   pow(a, b) = exp(b log a) for finite a > 0.
__device__ __forceinline__ float fs_exp(float x) {
    if (x != x) return x;
    if (x > 88.72283f) return __int_as_float(0x7f800000);
    if (x < -103.97208f) return 0.0f;
    const float k = rintf(x * 1.44269504f);
    const float r = (x - k * 0.693145751953125f) - k * 1.42860682e-6f;
    float p = 1.0f / 5040.0f;
    p = 1.0f / 720.0f + r * p;
    p = 1.0f / 120.0f + r * p;
    p = 1.0f / 24.0f + r * p;
    p = 1.0f / 6.0f + r * p;
    p = 0.5f + r * p;
    p = 1.0f + r * p;
    p = 1.0f + r * p;
    return ldexpf(p, (int)k);
}
__device__ __forceinline__ float fs_log(float a) {
    if (a != a || a < 0.0f) return __int_as_float(0x7fc00000);
    if (a == 0.0f) return -__int_as_float(0x7f800000);
    if (a == __int_as_float(0x7f800000)) return a;
    int e;
    float m = frexpf(a, &e);
    if (m < 0.70710678f) { m = m * 2.0f; e = e - 1; }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float q = 1.0f / 9.0f;
    q = 1.0f / 7.0f + s2 * q;
    q = 1.0f / 5.0f + s2 * q;
    q = 1.0f / 3.0f + s2 * q;
    const float t = 2.0f * s;
    const float fe = (float)e;
    return fe * 0.693145751953125f + (fe * 1.42860682e-6f + (t + t * (s2 * q)));
}
__device__ __forceinline__ float fs_pow(float a, float b) {
    if (a == 1.0f || b == 0.0f) return 1.0f;
    if (a == 0.0f) return b > 0.0f ? 0.0f : __int_as_float(0x7f800000);
    return fs_exp(b * fs_log(a));
}

// ---- tiling -----------------------------------------------------------------
__device__ __forceinline__ int xcd_remap(int hw, int total) {
    const int xcd = hw & 7, q = total >> 3, r = total & 7;
    return xcd * q + min(xcd, r) + (hw >> 3);
}
struct Cell { int x, y, z, i; };
__device__ __forceinline__ bool cell_of(const Dims& g, Cell& c) {
    const int lb = xcd_remap((int)blockIdx.x, g.nblocks);
    const int bx = lb % g.nbx;
    const int t = lb / g.nbx;
    const int by = t % g.nby;
    c.z = t / g.nby;
    c.x = bx * TILE_X + ((int)threadIdx.x & 63);
    c.y = by * TILE_Y + ((int)threadIdx.x >> 6);
    c.i = (c.z * g.ny + c.y) * g.nx + c.x;
    return c.x < g.nx && c.y < g.ny;
}
Dims make_dims(int nx, int ny, int nz) {
    Dims g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.nbx = (nx + TILE_X - 1) / TILE_X;
    g.nby = (ny + TILE_Y - 1) / TILE_Y;
    g.nblocks = g.nbx * g.nby * nz;
    return g;
}

// clamped texel load (ld(), src/sim.js:114)
template <class S>
__device__ __forceinline__ float4 ldc(const typename S::raw* f, const Dims& g, int x, int y, int z) {
    x = clampi(x, 0, g.nx - 1); y = clampi(y, 0, g.ny - 1); z = clampi(z, 0, g.nz - 1);
    return Codec<S>::ld(f, (z * g.ny + y) * g.nx + x);
}
// trilinear, clamp-to-edge, at cell-space position b (texel centres at i + 0.5).
// Clamping the coordinate to [-1, n] first leaves the clamped-texel result unchanged.
template <class S>
__device__ __forceinline__ float4 sample_clamp(const typename S::raw* f, const Dims& g, float bx, float by, float bz) {
    float tx = wclamp(bx - 0.5f, -1.f, (float)g.nx);
    float ty = wclamp(by - 0.5f, -1.f, (float)g.ny);
    float tz = wclamp(bz - 0.5f, -1.f, (float)g.nz);
    const float flx = floorf(tx), fly = floorf(ty), flz = floorf(tz);
    const int ix = (int)flx, iy = (int)fly, iz = (int)flz;
    const float fx = tx - flx, fy = ty - fly, fz = tz - flz;
    const int x0 = clampi(ix, 0, g.nx - 1), x1 = clampi(ix + 1, 0, g.nx - 1);
    const int y0 = clampi(iy, 0, g.ny - 1), y1 = clampi(iy + 1, 0, g.ny - 1);
    const int z0 = clampi(iz, 0, g.nz - 1), z1 = clampi(iz + 1, 0, g.nz - 1);
    const int r00 = (z0 * g.ny + y0) * g.nx, r10 = (z0 * g.ny + y1) * g.nx;
    const int r01 = (z1 * g.ny + y0) * g.nx, r11 = (z1 * g.ny + y1) * g.nx;
    const float4 c000 = Codec<S>::ld(f, r00 + x0), c100 = Codec<S>::ld(f, r00 + x1);
    const float4 c010 = Codec<S>::ld(f, r10 + x0), c110 = Codec<S>::ld(f, r10 + x1);
    const float4 c001 = Codec<S>::ld(f, r01 + x0), c101 = Codec<S>::ld(f, r01 + x1);
    const float4 c011 = Codec<S>::ld(f, r11 + x0), c111 = Codec<S>::ld(f, r11 + x1);
    const float4 x00 = mix4(c000, c100, fx), x10 = mix4(c010, c110, fx);
    const float4 x01 = mix4(c001, c101, fx), x11 = mix4(c011, c111, fx);
    return mix4(mix4(x00, x10, fy), mix4(x01, x11, fy), fz);
}
// trilinear with repeat addressing at normalised coordinate q (noise textures, n = 2^k)
__device__ __forceinline__ float4 ld_h4(const half4_t* t, int i) {
    half4_t h = t[i];
    return make_float4((float)h.x, (float)h.y, (float)h.z, (float)h.w);
}
__device__ __forceinline__ float4 sample_rep(const half4_t* t, int n, float qx, float qy, float qz) {
    const float tx = qx * (float)n - 0.5f, ty = qy * (float)n - 0.5f, tz = qz * (float)n - 0.5f;
    const float flx = floorf(tx), fly = floorf(ty), flz = floorf(tz);
    const int ix = (int)flx, iy = (int)fly, iz = (int)flz;
    const float fx = tx - flx, fy = ty - fly, fz = tz - flz;
    const int m = n - 1;
    const int x0 = ix & m, x1 = (ix + 1) & m, y0 = iy & m, y1 = (iy + 1) & m, z0 = iz & m, z1 = (iz + 1) & m;
    const int r00 = (z0 * n + y0) * n, r10 = (z0 * n + y1) * n, r01 = (z1 * n + y0) * n, r11 = (z1 * n + y1) * n;
    const float4 x00 = mix4(ld_h4(t, r00 + x0), ld_h4(t, r00 + x1), fx);
    const float4 x10 = mix4(ld_h4(t, r10 + x0), ld_h4(t, r10 + x1), fx);
    const float4 x01 = mix4(ld_h4(t, r01 + x0), ld_h4(t, r01 + x1), fx);
    const float4 x11 = mix4(ld_h4(t, r11 + x0), ld_h4(t, r11 + x1), fx);
    return mix4(mix4(x00, x10, fy), mix4(x01, x11, fy), fz);
}

// ---- geo: emitter geometry (src/sim.js:498-523, sdCapsule src/shaders/lib.js:58-62)
template <class S>
__global__ __launch_bounds__(256) void k_geo(const DevParams* __restrict__ P, Dims g, const float* __restrict__ logs,
                                             typename S::raw* __restrict__ geo) {
    Cell c;
    if (!cell_of(g, c)) return;
    const float dx = P->f[P_dx];
    const float px = P->f[P_ox] + ((float)c.x + 0.5f) * dx;
    const float py = P->f[P_oy] + ((float)c.y + 0.5f) * dx;
    const float pz = P->f[P_oz] + ((float)c.z + 0.5f) * dx;
    const float pxz = sqrtf(px * px + pz * pz);
    const int nl = P->n_logs;
    float dmin = 1e9f, band = 0.f;
    for (int l = 0; l < nl; l++) {
        const float* A = logs + l * 8;
        const float* B = A + 4;
        const float pax = px - A[0], pay = py - A[1], paz = pz - A[2];
        const float bax = B[0] - A[0], bay = B[1] - A[1], baz = B[2] - A[2];
        const float h = wclamp((pax * bax + pay * bay + paz * baz) / (bax * bax + bay * bay + baz * baz), 0.f, 1.f);
        const float ex = pax - bax * h, ey = pay - bay * h, ez = paz - baz * h;
        float dd = sqrtf(ex * ex + ey * ey + ez * ez) - A[3];
        dd += A[3] * 0.3f * (1.f - sstep(0.f, 0.16f, pxz)) * B[3];
        dmin = fminf(dmin, dd);
        const float bnd = 1.f - sstep(0.f, 1.6f * dx, fabsf(dd - 0.6f * dx));
        const float hot = B[3] * (1.f - sstep(P->f[P_hotR0], P->f[P_hotR1], pxz)) * (1.f - sstep(0.08f, P->f[P_hotY1], py));
        band = fmaxf(band, bnd * hot);
    }
    const float bedR = P->f[P_bedRadius];
    const float bedH = 0.05f * (1.f - sstep(0.f, bedR, pxz)) + 0.02f;
    const float bed = (1.f - sstep(0.f, 2.f * dx, py - bedH)) * (1.f - sstep(bedR * 0.6f, bedR * 1.15f, pxz));
    const float solid = dmin < -0.5f * dx ? 1.f : 0.f;
    Codec<S>::st(geo, c.i, make_float4(band, bed, solid, dmin));
}

// ---- advection (src/sim.js:123-197) -----------------------------------------
template <class S>
__device__ __forceinline__ float3 backtrace(const typename S::raw* vel, const Dims& g, float cx, float cy, float cz,
                                            float dt, float dx) {   // sim.js:130-135
    const float4 v0 = sample_clamp<S>(vel, g, cx, cy, cz);
    const float hx = cx - v0.x * (0.5f * dt) / dx, hy = cy - v0.y * (0.5f * dt) / dx, hz = cz - v0.z * (0.5f * dt) / dx;
    const float4 v1 = sample_clamp<S>(vel, g, hx, hy, hz);
    return make_float3(cx - v1.x * dt / dx, cy - v1.y * dt / dx, cz - v1.z * dt / dx);
}
__device__ __forceinline__ float outside_fade(const Dims& g, float x, float y, float z) {   // sim.js:137-140
    const float lo = fminf(x, z);
    const float hi = fmaxf(x - (float)g.nx, fmaxf(y - (float)g.ny, z - (float)g.nz));
    return wclamp(1.f - fmaxf(-lo, hi), 0.f, 1.f);
}

template <class S>
__global__ __launch_bounds__(256) void k_adv1(const DevParams* __restrict__ P, Dims g,
                                              const typename S::raw* __restrict__ vel, const typename S::raw* __restrict__ sF,
                                              const typename S::raw* __restrict__ s2F, typename S::raw* __restrict__ velX,
                                              typename S::raw* __restrict__ sX, typename S::raw* __restrict__ s2X) {
    Cell c;
    if (!cell_of(g, c)) return;
    const float dt = (float)P->dt, dx = P->f[P_dx];
    const float cx = (float)c.x + 0.5f, cy = (float)c.y + 0.5f, cz = (float)c.z + 0.5f;
    const float3 b = backtrace<S>(vel, g, cx, cy, cz, dt, dx);
    const float f = outside_fade(g, b.x, b.y, b.z);
    const float ws = sstep(0.03f, 0.4f, b.y * dx);
    const float wx = P->f[P_windX] * ws, wy = 0.f * ws, wz = P->f[P_windZ] * ws;
    const float4 sv = sample_clamp<S>(vel, g, b.x, b.y, b.z);
    Codec<S>::st(velX, c.i, make_float4(mixf(wx, sv.x, f), mixf(wy, sv.y, f), mixf(wz, sv.z, f), 0.f));
    Codec<S>::st(sX, c.i, scale4(sample_clamp<S>(sF, g, b.x, b.y, b.z), f));
    Codec<S>::st(s2X, c.i, scale4(sample_clamp<S>(s2F, g, b.x, b.y, b.z), f));
}

__device__ __forceinline__ float clamp_lim(float v, float mn, float mx, float h) { return wclamp(v, fminf(mn, h), fmaxf(mx, h)); }

template <class S>
__global__ __launch_bounds__(256) void k_adv2(const DevParams* __restrict__ P, Dims g,
                                              const typename S::raw* __restrict__ vel, const typename S::raw* __restrict__ sF,
                                              const typename S::raw* __restrict__ s2F, const typename S::raw* __restrict__ velH,
                                              const typename S::raw* __restrict__ sH, const typename S::raw* __restrict__ s2H,
                                              typename S::raw* __restrict__ velY, typename S::raw* __restrict__ sY,
                                              typename S::raw* __restrict__ s2Y) {
    Cell c;
    if (!cell_of(g, c)) return;
    const float dt = (float)P->dt, dx = P->f[P_dx];
    const float cx = (float)c.x + 0.5f, cy = (float)c.y + 0.5f, cz = (float)c.z + 0.5f;
    const float4 v = Codec<S>::ld(vel, c.i);
    // reverse advection of the forward result (sim.js:172-178)
    const float fwx = cx + dt * v.x / dx, fwy = cy + dt * v.y / dx, fwz = cz + dt * v.z / dx;
    const float4 vH = Codec<S>::ld(velH, c.i);
    const float4 sHc = Codec<S>::ld(sH, c.i);
    const float4 s2Hc = Codec<S>::ld(s2H, c.i);
    const float4 vF = sample_clamp<S>(velH, g, fwx, fwy, fwz);
    const float4 sF4 = sample_clamp<S>(sH, g, fwx, fwy, fwz);
    const float4 s2F4 = sample_clamp<S>(s2H, g, fwx, fwy, fwz);
    const float4 s0 = Codec<S>::ld(sF, c.i), s20 = Codec<S>::ld(s2F, c.i);
    float vNx = vH.x + (v.x - vF.x) * 0.5f, vNy = vH.y + (v.y - vF.y) * 0.5f, vNz = vH.z + (v.z - vF.z) * 0.5f;
    float4 sN = make_float4(sHc.x + (s0.x - sF4.x) * 0.5f, sHc.y + (s0.y - sF4.y) * 0.5f,
                            sHc.z + (s0.z - sF4.z) * 0.5f, sHc.w + (s0.w - sF4.w) * 0.5f);
    float4 s2N = make_float4(s2Hc.x + (s20.x - s2F4.x) * 0.5f, s2Hc.y + (s20.y - s2F4.y) * 0.5f,
                             s2Hc.z + (s20.z - s2F4.z) * 0.5f, s2Hc.w + (s20.w - s2F4.w) * 0.5f);
    // Selle et al. 2008 limiter over the 8 texels at floor(backtrace - 0.5) (sim.js:180-193)
    const float3 bt = backtrace<S>(vel, g, cx, cy, cz, dt, dx);
    const float bx = bt.x - 0.5f, by = bt.y - 0.5f, bz = bt.z - 0.5f;
    const int b0x = (int)floorf(wclamp(bx, -2.f, (float)g.nx + 1.f));
    const int b0y = (int)floorf(wclamp(by, -2.f, (float)g.ny + 1.f));
    const int b0z = (int)floorf(wclamp(bz, -2.f, (float)g.nz + 1.f));
    float4 vmn = make_float4(1e9f, 1e9f, 1e9f, 0.f), vmx = make_float4(-1e9f, -1e9f, -1e9f, 0.f);
    float4 smn = make_float4(1e9f, 1e9f, 1e9f, 1e9f), smx = make_float4(-1e9f, -1e9f, -1e9f, -1e9f);
    float4 tmn = smn, tmx = smx;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int qx = b0x + (k & 1), qy = b0y + ((k >> 1) & 1), qz = b0z + ((k >> 2) & 1);
        const float4 vv = ldc<S>(vel, g, qx, qy, qz), ss = ldc<S>(sF, g, qx, qy, qz), tt = ldc<S>(s2F, g, qx, qy, qz);
        vmn.x = fminf(vmn.x, vv.x); vmn.y = fminf(vmn.y, vv.y); vmn.z = fminf(vmn.z, vv.z);
        vmx.x = fmaxf(vmx.x, vv.x); vmx.y = fmaxf(vmx.y, vv.y); vmx.z = fmaxf(vmx.z, vv.z);
        smn.x = fminf(smn.x, ss.x); smn.y = fminf(smn.y, ss.y); smn.z = fminf(smn.z, ss.z); smn.w = fminf(smn.w, ss.w);
        smx.x = fmaxf(smx.x, ss.x); smx.y = fmaxf(smx.y, ss.y); smx.z = fmaxf(smx.z, ss.z); smx.w = fmaxf(smx.w, ss.w);
        tmn.x = fminf(tmn.x, tt.x); tmn.y = fminf(tmn.y, tt.y); tmn.z = fminf(tmn.z, tt.z); tmn.w = fminf(tmn.w, tt.w);
        tmx.x = fmaxf(tmx.x, tt.x); tmx.y = fmaxf(tmx.y, tt.y); tmx.z = fmaxf(tmx.z, tt.z); tmx.w = fmaxf(tmx.w, tt.w);
    }
    if (outside_fade(g, bx + 0.5f, by + 0.5f, bz + 0.5f) < 1.f) {   // sim.js:190
        vNx = vH.x; vNy = vH.y; vNz = vH.z; sN = sHc; s2N = s2Hc;
    }
    vNx = clamp_lim(vNx, vmn.x, vmx.x, vH.x); vNy = clamp_lim(vNy, vmn.y, vmx.y, vH.y); vNz = clamp_lim(vNz, vmn.z, vmx.z, vH.z);
    sN.x = clamp_lim(sN.x, smn.x, smx.x, sHc.x); sN.y = clamp_lim(sN.y, smn.y, smx.y, sHc.y);
    sN.z = clamp_lim(sN.z, smn.z, smx.z, sHc.z); sN.w = clamp_lim(sN.w, smn.w, smx.w, sHc.w);
    s2N.x = clamp_lim(s2N.x, tmn.x, tmx.x, s2Hc.x); s2N.y = clamp_lim(s2N.y, tmn.y, tmx.y, s2Hc.y);
    s2N.z = clamp_lim(s2N.z, tmn.z, tmx.z, s2Hc.z); s2N.w = clamp_lim(s2N.w, tmn.w, tmx.w, s2Hc.w);
    Codec<S>::st(velY, c.i, make_float4(vNx, vNy, vNz, 0.f));
    Codec<S>::st(sY, c.i, make_float4(fmaxf(sN.x, 0.f), fmaxf(sN.y, 0.f), fmaxf(sN.z, 0.f), fmaxf(sN.w, 0.f)));
    Codec<S>::st(s2Y, c.i, make_float4(fmaxf(s2N.x, 0.f), fmaxf(s2N.y, 0.f), fmaxf(s2N.z, 0.f), fmaxf(s2N.w, 0.f)));
}

// ---- curl: w = curl(v) * dx, |w| (src/sim.js:203-211) ---------------------------
template <class S>
__global__ __launch_bounds__(256) void k_curl(Dims g, const typename S::raw* __restrict__ v, typename S::raw* __restrict__ vort) {
    Cell c;
    if (!cell_of(g, c)) return;
    const float4 L = ldc<S>(v, g, c.x - 1, c.y, c.z), R = ldc<S>(v, g, c.x + 1, c.y, c.z);
    const float4 B = ldc<S>(v, g, c.x, c.y - 1, c.z), T = ldc<S>(v, g, c.x, c.y + 1, c.z);
    const float4 K = ldc<S>(v, g, c.x, c.y, c.z - 1), F = ldc<S>(v, g, c.x, c.y, c.z + 1);
    const float wx = ((T.z - B.z) - (F.y - K.y)) * 0.5f;
    const float wy = ((F.x - K.x) - (R.z - L.z)) * 0.5f;
    const float wz = ((R.y - L.y) - (T.x - B.x)) * 0.5f;
    Codec<S>::st(vort, c.i, make_float4(wx, wy, wz, sqrtf(wx * wx + wy * wy + wz * wz)));
}

// ---- sources, combustion, soot, cooling, buoyancy, confinement, curl noise,
//      wind, sponges and clamps: one fused cell update (src/sim.js:230-396) -------
template <class S>
__global__ __launch_bounds__(256) void k_forces(const DevParams* __restrict__ P, Dims g,
                                                const typename S::raw* __restrict__ velY, const typename S::raw* __restrict__ sY,
                                                const typename S::raw* __restrict__ s2Y, const typename S::raw* __restrict__ vort,
                                                const typename S::raw* __restrict__ geo, const half4_t* __restrict__ noise,
                                                const half4_t* __restrict__ curlT, int nn, const fs_spot* __restrict__ spots,
                                                const float* __restrict__ faceFlux, typename S::raw* __restrict__ velOut,
                                                typename S::raw* __restrict__ sOut, typename S::raw* __restrict__ s2Out) {
    Cell c;
    if (!cell_of(g, c)) return;
    const float* U = P->f;
    const float dt = (float)P->dt, dx = U[P_dx], tm = (float)P->time, I = U[P_intensity];
    const float px = U[P_ox] + ((float)c.x + 0.5f) * dx;
    const float py = U[P_oy] + ((float)c.y + 0.5f) * dx;
    const float pz = U[P_oz] + ((float)c.z + 0.5f) * dx;
    const float4 v4 = Codec<S>::ld(velY, c.i);
    float vx = v4.x, vy = v4.y, vz = v4.z;
    float4 s = Codec<S>::ld(sY, c.i);
    float4 s2 = Codec<S>::ld(s2Y, c.i);
    const float4 gg = Codec<S>::ld(geo, c.i);

    // ---- sources (sim.js:246-297)
    float src = 0.f, anchor = 0.f;
    if (gg.x > 0.001f) {
        const float af = U[P_anchorFreq];
        const float4 na = sample_rep(noise, nn, px * af + U[P_seed] * 0.37f, py * af + 0.f, pz * af + 0.11f);
        anchor = sstep(-0.05f, 0.25f, na.z * 2.f - 1.f);
        const float sf = U[P_srcFreq] * 0.06f;
        const float4 nq = sample_rep(noise, nn, px * sf + U[P_seed] * 0.131f, py * sf + 0.f, pz * sf + tm * U[P_srcSpeed]);
        const float pulse = sstep(U[P_srcThresh] - 0.25f, U[P_srcThresh] + 0.25f, nq.x * 2.f - 1.f);
        src = wclamp(gg.x * U[P_logSrc] * (U[P_srcFloor] + U[P_anchorGain] * anchor) * mixf(0.55f, 1.f, pulse) * 1.2f, 0.f, 1.f) * I;
        src *= 1.f - gg.z;
    }
    const float srcW = fminf(src, 1.f);
    const float bedHot = gg.y;
    float bed = gg.y * wclamp(I * 1.4f, 0.3f, 1.f);
    if (bed > 0.001f) {
        const float4 nb = sample_rep(noise, nn, px * 2.4f + 0.17f, py * 2.4f + -tm * 0.12f, pz * 2.4f + tm * 0.05f);
        bed *= sstep(0.f, 0.3f, nb.y * 2.f - 1.f);
    }
    s.y = fmaxf(s.y, src * U[P_fuelSrc]);
    s.y = fmaxf(s.y, bed * U[P_bedFuel]);
    s.x = fmaxf(s.x, src * U[P_srcTemp]);
    s.x = fmaxf(s.x, bedHot * U[P_bedTemp]);
    float srcAll = fmaxf(srcW, bed);
    s2.w = mixf(s2.w, 1.f, bed * (1.f - srcW));
    s2.w = mixf(s2.w, 0.f, srcW);
    s2.y = s2.y * (1.f - srcAll);
    s2.x = fmaxf(s2.x, (1.f - s.y) * srcW);
    s2.x = fmaxf(s2.x, (1.f - s.y) * bed * 0.6f);
    const float wsrc = 1.f - fs_pow(1.f - wclamp(srcW + bed * 0.4f, 0.f, 1.f) * 0.5f, dt * 90.f);
    vy = mixf(vy, fmaxf(vy, U[P_srcVel]), wsrc);
    s2.z += gg.x * (1.f - anchor) * U[P_smolder] * dt * I;
    if (U[P_burstAmt] > 0.f) {
        const float ddx = px - U[P_burstX], ddy = py - U[P_burstY], ddz = pz - U[P_burstZ];
        const float bd = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
        const float bw = wclamp(fs_exp(-bd * bd / 0.006f) * U[P_burstAmt], 0.f, 1.f);
        s.y = fmaxf(s.y, 0.17f * bw);
        s.x = fmaxf(s.x, 0.9f * bw);
        s2.x *= 1.f - bw;
        s2.w *= 1.f - bw;
        s2.y = fmaxf(s2.y, 0.05f * bw * wclamp(s.y + s2.x, 0.f, 1.f));
    }
    // ---- pyrolysis spots and floor fuel flux (build extension: mass-flux sources)
    for (int k = 0; k < P->n_spots; k++) {
        const fs_spot sp = spots[k];
        const float ddx = px - sp.pos[0], ddy = py - sp.pos[1], ddz = pz - sp.pos[2];
        const float d = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
        if (d < sp.radius) {
            const float w = 1.f - sstep(0.5f * sp.radius, sp.radius, d);
            s.y = fminf(s.y + sp.flux * dt * sp.norm * w, 1.f);
            s.x = fmaxf(s.x, w * sp.temp);
            s2.w = mixf(s2.w, 0.f, w);
            s2.y = s2.y * (1.f - w);
            srcAll = fmaxf(srcAll, w);
        }
    }
    if (P->has_face_flux && c.y == 0) s.y = fminf(s.y + faceFlux[c.z * g.nx + c.x] * dt / (FS_RHO0 * dx), 1.f);
    float ageF = s2.y / fmaxf(s.y + s2.x, 1e-4f);

    // ---- mixing-limited combustion (sim.js:300-309)
    const float T0 = 293.f + 1000.f * s.x;
    const float o = wclamp(1.f - s.y - s2.x, 0.f, 1.f);
    const float ign = sstep(U[P_ignT] - 50.f, U[P_ignT] + 50.f, T0);
    const float b = fminf(s.y, o / U[P_rs]) * (1.f - fs_exp(-U[P_kMix] * ign * dt));
    s.y -= b;
    s2.x = fminf(s2.x + (1.f + U[P_rs]) * b, 1.f);
    const float dT = fminf(U[P_heatOut] * b, fmaxf(2100.f - T0, 0.f));
    float T = T0 + dT;
    const float q = dT / dt;
    const float phi = gg.z > 0.5f ? 0.f : fminf(U[P_expansion] * q / T, 60.f);

    // ---- soot (sim.js:312-318)
    ageF = fminf(ageF + dt, 4.f);
    s.z += U[P_sootForm] * b * sstep(0.015f, 0.05f, ageF) * sstep(1000.f, 1400.f, T) * (1.f - s2.w);
    s.z *= fs_exp(-U[P_sootOx] * (0.25f + o) * sstep(1100.f, 1400.f, T) * dt);
    s.z *= fs_exp(-U[P_sootOx] * 0.5f * o * sstep(850.f, 1100.f, T) * (1.f - sstep(1100.f, 1300.f, T)) * dt);
    s.z *= fs_exp(-U[P_sootDissip] * dt);

    // ---- cooling (sim.js:321-330)
    const float th = T * 0.001f;
    T -= fminf(U[P_radCool] * 970.f * s.z * th * th * (th * th * th * th - 0.00737f), 6000.f) * dt;
    const float notBurning = 1.f - sstep(500.f, 5000.f, q);
    T = 293.f + fmaxf(T - 293.f, 0.f) * fs_exp(-U[P_coolRate] * notBurning * dt);
    s.x = (T - 293.f) * 0.001f;
    s.w = q * 1e-4f;
    s.y *= fs_exp(-U[P_fuelDissip] * dt);
    s2.x *= fs_exp(-U[P_vitDecay] * dt);
    s2.z *= fs_exp(-U[P_smokeDissip] * dt);
    s2.y = ageF * wclamp(s.y + s2.x, 0.f, 1.f);

    // ---- buoyancy (sim.js:333)
    vy += U[P_buoyancy] * s.x * dt;

    // ---- hot-gas viscosity (sim.js:339-346)
    if (U[P_visc] > 0.f) {
        const float4 a = ldc<S>(velY, g, c.x + 1, c.y, c.z), bb = ldc<S>(velY, g, c.x - 1, c.y, c.z);
        const float4 cc = ldc<S>(velY, g, c.x, c.y + 1, c.z), dd = ldc<S>(velY, g, c.x, c.y - 1, c.z);
        const float4 ee = ldc<S>(velY, g, c.x, c.y, c.z + 1), ff = ldc<S>(velY, g, c.x, c.y, c.z - 1);
        const float4 me = Codec<S>::ld(velY, c.i);
        const float lx = a.x + bb.x + cc.x + dd.x + ee.x + ff.x - me.x * 6.f;
        const float ly = a.y + bb.y + cc.y + dd.y + ee.y + ff.y - me.y * 6.f;
        const float lz = a.z + bb.z + cc.z + dd.z + ee.z + ff.z - me.z * 6.f;
        const float nu = U[P_visc] * 1.5e-5f * fs_pow(fmaxf(T, 293.f) / 293.f, 1.7f);
        const float k = fminf(nu * dt / (dx * dx), 0.15f);
        vx = vx + lx * k; vy = vy + ly * k; vz = vz + lz * k;
    }
    // ---- laminar base gate (sim.js:350)
    const float lamAge = U[P_lamAge];
    const float lam = lamAge > 0.f ? sstep(0.5f * lamAge, fmaxf(lamAge, 1e-3f), ageF) : 1.f;

    // ---- vorticity confinement (sim.js:353-361)
    const float4 w = Codec<S>::ld(vort, c.i);
    const float gx = ldc<S>(vort, g, c.x + 1, c.y, c.z).w - ldc<S>(vort, g, c.x - 1, c.y, c.z).w;
    const float gy = ldc<S>(vort, g, c.x, c.y + 1, c.z).w - ldc<S>(vort, g, c.x, c.y - 1, c.z).w;
    const float gz = ldc<S>(vort, g, c.x, c.y, c.z + 1).w - ldc<S>(vort, g, c.x, c.y, c.z - 1).w;
    const float gl = sqrtf(gx * gx + gy * gy + gz * gz);
    if (gl > 1e-6f) {
        const float gate = sstep(450.f, 900.f, T) * (1.f - srcAll) * sstep(0.02f, 0.08f, py) * lam;
        const float nx_ = gx / gl, ny_ = gy / gl, nz_ = gz / gl;
        const float k = U[P_vortEps] * gate * dt;
        vx = vx + (ny_ * w.z - nz_ * w.y) * k;
        vy = vy + (nz_ * w.x - nx_ * w.z) * k;
        vz = vz + (nx_ * w.y - ny_ * w.x) * k;
    }
    // ---- rising curl noise (sim.js:364-368)
    const float tf = U[P_turbFreq];
    const float tqx = (px - 0.f) * tf + 0.f;
    const float tqy = (py - tm * U[P_turbRise]) * tf + 0.f;
    const float tqz = (pz - 0.f) * tf + tm * U[P_turbEvolve];
    const float4 cn = sample_rep(curlT, nn, tqx, tqy, tqz);
    const float tgate = fmaxf(sstep(500.f, 1100.f, T), sstep(0.f, 3.f, s.w)) * (1.f - srcAll) *
                        mixf(0.6f, 1.8f, sstep(0.2f, 0.75f, py)) * lam;
    {
        const float k = (U[P_turbAmp] / 12.f) * tgate * dt * 10.f;
        vx = vx + cn.x * k; vy = vy + cn.y * k; vz = vz + cn.z * k;
    }
    // ---- wind (sim.js:372-375)
    const float windK = sstep(0.05f, 0.5f, py) * (1.f + U[P_gust]);
    vx += (U[P_windX] * windK - vx) * dt * 0.35f * windK;
    vz += (U[P_windZ] * windK - vz) * dt * 0.35f * windK;
    {
        const float e = fs_exp(-U[P_velDissip] * dt);
        vx = vx * e; vy = vy * e; vz = vz * e;
    }
    // ---- solids and sponge layers (sim.js:378-388)
    if (gg.z > 0.5f) { vx = 0.f; vy = 0.f; vz = 0.f; s.y = 0.f; s2.y = 0.f; }
    const float fpx = (float)c.x + 0.5f, fpy = (float)c.y + 0.5f, fpz = (float)c.z + 0.5f;
    const float side = fminf(fminf(fpx, (float)g.nx - fpx), fminf(fpz, (float)g.nz - fpz));
    const float sp = (1.f - sstep(0.f, 5.f, side)) * U[P_sponge];
    const float wsg = sstep(0.03f, 0.4f, py);
    const float wSx = U[P_windX] * wsg, wSz = U[P_windZ] * wsg;
    const float ks = 1.f - fs_pow(1.f - 0.15f * wclamp(sp, 0.f, 1.f), dt * 90.f);
    vx = mixf(vx, wSx, ks);
    vz = mixf(vz, wSz, ks);
    const float top = sstep((float)g.ny - 10.f, (float)g.ny, fpy);
    const float kt = fs_pow(1.f - 0.1f * top, dt * 90.f);
    s.z *= kt;
    s2.z *= kt;
    // ---- clamps and NaN reset (sim.js:390-392)
    const float vl = sqrtf(vx * vx + vy * vy + vz * vz);
    if (vl > 15.f) { const float k = 15.f / vl; vx = vx * k; vy = vy * k; vz = vz * k; }
    const bool bad = isnan(s.x) || isnan(s.y) || isnan(s.z) || isnan(s.w) || isnan(s2.x) || isnan(s2.y) ||
                     isnan(s2.z) || isnan(s2.w) || isnan(vx) || isnan(vy) || isnan(vz);
    if (bad) { s = make_float4(0.f, 0.f, 0.f, 0.f); s2 = s; vx = 0.f; vy = 0.f; vz = 0.f; }
    Codec<S>::st(velOut, c.i, make_float4(vx, vy, vz, phi));
    Codec<S>::st(sOut, c.i, s);
    Codec<S>::st(s2Out, c.i, s2);
}

// ---- divergence minus the expansion target (src/sim.js:403-413) ----------------
template <class S>
__global__ __launch_bounds__(256) void k_div(DevParams* __restrict__ P, Dims g, const typename S::raw* __restrict__ v,
                                             float* __restrict__ bOut, int advance_clock) {
    Cell c;
    const bool in = cell_of(g, c);
    if (advance_clock && blockIdx.x == 0 && threadIdx.x == 0) {
        // time += dt; frame++ (sim.js:646-647).  Nothing after the forces kernel in a step reads the clock.
        P->time = P->time + P->dt;
        P->frame = P->frame + 1;
    }
    if (!in) return;
    const float dx = P->f[P_dx];
    const float4 me = Codec<S>::ld(v, c.i);
    const float L = ldc<S>(v, g, c.x - 1, c.y, c.z).x, R = ldc<S>(v, g, c.x + 1, c.y, c.z).x;
    float B = ldc<S>(v, g, c.x, c.y - 1, c.z).y;
    const float T = ldc<S>(v, g, c.x, c.y + 1, c.z).y;
    const float K = ldc<S>(v, g, c.x, c.y, c.z - 1).z, F = ldc<S>(v, g, c.x, c.y, c.z + 1).z;
    if (c.y == 0) B = -me.y;
    bOut[rb_index(c.x, c.y, c.z, g.nx, g.ny)] = ((R - L) + (T - B) + (F - K)) / (2.f * dx) - me.w;   // parity-split layout
}

// ---- projection (src/sim.js:532-550) ------------------------------------------
template <class S, class PT>
__global__ __launch_bounds__(256) void k_project(const DevParams* __restrict__ P, Dims g, const typename S::raw* __restrict__ vIn,
                                                 const PT* __restrict__ p, const typename S::raw* __restrict__ geo,
                                                 typename S::raw* __restrict__ vOut) {
    Cell c;
    if (!cell_of(g, c)) return;
    const float dx = P->f[P_dx];
    // p is in the multigrid's parity-split layout; all seven values are loaded (an outside
    // neighbour from the cell itself, unused) so the loads issue together
    const int io = rb_index(c.x, c.y, c.z, g.nx, g.ny);
    const float me = (float)p[io];
    const float vxp = (float)p[fs_opaque(c.x + 1 < g.nx ? rb_index(c.x + 1, c.y, c.z, g.nx, g.ny) : io)];
    const float vxm = (float)p[fs_opaque(c.x > 0 ? rb_index(c.x - 1, c.y, c.z, g.nx, g.ny) : io)];
    const float vyp = (float)p[fs_opaque(c.y + 1 < g.ny ? rb_index(c.x, c.y + 1, c.z, g.nx, g.ny) : io)];
    const float vym = (float)p[fs_opaque(c.y > 0 ? rb_index(c.x, c.y - 1, c.z, g.nx, g.ny) : io)];
    const float vzp = (float)p[fs_opaque(c.z + 1 < g.nz ? rb_index(c.x, c.y, c.z + 1, g.nx, g.ny) : io)];
    const float vzm = (float)p[fs_opaque(c.z > 0 ? rb_index(c.x, c.y, c.z - 1, g.nx, g.ny) : io)];
    const float4 vi = Codec<S>::ld(vIn, c.i);
    const float solid = Codec<S>::ld(geo, c.i).z;
    // ghost rules: -p outside x, z and above the top (Dirichlet), +p below the floor (Neumann)
    const float pxp = c.x + 1 < g.nx ? vxp : -me;
    const float pxm = c.x > 0 ? vxm : -me;
    const float pyp = c.y + 1 < g.ny ? vyp : -me;
    const float pym = c.y > 0 ? vym : me;
    const float pzp = c.z + 1 < g.nz ? vzp : -me;
    const float pzm = c.z > 0 ? vzm : -me;
    const float gx = (pxp - pxm) / (2.f * dx), gy = (pyp - pym) / (2.f * dx), gz = (pzp - pzm) / (2.f * dx);
    float vx = vi.x - gx, vy = vi.y - gy, vz = vi.z - gz;
    if (solid > 0.5f) { vx = 0.f; vy = 0.f; vz = 0.f; }
    if (c.y == 0) vy = fmaxf(vy, 0.f);
    const float vl = sqrtf(vx * vx + vy * vy + vz * vz);
    if (vl > 20.f || vl != vl) {
        if (vl != vl) { vx = 0.f; vy = 0.f; vz = 0.f; }
        else { vx = vx * 20.f / vl; vy = vy * 20.f / vl; vz = vz * 20.f / vl; }
    }
    Codec<S>::st(vOut, c.i, make_float4(vx, vy, vz, 0.f));
}

// ---- init-time noise and curl textures (src/noise.js:8-54, 69-80) ---------------
__device__ __forceinline__ void pcg3d(unsigned& x, unsigned& y, unsigned& z) {   // src/shaders/lib.js:4-10
    x = x * 1664525u + 1013904223u; y = y * 1664525u + 1013904223u; z = z * 1664525u + 1013904223u;
    x += y * z; y += z * x; z += x * y;
    x ^= x >> 16; y ^= y >> 16; z ^= z >> 16;
    x += y * z; y += z * x; z += x * y;
}
__device__ __forceinline__ float3 hash3f(int px, int py, int pz) {   // lib.js:16
    unsigned x = (unsigned)px, y = (unsigned)py, z = (unsigned)pz;
    pcg3d(x, y, z);
    const float k = (float)(1.0 / 4294967295.0);
    return make_float3((float)x * k, (float)y * k, (float)z * k);
}
__device__ __forceinline__ int wrapi(int c, int per) { return ((c % per) + per) % per; }
__device__ float tnoise(float px, float py, float pz, int per, int ch, int seed) {
    const float flx = floorf(px), fly = floorf(py), flz = floorf(pz);
    const int ix = (int)flx, iy = (int)fly, iz = (int)flz;
    const float fx = px - flx, fy = py - fly, fz = pz - flz;
    const float ux = fx * fx * fx * (fx * (fx * 6.f - 15.f) + 10.f);
    const float uy = fy * fy * fy * (fy * (fy * 6.f - 15.f) + 10.f);
    const float uz = fz * fz * fz * (fz * (fz * 6.f - 15.f) + 10.f);
    float acc[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
        const float3 h = hash3f(wrapi(ix + ox, per) + ch * 1013 + seed, wrapi(iy + oy, per) + ch * 7919,
                                wrapi(iz + oz, per) + ch * 104729);
        const float a = h.x * 2.f - 1.f + 1e-5f, b = h.y * 2.f - 1.f + 1e-5f, c = h.z * 2.f - 1.f + 1e-5f;
        const float l = sqrtf(a * a + b * b + c * c);
        acc[k] = (a / l) * (fx - (float)ox) + (b / l) * (fy - (float)oy) + (c / l) * (fz - (float)oz);
    }
    const float x0 = mixf(acc[0], acc[1], ux), x1 = mixf(acc[2], acc[3], ux);
    const float x2 = mixf(acc[4], acc[5], ux), x3 = mixf(acc[6], acc[7], ux);
    return mixf(mixf(x0, x1, uy), mixf(x2, x3, uy), uz);
}
__device__ float tfbm(float px, float py, float pz, int ch, int seed) {
    float s = 0.f, a = 0.5f;
    int per = 4;
    for (int o = 0; o < 5; o++) {
        const float fp = (float)per;
        s += a * tnoise(px * fp, py * fp, pz * fp, per, ch, seed);
        per *= 2; a *= 0.5f;
    }
    return s * 1.5f;
}
__device__ float worley(float px, float py, float pz, int per, int seed) {
    const float fp = (float)per;
    const float qx = px * fp, qy = py * fp, qz = pz * fp;
    const float flx = floorf(qx), fly = floorf(qy), flz = floorf(qz);
    const int ix = (int)flx, iy = (int)fly, iz = (int)flz;
    const float fx = qx - flx, fy = qy - fly, fz = qz - flz;
    float d = 9.f;
    for (int k = 0; k < 27; k++) {
        const int ox = k % 3 - 1, oy = (k / 3) % 3 - 1, oz = k / 9 - 1;
        const float3 h = hash3f(wrapi(ix + ox, per) + 77, wrapi(iy + oy, per) + 33, wrapi(iz + oz, per) + seed);
        const float a = (float)ox + h.x - fx, b = (float)oy + h.y - fy, c = (float)oz + h.z - fz;
        d = fminf(d, sqrtf(a * a + b * b + c * c));
    }
    return d;
}
__global__ __launch_bounds__(256) void k_noise(half4_t* __restrict__ out, int n, int seed) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n * n * n) return;
    const int x = i % n, y = (i / n) % n, z = i / (n * n);
    const float px = ((float)x + 0.5f) / (float)n, py = ((float)y + 0.5f) / (float)n, pz = ((float)z + 0.5f) / (float)n;
    const float a = tfbm(px, py, pz, 0, seed), b = tfbm(px, py, pz, 1, seed), c = tfbm(px, py, pz, 2, seed);
    const float w = 0.65f * worley(px, py, pz, 8, seed) + 0.35f * worley(px, py, pz, 16, seed);
    half4_t h;
    h.x = (_Float16)(a * 0.5f + 0.5f); h.y = (_Float16)(b * 0.5f + 0.5f);
    h.z = (_Float16)(c * 0.5f + 0.5f); h.w = (_Float16)wclamp(w, 0.f, 1.f);
    out[i] = h;
}
__global__ __launch_bounds__(256) void k_curl_tex(const half4_t* __restrict__ src, half4_t* __restrict__ out, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n * n * n) return;
    const int x = i % n, y = (i / n) % n, z = i / (n * n);
    auto L = [&](int xx, int yy, int zz) {
        xx = (xx + n) % n; yy = (yy + n) % n; zz = (zz + n) % n;
        const float4 v = ld_h4(src, (zz * n + yy) * n + xx);
        return make_float3(v.x * 2.f - 1.f, v.y * 2.f - 1.f, v.z * 2.f - 1.f);
    };
    const float3 ap = L(x + 1, y, z), am = L(x - 1, y, z), bp = L(x, y + 1, z), bm = L(x, y - 1, z), cp = L(x, y, z + 1), cm = L(x, y, z - 1);
    const float3 ddx = make_float3(ap.x - am.x, ap.y - am.y, ap.z - am.z);
    const float3 ddy = make_float3(bp.x - bm.x, bp.y - bm.y, bp.z - bm.z);
    const float3 ddz = make_float3(cp.x - cm.x, cp.y - cm.y, cp.z - cm.z);
    const float k = (float)n / 2.0f;
    const float cx = (ddy.z - ddz.y) * k, cy = (ddz.x - ddx.z) * k, cz = (ddx.y - ddy.x) * k;
    half4_t h;
    h.x = (_Float16)cx; h.y = (_Float16)cy; h.z = (_Float16)cz; h.w = (_Float16)sqrtf(cx * cx + cy * cy + cz * cz);
    out[i] = h;
}

// ---- small utility kernels ------------------------------------------------------
__global__ void k_set_params(DevParams* __restrict__ dst, DevParams src) { *dst = src; }
__global__ __launch_bounds__(256) void k_fill_zero(uint4* __restrict__ p, size_t n16, unsigned char* __restrict__ tail, size_t ntail) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n16) p[i] = make_uint4(0u, 0u, 0u, 0u);
    if (i < ntail) tail[i] = 0;
}
// kind 0: double -> float, 1: float -> double
__global__ __launch_bounds__(256) void k_convert(int kind, const void* __restrict__ src, void* __restrict__ dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (kind == 0) ((float*)dst)[i] = (float)((const double*)src)[i];
    else ((double*)dst)[i] = (double)((const float*)src)[i];
}

// plain streaming copy (bandwidth probe and counter calibration)
__global__ __launch_bounds__(256) void k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}
void launch_copy_probe(hipStream_t s, const void* src, void* dst, size_t bytes) {
    const size_t n = bytes / 16;
    k_copy_probe<<<2048, 256, 0, s>>>((const float4*)src, (float4*)dst, n);
}

// ---- host launchers ---------------------------------------------------------------
#define DISPATCH_PREC(prec, BODY)                                  \
    do {                                                           \
        if ((prec) == FS_PREC_FP32) { typedef StF32 S; BODY; }     \
        else if ((prec) == FS_PREC_FP16) { typedef StF16 S; BODY; } \
        else { typedef StBF16 S; BODY; }                           \
    } while (0)
#define RAW(p) (typename S::raw*)(p)
#define CRAW(p) (const typename S::raw*)(p)

template <class S> static void geo_t(const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F) {
    k_geo<S><<<g.nblocks, 256, 0, s>>>(P, g, F.logs, RAW(F.geo));
}
void launch_geo(int prec, const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F) {
    DISPATCH_PREC(prec, geo_t<S>(g, s, P, F));
}
template <class S> static void adv1_t(const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F) {
    k_adv1<S><<<g.nblocks, 256, 0, s>>>(P, g, CRAW(F.vel), CRAW(F.S), CRAW(F.S2), RAW(F.velX), RAW(F.SX), RAW(F.S2X));
}
void launch_adv1(int prec, const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F) {
    DISPATCH_PREC(prec, adv1_t<S>(g, s, P, F));
}
template <class S> static void adv2_t(const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F) {
    k_adv2<S><<<g.nblocks, 256, 0, s>>>(P, g, CRAW(F.vel), CRAW(F.S), CRAW(F.S2), CRAW(F.velX), CRAW(F.SX), CRAW(F.S2X),
                                        RAW(F.velY), RAW(F.SY), RAW(F.S2Y));
}
void launch_adv2(int prec, const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F) {
    DISPATCH_PREC(prec, adv2_t<S>(g, s, P, F));
}
template <class S> static void curl_t(const Dims& g, hipStream_t s, const FieldSet& F) {
    k_curl<S><<<g.nblocks, 256, 0, s>>>(g, CRAW(F.velY), RAW(F.vort));
}
void launch_curl(int prec, const Dims& g, hipStream_t s, const FieldSet& F) { DISPATCH_PREC(prec, curl_t<S>(g, s, F)); }
template <class S> static void forces_t(const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F) {
    k_forces<S><<<g.nblocks, 256, 0, s>>>(P, g, CRAW(F.velY), CRAW(F.SY), CRAW(F.S2Y), CRAW(F.vort), CRAW(F.geo), F.noise,
                                          F.curl, F.noise_n, F.spots, F.face_flux, RAW(F.velX), RAW(F.S), RAW(F.S2));
}
void launch_forces(int prec, const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F) {
    DISPATCH_PREC(prec, forces_t<S>(g, s, P, F));
}
template <class S> static void divk_t(const Dims& g, hipStream_t s, DevParams* P, const FieldSet& F, float* b, int adv) {
    k_div<S><<<g.nblocks, 256, 0, s>>>(P, g, CRAW(F.velX), b, adv);
}
void launch_div(int prec, const Dims& g, hipStream_t s, DevParams* P, const FieldSet& F, float* b_out, int advance_clock) {
    DISPATCH_PREC(prec, divk_t<S>(g, s, P, F, b_out, advance_clock));
}
template <class S> static void project_t(const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F, const float* p32,
                                         const double* p64) {
    if (p64) k_project<S, double><<<g.nblocks, 256, 0, s>>>(P, g, CRAW(F.velX), p64, CRAW(F.geo), RAW(F.vel));
    else k_project<S, float><<<g.nblocks, 256, 0, s>>>(P, g, CRAW(F.velX), p32, CRAW(F.geo), RAW(F.vel));
}
void launch_project(int prec, const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F, const float* p32,
                    const double* p64) {
    DISPATCH_PREC(prec, project_t<S>(g, s, P, F, p32, p64));
}
void launch_noise_bake(hipStream_t s, half4_t* noise, half4_t* curl, int n, int seed) {
    const int cells = n * n * n;
    k_noise<<<(cells + 255) / 256, 256, 0, s>>>(noise, n, seed);
    k_curl_tex<<<(cells + 255) / 256, 256, 0, s>>>(noise, curl, n);
}
void launch_set_params(hipStream_t s, DevParams* dst, const DevParams& src) { k_set_params<<<1, 1, 0, s>>>(dst, src); }
void launch_fill_zero(hipStream_t s, void* ptr, size_t bytes) {
    const size_t n16 = bytes / 16, ntail = bytes % 16;
    const size_t n = n16 > ntail ? n16 : ntail;
    if (n == 0) return;
    k_fill_zero<<<(unsigned)((n + 255) / 256), 256, 0, s>>>((uint4*)ptr, n16, (unsigned char*)ptr + n16 * 16, ntail);
}
void launch_convert(hipStream_t s, int kind, const void* src, void* dst, size_t n) {
    if (n == 0) return;
    k_convert<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(kind, src, dst, n);
}
