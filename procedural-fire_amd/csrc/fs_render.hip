// Render-volume consumers of the step's fields on gfx950: the fire-light extraction
// (src/renderer.js:104-164), the flame-field prefilter (src/renderer.js:173-191) and the
// volumetric ray-march (src/renderer.js:194-326) that the reference runs each frame after the
// simulation step.
//
// Lights: one 256-thread block per height band, a strided pass over the 2x2x2-subsampled
// band then the halving tree of the reference's workgroup reduction (same summation order).
// Prefilter: one thread per cell, s + k (s - mean of the 6 neighbours) clamped to the local
// range, stored as binary16 like the reference's rgba16float target.
// Ray-march: one thread per volume pixel; a 256-thread block covers 16x16 pixels with each
// wave on an 8x8 quad, so neighbouring rays walk neighbouring cells and share cache lines,
// and blocks are dealt to the XCDs in contiguous bands of tiles (each XCD's L2 then holds
// the part of the field its rays cross).  The march, its sub-voxel curl warp, soot and
// blue emission, single scattering of the 8 lights and the Beer-Lambert accumulation follow
// the reference's order of operations.  Samples use the step's trilinear (clamp to edge) at
// cell-space positions and the repeat-addressed curl sample; exp/pow are the library's
// textbook ones, normalize is v / sqrt(v.v), inverseSqrt is 1 / sqrt.
#include <hip/hip_runtime.h>

#include "fs_codec.h"
#include "fs_internal.h"
#include "fs_math.h"

namespace {

__device__ __forceinline__ float fractf(float x) { return x - floorf(x); }
struct F3 { float x, y, z; };
__device__ __forceinline__ F3 v3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 vadd(F3 a, F3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 vsub(F3 a, F3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 vmul(F3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float vdot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 vnorm(F3 a) {
    const float l = sqrtf(vdot(a, a));
    return v3(a.x / l, a.y / l, a.z / l);
}
__device__ __forceinline__ float lum(float r, float g, float b) { return r * 0.2126f + g * 0.7152f + b * 0.0722f; }

// bbRow (src/renderer.js:74-80): linear interpolation along one LUT row
__device__ __forceinline__ F3 bb_row(const RenderU& R, const float4* lut, float T, int row) {
    const float x = wclamp((T - R.f[RP_bbTmin]) / (R.f[RP_bbTmax] - R.f[RP_bbTmin]), 0.0f, 1.0f) * (R.f[RP_bbN] - 1.0f);
    const unsigned i = (unsigned)floorf(x);
    const float f = x - floorf(x);
    const unsigned n = (unsigned)R.f[RP_bbN];
    const float4 a = lut[row * n + i], b = lut[row * n + min(i + 1u, n - 1u)];
    return v3(mixf(a.x, b.x, f), mixf(a.y, b.y, f), mixf(a.z, b.z, f));
}

// ---- light extraction (src/renderer.js:112-164) -----------------------------------
template <class S>
__global__ __launch_bounds__(256) void k_lights(RenderU R, Dims g, const typename S::raw* __restrict__ sT,
                                                const float4* __restrict__ lut, float4* __restrict__ lstate,
                                                float4* __restrict__ lights) {
    __shared__ float4 accP[256], accC[256];
    const unsigned li = threadIdx.x, wg = blockIdx.x;
    const unsigned nx = (unsigned)R.f[RP_nx], ny = (unsigned)R.f[RP_ny], nz = (unsigned)R.f[RP_nz];
    const unsigned bandH = ny / (unsigned)FS_NLIGHTS;
    const unsigned y0 = wg * bandH;
    const unsigned hx = nx / 2u, hb = bandH / 2u;
    const unsigned total = hx * hb * (nz / 2u);
    float4 P = make_float4(0.f, 0.f, 0.f, 0.f), C = make_float4(0.f, 0.f, 0.f, 0.f);
    for (unsigned i = li; i < total; i += 256u) {
        const unsigned ix = (i % hx) * 2u;
        const unsigned iy = ((i / hx) % hb) * 2u + y0;
        const unsigned iz = (i / (hx * hb)) * 2u;
        const float4 s = Codec<S>::ld(sT, (int)((iz * (unsigned)g.ny + iy) * (unsigned)g.nx + ix));
        const float T = 293.0f + s.x * R.f[RP_tempScale];
        if (T < 700.0f) continue;
        const F3 bs = bb_row(R, lut, T, 1);
        const float k = s.z * R.f[RP_sootEmit];
        const F3 e = v3(k * bs.x * R.f[RP_emitK], k * bs.y * R.f[RP_emitK], k * bs.z * R.f[RP_emitK]);
        const float w = lum(e.x, e.y, e.z);
        const F3 pos = v3(R.f[RP_ox] + ((float)ix + 1.0f) * R.f[RP_dx], R.f[RP_oy] + ((float)iy + 1.0f) * R.f[RP_dx],
                          R.f[RP_oz] + ((float)iz + 1.0f) * R.f[RP_dx]);
        P.x += pos.x * w; P.y += pos.y * w; P.z += pos.z * w; P.w += w;
        C.x += e.x; C.y += e.y; C.z += e.z; C.w += vdot(pos, pos) * w;
    }
    accP[li] = P;
    accC[li] = C;
    __syncthreads();
    for (unsigned s = 128u; s > 0u; s >>= 1u) {
        if (li < s) {
            const float4 a = accP[li + s], c = accC[li + s];
            accP[li].x += a.x; accP[li].y += a.y; accP[li].z += a.z; accP[li].w += a.w;
            accC[li].x += c.x; accC[li].y += c.y; accC[li].z += c.z; accC[li].w += c.w;
        }
        __syncthreads();
    }
    if (li != 0u) return;
    const float4 AP = accP[0], AC = accC[0];
    const float w = AP.w;
    const bool some = w > 1e-6f;
    const F3 cen = some ? v3(AP.x / w, AP.y / w, AP.z / w) : v3(0.0f, ((float)y0 + (float)bandH * 0.5f) * R.f[RP_dx], 0.0f);
    const float spread = some ? fmaxf(AC.w / w - vdot(cen, cen), 0.0f) : 0.04f;
    const float sc_ = (float)(nx * bandH * nz) / (float)total;
    const float dx = R.f[RP_dx];
    const F3 J = v3(AC.x * sc_ * dx * dx * dx, AC.y * sc_ * dx * dx * dx, AC.z * sc_ * dx * dx * dx);
    const int N = FS_NLIGHTS, k = (int)wg;
    float4 sp = lstate[k], sc = lstate[N + k], sm = lstate[2 * N + k];
    const float lightDt = R.f[RP_lightDt];
    if (lightDt < 0.0f || sm.w <= 0.0f) {
        sp = make_float4(cen.x, cen.y, cen.z, spread);
        sc = make_float4(J.x, J.y, J.z, 0.0f);
        sm = make_float4(J.x, J.y, J.z, 1.0f);
    } else {
        const float a = 1.0f - fs_exp(-lightDt / 0.08f), ap = 1.0f - fs_exp(-lightDt / 0.15f),
                    am = 1.0f - fs_exp(-lightDt / 1.5f);
        sp = make_float4(mixf(sp.x, cen.x, ap), mixf(sp.y, cen.y, ap), mixf(sp.z, cen.z, ap), mixf(sp.w, spread, ap));
        sc = make_float4(mixf(sc.x, J.x, a), mixf(sc.y, J.y, a), mixf(sc.z, J.z, a), 0.0f);
        sm = make_float4(mixf(sm.x, J.x, am), mixf(sm.y, J.y, am), mixf(sm.z, J.z, am), 1.0f);
    }
    lstate[k] = sp;
    lstate[N + k] = sc;
    lstate[2 * N + k] = sm;
    const float ratio = lum(sc.x, sc.y, sc.z) / fmaxf(lum(sm.x, sm.y, sm.z), 1e-9f);
    const float fk = fs_pow(fmaxf(ratio, 1e-4f), R.f[RP_flickerK] - 1.0f);
    const F3 Jc = v3(sc.x * fk, sc.y * fk, sc.z * fk);
    const float wl = lum(Jc.x, Jc.y, Jc.z);
    const bool on = wl > 1e-3f * lum(sm.x, sm.y, sm.z) && wl > 0.0f;
    lights[2 * k] = make_float4(sp.x, sp.y, sp.z, on ? wl : 0.0f);
    lights[2 * k + 1] = on ? make_float4(Jc.x, Jc.y, Jc.z, sp.w) : make_float4(0.0f, 0.0f, 0.0f, sp.w);
}

// ---- prefilter (src/renderer.js:176-191) ----------------------------------------------
template <class S>
__global__ __launch_bounds__(256) void k_sharpen(float k, Dims g, const typename S::raw* __restrict__ sIn,
                                                 half4_t* __restrict__ sOut) {
    Cell c;
    if (!cell_of(g, c)) return;
    const float4 s = Codec<S>::ld(sIn, c.i);
    float4 n = make_float4(0.f, 0.f, 0.f, 0.f), lo = s, up = s;
#pragma unroll
    for (int q = 0; q < 6; q++) {   // +x, -x, +y, -y, +z, -z
        const int d = (q & 1) ? -1 : 1;
        const float4 v = ldc<S>(sIn, g, c.x + (q / 2 == 0 ? d : 0), c.y + (q / 2 == 1 ? d : 0), c.z + (q / 2 == 2 ? d : 0));
        n.x += v.x; n.y += v.y; n.z += v.z; n.w += v.w;
        lo.x = fminf(lo.x, v.x); lo.y = fminf(lo.y, v.y); lo.z = fminf(lo.z, v.z); lo.w = fminf(lo.w, v.w);
        up.x = fmaxf(up.x, v.x); up.y = fmaxf(up.y, v.y); up.z = fmaxf(up.z, v.z); up.w = fmaxf(up.w, v.w);
    }
    const float4 o = make_float4(wclamp(s.x + k * (s.x - n.x / 6.0f), lo.x, up.x), wclamp(s.y + k * (s.y - n.y / 6.0f), lo.y, up.y),
                                 wclamp(s.z + k * (s.z - n.z / 6.0f), lo.z, up.z), wclamp(s.w + k * (s.w - n.w / 6.0f), lo.w, up.w));
    Codec<StF16>::st(sOut, c.i, o);
}

// ---- empty-space map of the march ---------------------------------------------------------
// A march sample changes nothing (L, Tr and haze stay bit for bit) when its temperature,
// soot and heat release after the fade, and its smoke, are at or below the thresholds of
// src/renderer.js:292 and :320.  The cubic and trilinear weights are non-negative and sum to
// 1 and the fade is at most 1, so a sample whose texel footprint holds no texel above
// (1 - 1e-5) x threshold is such a no-op (the margin covers the rounding of the weighted sum).
// Bricks of 4^3 cells are flagged when any cell within `reach` cells (the cubic footprint, the
// largest curl-warp displacement and one cell of slack) exceeds that bound.
constexpr float kActiveMargin = 0.99999f;

template <class SS, class S2S>
__global__ __launch_bounds__(256) void k_active(Dims g, const typename SS::raw* __restrict__ sT,
                                                const typename S2S::raw* __restrict__ s2T, int smoke_on,
                                                unsigned char* __restrict__ raw, int bx, int by) {
    Cell c;
    if (!cell_of(g, c)) return;
    const float4 s = Codec<SS>::ld(sT, c.i);
    bool on = s.x > 0.05f * kActiveMargin || s.z > 1e-4f * kActiveMargin || s.w > 1e-3f * kActiveMargin;
    if (smoke_on) on = on || Codec<S2S>::ld(s2T, c.i).z > 1e-4f * kActiveMargin;
    if (on) raw[((c.z >> 2) * by + (c.y >> 2)) * bx + (c.x >> 2)] = 1;   // every writer stores 1
}
__global__ __launch_bounds__(256) void k_dilate(const unsigned char* __restrict__ raw, unsigned char* __restrict__ flags,
                                                int bx, int by, int bz, int rb) {
    const int i = (int)(blockIdx.x * 256 + threadIdx.x);
    if (i >= bx * by * bz) return;
    const int x = i % bx, y = (i / bx) % by, z = i / (bx * by);
    unsigned char f = 0;
    for (int dz = max(z - rb, 0); dz <= min(z + rb, bz - 1); dz++)
        for (int dy = max(y - rb, 0); dy <= min(y + rb, by - 1); dy++)
            for (int dx = max(x - rb, 0); dx <= min(x + rb, bx - 1); dx++) f |= raw[(dz * by + dy) * bx + dx];
    flags[i] = f;
}

// ---- volume ray-march (src/renderer.js:207-326) -----------------------------------------
struct VolIn {
    const float4* lut;     // 2 rows x bbN texels: blackbody, soot
    const float4* lights;  // 8 x (pos, col)
    const float* depth;    // surface depth, dw x dh
    int dw, dh;
    const half4_t* curl;
    int nn;
    const unsigned char* flags;   // empty-space map (bricks of 4^3 cells), or null
    int bx, by, bz;
    float inv_dx;                 // for the map lookup only (the march itself divides by dx)
};

__device__ __forceinline__ float hg(float c, float g) {   // Henyey-Greenstein phase function
    const float g2 = g * g;
    return (1.0f - g2) / (4.0f * 3.14159265f * fs_pow(1.0f + g2 - 2.0f * g * c, 1.5f));
}
// cubic B-spline reconstruction from 8 trilinear fetches (Sigg & Hadwiger 2005); g is the
// cell-space position (texel centres at i + 0.5), and so are the fetch positions h0, h1
template <class SS, class F>
__device__ __forceinline__ float4 sample_cubic(const F& sT, const Dims& d, F3 g) {
    const float cs[3] = {g.x - 0.5f, g.y - 0.5f, g.z - 0.5f};
    float h0[3], h1[3], g0[3], g1[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float i = floorf(cs[a]), f = cs[a] - i;
        const float f2 = f * f, f3 = f2 * f;
        const float w0 = (1.0f - 3.0f * f + 3.0f * f2 - f3) / 6.0f;
        const float w1 = (4.0f - 6.0f * f2 + 3.0f * f3) / 6.0f;
        const float w2 = (1.0f + 3.0f * f + 3.0f * f2 - 3.0f * f3) / 6.0f;
        const float w3 = f3 / 6.0f;
        g0[a] = w0 + w1;
        g1[a] = w2 + w3;
        h0[a] = i - 0.5f + w1 / g0[a];
        h1[a] = i + 1.5f + w3 / g1[a];
    }
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const bool ox = k & 1, oy = k & 2, oz = k & 4;
        const float4 s = sample_clamp_f<SS>(sT, d, ox ? h1[0] : h0[0], oy ? h1[1] : h0[1], oz ? h1[2] : h0[2]);
        const float wv = (ox ? g1[0] : g0[0]) * (oy ? g1[1] : g0[1]) * (oz ? g1[2] : g0[2]);
        acc.x += s.x * wv; acc.y += s.y * wv; acc.z += s.z * wv; acc.w += s.w * wv;
    }
    return acc;
}

template <class SS, class S2S>
__global__ __launch_bounds__(256) void k_volume(RenderU R, Dims d, const typename SS::raw* __restrict__ sT,
                                                const typename S2S::raw* __restrict__ s2T, VolIn in,
                                                half4_t* __restrict__ volOut, float2* __restrict__ auxOut) {
    const float vw = R.f[RP_vw], vh = R.f[RP_vh];
    const int tilesX = ((int)vw + 15) / 16, tilesY = ((int)vh + 15) / 16;
    const int tile = xcd_remap((int)blockIdx.x, tilesX * tilesY);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const FPair<SS> ST(sT, d);        // x-neighbour texel pairs through buffer resources for 8-byte texels
    const FPair<S2S> S2T(s2T, d);
    const RepTex CT(in.curl, in.nn);
    const int ix = (tile % tilesX) * 16 + (wv & 1) * 8 + (lane & 7);
    const int iy = (tile / tilesX) * 16 + (wv >> 1) * 8 + (lane >> 3);
    if ((float)ix >= vw || (float)iy >= vh) return;
    const float pxx = (float)ix + 0.5f + R.f[RP_jitX] * (vw / R.f[RP_width]);
    const float pxy = (float)iy + 0.5f + R.f[RP_jitY] * (vw / R.f[RP_width]);
    // camRay + lensRay (src/renderer.js:87-94, 207-211)
    const float ndx = pxx / vw * 2.0f - 1.0f, ndy = 1.0f - pxy / vh * 2.0f;
    const F3 fwd = v3(R.f[RP_fwdX], R.f[RP_fwdY], R.f[RP_fwdZ]);
    const F3 rgt = v3(R.f[RP_rgtX], R.f[RP_rgtY], R.f[RP_rgtZ]);
    const F3 upv = v3(R.f[RP_upX], R.f[RP_upY], R.f[RP_upZ]);
    const F3 dPin = vnorm(vadd(vadd(fwd, vmul(vmul(vmul(rgt, ndx), R.f[RP_tanFov]), R.f[RP_aspect])),
                               vmul(vmul(upv, ndy), R.f[RP_tanFov])));
    const F3 cam = v3(R.f[RP_camX], R.f[RP_camY], R.f[RP_camZ]);
    F3 ro = cam, rd = dPin;
    if (R.f[RP_focus] > 0.0f) {
        ro = vadd(vadd(cam, vmul(rgt, R.f[RP_lensX])), vmul(upv, R.f[RP_lensY]));
        const F3 q = vadd(cam, vmul(dPin, R.f[RP_focus] / vdot(dPin, fwd)));
        rd = vnorm(vsub(q, ro));
    }
    // surface depth: centre sample of the full-resolution buffer
    const int fx = (int)(pxx / vw * (float)in.dw), fy = (int)(pxy / vh * (float)in.dh);
    const float sd = in.depth[clampi(fy, 0, in.dh - 1) * in.dw + clampi(fx, 0, in.dw - 1)];
    const F3 bmin = v3(R.f[RP_ox], R.f[RP_oy], R.f[RP_oz]);
    const F3 bmax = vadd(bmin, v3(R.f[RP_sx], R.f[RP_sy], R.f[RP_sz]));
    const F3 inv = v3(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    const F3 t0 = v3((bmin.x - ro.x) * inv.x, (bmin.y - ro.y) * inv.y, (bmin.z - ro.z) * inv.z);
    const F3 t1 = v3((bmax.x - ro.x) * inv.x, (bmax.y - ro.y) * inv.y, (bmax.z - ro.z) * inv.z);
    const float tn = fmaxf(fmaxf(fmaxf(fminf(t0.x, t1.x), fminf(t0.y, t1.y)), fminf(t0.z, t1.z)), 0.0f);
    const float tf = fminf(fminf(fminf(fmaxf(t0.x, t1.x), fmaxf(t0.y, t1.y)), fmaxf(t0.z, t1.z)), sd);
    float Lr = 0.f, Lg = 0.f, Lb = 0.f, Tr = 1.0f, haze = 0.0f;
    if (tf > tn) {
        const float dx = R.f[RP_dx];
        const float ds = dx * R.f[RP_stepMul];
        // interleaved gradient noise (src/shaders/lib.js:22-27)
        const float fr = R.f[RP_frame] - 64.0f * truncf(R.f[RP_frame] / 64.0f);
        const float qx = (float)ix + 5.588238f * fr, qy = (float)iy + 5.588238f * fr;
        const float jit = fractf(52.9829189f * fractf(qx * 0.06711056f + qy * 0.00583715f));
        float t = tn + ds * jit;
        const float nxf = R.f[RP_nx], nyf = R.f[RP_ny], nzf = R.f[RP_nz];
        const int nlit = (int)R.f[RP_nLights];
        const float detail = R.f[RP_detailAmt];
        for (int i = 0; i < 1200; i++) {
            if (t > tf || Tr < 0.003f) break;
            const F3 p = vadd(ro, vmul(rd, t));
            if (in.flags) {   // a sample in an empty brick is a no-op: advance only
                const int cx = clampi((int)floorf((p.x - bmin.x) * in.inv_dx), 0, in.bx * 4 - 1);
                const int cy = clampi((int)floorf((p.y - bmin.y) * in.inv_dx), 0, in.by * 4 - 1);
                const int cz = clampi((int)floorf((p.z - bmin.z) * in.inv_dx), 0, in.bz * 4 - 1);
                if (!in.flags[((cz >> 2) * in.by + (cy >> 2)) * in.bx + (cx >> 2)]) {
                    t += ds;
                    continue;
                }
            }
            F3 g = v3((p.x - bmin.x) / dx, (p.y - bmin.y) / dx, (p.z - bmin.z) / dx);
            float mottle = 0.0f;
            if (detail > 0.0f) {   // divergence-free warp of the lookup in two cross-faded phases
                const float rise = R.f[RP_detailSpeed] * (0.5f + 1.5f * sstep(0.0f, 1.0f, p.y));
                const float ph = R.f[RP_time] / 0.7f;
                const float f0 = fractf(ph), f1 = fractf(ph + 0.5f);
                const float w0 = 1.0f - fabsf(2.0f * f0 - 1.0f);
                const float fl0 = floorf(ph), fl1 = floorf(ph + 0.5f);
                const float fq = R.f[RP_detailFreq];
                const float q0x = (p.x - 0.0f) * fq * 0.25f + 0.37f * fl0;
                const float q0y = (p.y - rise * f0 * 0.7f) * fq * 0.25f + 0.0f * fl0;
                const float q0z = (p.z - 0.0f) * fq * 0.25f + 0.61f * fl0;
                const float q1x = (p.x - 0.0f) * fq * 0.25f + 0.71f * fl1;
                const float q1y = (p.y - rise * f1 * 0.7f) * fq * 0.25f + 0.0f * fl1;
                const float q1z = (p.z - 0.0f) * fq * 0.25f + 0.29f * fl1;
                const float4 c0 = sample_rep_b(CT, q0x, q0y, q0z);
                const float4 c1 = sample_rep_b(CT, q1x, q1y, q1z);
                const float cwx = c0.x * w0 + c1.x * (1.0f - w0), cwy = c0.y * w0 + c1.y * (1.0f - w0);
                const float cwz = c0.z * w0 + c1.z * (1.0f - w0), cww = c0.w * w0 + c1.w * (1.0f - w0);
                const float k12 = detail * 0.12f;
                g = v3(g.x + cwx * k12 * 1.0f, g.y + cwy * k12 * 0.6f, g.z + cwz * k12 * 1.0f);
                mottle = cww * 0.08f - 0.6f;
            }
            float4 s = R.f[RP_cubic] > 0.5f ? sample_cubic<SS>(ST, d, g) : sample_clamp_f<SS>(ST, d, g.x, g.y, g.z);
            // fade towards the open sides and top
            const float edge = fminf(fminf(g.x, nxf - g.x), fminf(g.z, nzf - g.z));
            const float fade = sstep(0.02f, 0.15f, edge * dx) * sstep(0.0f, 0.4f, (nyf - g.y) * dx);
            s = make_float4(s.x * fade, s.y * fade, s.z * fade, s.w * fade);
            const float T = 293.0f + s.x * R.f[RP_tempScale] + mottle * 60.0f * detail * (s.x >= 0.3f ? 1.0f : 0.0f);
            float soot = s.z;
            if (R.f[RP_sootEdge] > 0.0f) soot = soot * sstep(0.6f * R.f[RP_sootEdge], 1.4f * R.f[RP_sootEdge], soot);
            float smoke = 0.0f;
            if (R.f[RP_smokeScat] > 0.0f) smoke = sample_clamp_f<S2S>(S2T, d, g.x, g.y, g.z).z * fade;
            if (soot > 1e-4f || s.w > 1e-3f || smoke > 1e-4f) {
                const float sigSoot = soot * R.f[RP_sootAbs];
                const float sigS = sigSoot * R.f[RP_sootScat] + smoke * R.f[RP_smokeScat];
                const float sigT = sigSoot + smoke * R.f[RP_smokeScat] * 1.1f + 1e-6f;
                float er = 0.f, eg = 0.f, eb = 0.f;
                if (T > R.f[RP_emitTempLo]) {
                    const F3 bs = bb_row(R, in.lut, T, 1);
                    const float k = soot * R.f[RP_sootEmit];
                    er = k * bs.x * R.f[RP_emitK]; eg = k * bs.y * R.f[RP_emitK]; eb = k * bs.z * R.f[RP_emitK];
                }
                const float pw = fs_pow(s.w, 1.6f), bk = R.f[RP_blueK] / 3.0f, ex = fs_exp(-soot * 2.0f);
                er += 0.15f * pw * bk * ex; eg += 0.3f * pw * bk * ex; eb += 1.0f * pw * bk * ex;
                float ir = 0.f, ig = 0.f, ib = 0.f;
                if (sigS * Tr > 1e-6f) {
                    for (int k = 0; k < nlit; k++) {
                        const float4 lp = in.lights[2 * k], lc = in.lights[2 * k + 1];
                        if (lp.w == 0.0f) continue;   // a dark light (colour 0) adds exactly +0
                        const F3 toL = vsub(v3(lp.x, lp.y, lp.z), p);
                        const float r2 = vdot(toL, toL) + lc.w * 0.6f + 0.004f;
                        const float ph_ = hg(vdot(rd, vmul(toL, 1.0f / sqrtf(r2))), R.f[RP_smokeG]);
                        ir += lc.x * ph_ / r2; ig += lc.y * ph_ / r2; ib += lc.z * ph_ / r2;
                    }
                    const float ks = sigS * R.f[RP_lightK];
                    ir *= ks; ig *= ks; ib *= ks;
                }
                const float a = fs_exp(-sigT * ds);
                const float w = sigT * ds < 1e-4f ? ds : (1.0f - a) / sigT;
                Lr += Tr * (er + ir) * w; Lg += Tr * (eg + ig) * w; Lb += Tr * (eb + ib) * w;
                Tr *= a;
            }
            haze += fmaxf(s.x - 0.05f, 0.0f) * ds;
            t += ds;
        }
    }
    const int o = iy * (int)vw + ix;
    Codec<StF16>::st(volOut, o, make_float4(Lr, Lg, Lb, Tr));
    auxOut[o] = make_float2(haze, sd);
}

template <class SS, class S2S>
void volume_t(hipStream_t s, const RenderU& R, const Dims& d, const void* sT, const void* s2T, const VolIn& in,
              half4_t* vol, float2* aux) {
    const int tx = ((int)R.f[RP_vw] + 15) / 16, ty = ((int)R.f[RP_vh] + 15) / 16;
    k_volume<SS, S2S><<<tx * ty, 256, 0, s>>>(R, d, (const typename SS::raw*)sT, (const typename S2S::raw*)s2T, in, vol, aux);
}

}  // namespace

void launch_lights(int prec, hipStream_t s, const RenderU& R, const Dims& g, const void* S_field, const float4* lut,
                   float4* lstate, float4* lights) {
    switch (prec) {
        case FS_PREC_FP32: k_lights<StF32><<<FS_NLIGHTS, 256, 0, s>>>(R, g, (const float4*)S_field, lut, lstate, lights); break;
        case FS_PREC_FP16: k_lights<StF16><<<FS_NLIGHTS, 256, 0, s>>>(R, g, (const half4_t*)S_field, lut, lstate, lights); break;
        default: k_lights<StBF16><<<FS_NLIGHTS, 256, 0, s>>>(R, g, (const StBF16::raw*)S_field, lut, lstate, lights); break;
    }
}

void launch_sharpen(int prec, hipStream_t s, float k, const Dims& g, const void* S_in, half4_t* out) {
    switch (prec) {
        case FS_PREC_FP32: k_sharpen<StF32><<<g.nblocks, 256, 0, s>>>(k, g, (const float4*)S_in, out); break;
        case FS_PREC_FP16: k_sharpen<StF16><<<g.nblocks, 256, 0, s>>>(k, g, (const half4_t*)S_in, out); break;
        default: k_sharpen<StBF16><<<g.nblocks, 256, 0, s>>>(k, g, (const StBF16::raw*)S_in, out); break;
    }
}

void launch_occupancy(int prec, bool sharpened, hipStream_t s, const Dims& g, const void* S_field, const void* S2_field,
                      bool smoke_on, unsigned char* raw, unsigned char* flags, int bx, int by, int bz, int rb) {
    const int nb = bx * by * bz;
    launch_fill_zero(s, raw, (size_t)nb);
    const int so = smoke_on ? 1 : 0;
    if (sharpened) {
        switch (prec) {
            case FS_PREC_FP32: k_active<StF16, StF32><<<g.nblocks, 256, 0, s>>>(g, (const half4_t*)S_field, (const float4*)S2_field, so, raw, bx, by); break;
            case FS_PREC_FP16: k_active<StF16, StF16><<<g.nblocks, 256, 0, s>>>(g, (const half4_t*)S_field, (const half4_t*)S2_field, so, raw, bx, by); break;
            default: k_active<StF16, StBF16><<<g.nblocks, 256, 0, s>>>(g, (const half4_t*)S_field, (const StBF16::raw*)S2_field, so, raw, bx, by); break;
        }
    } else {
        switch (prec) {
            case FS_PREC_FP32: k_active<StF32, StF32><<<g.nblocks, 256, 0, s>>>(g, (const float4*)S_field, (const float4*)S2_field, so, raw, bx, by); break;
            case FS_PREC_FP16: k_active<StF16, StF16><<<g.nblocks, 256, 0, s>>>(g, (const half4_t*)S_field, (const half4_t*)S2_field, so, raw, bx, by); break;
            default: k_active<StBF16, StBF16><<<g.nblocks, 256, 0, s>>>(g, (const StBF16::raw*)S_field, (const StBF16::raw*)S2_field, so, raw, bx, by); break;
        }
    }
    k_dilate<<<(nb + 255) / 256, 256, 0, s>>>(raw, flags, bx, by, bz, rb);
}

void launch_volume(int prec, bool sharpened, hipStream_t s, const RenderU& R, const Dims& g, const void* S_field,
                   const void* S2_field, const float4* lut, const float4* lights, const float* depth, int dw, int dh,
                   const half4_t* curl, int nn, const unsigned char* flags, int bx, int by, int bz, float inv_dx,
                   half4_t* vol, float2* aux) {
    const VolIn in{lut, lights, depth, dw, dh, curl, nn, flags, bx, by, bz, inv_dx};
    if (sharpened) {
        switch (prec) {
            case FS_PREC_FP32: volume_t<StF16, StF32>(s, R, g, S_field, S2_field, in, vol, aux); break;
            case FS_PREC_FP16: volume_t<StF16, StF16>(s, R, g, S_field, S2_field, in, vol, aux); break;
            default: volume_t<StF16, StBF16>(s, R, g, S_field, S2_field, in, vol, aux); break;
        }
    } else {
        switch (prec) {
            case FS_PREC_FP32: volume_t<StF32, StF32>(s, R, g, S_field, S2_field, in, vol, aux); break;
            case FS_PREC_FP16: volume_t<StF16, StF16>(s, R, g, S_field, S2_field, in, vol, aux); break;
            default: volume_t<StBF16, StBF16>(s, R, g, S_field, S2_field, in, vol, aux); break;
        }
    }
}
