// TEST INFRASTRUCTURE ONLY — CPU oracle of the render-volume consumers of the step's fields:
// the fire-light extraction (src/renderer.js:104-164), the flame-field prefilter
// (src/renderer.js:173-191) and the volume ray-march (src/renderer.js:194-326), in f32 with the
// reference's order of operations.  Inputs are plain arrays: the uniform block (RP_FIELDS order,
// src/renderer.js:8-17, as uploaded), the S / S2 fields as decoded floats (x fastest), the
// blackbody LUT (2 rows x bbN x 4), the 8 lights, the surface depth and the curl texture.
// Only tests/ load it, as the checker of the HIP library.
//
// Conventions where WGSL leaves the arithmetic to the implementation (the same ones the step's
// oracle fixes): trilinear weights are exact f32 at cell-space positions (texel centres at
// i + 0.5), clamp/min/max return the non-NaN operand, exp/pow are the textbook constructions of
// oracle_common.h, normalize(v) = v / sqrt(v.v), inverseSqrt(x) = 1 / sqrt(x), stores to
// rgba16float round to nearest even.
#include <algorithm>
#include <cmath>
#include <vector>

#include <omp.h>

#include "oracle_common.h"

using namespace orc;

namespace {

enum {   // RP_FIELDS indices used here
    width = 0, vwI = 2, vhI = 3, timeI = 4, frameI = 5,
    camX = 8, camY, camZ, tanFov, fwdX, fwdY, fwdZ, aspect, rgtX, rgtY, rgtZ, jitX, upX, upY, upZ, jitY,
    ox = 24, oy, oz, dxI, sx, sy, sz, tempScale, nxI, nyI, nzI, nLights,
    stepMul = 37, sootAbs, sootEmit, emitK, blueK, smokeScat, smokeG,
    detailAmt = 45, detailFreq, detailSpeed,
    lightK = 58,
    bbTmin = 60, bbTmax, bbN, flameSharp = 65,
    emitTempLo = 70,
    sootScat = 78, lensX = 80, lensY = 81, focus = 82, lightDt = 83, flickerK = 84, cubic = 85, sootEdge = 87,
};

typedef V3<float> F3;
typedef V4<float> F4;

struct Grid {
    const float* f;
    int nx, ny, nz;
    F4 ld(int x, int y, int z) const {
        x = std::min(std::max(x, 0), nx - 1); y = std::min(std::max(y, 0), ny - 1); z = std::min(std::max(z, 0), nz - 1);
        const float* s = f + (((size_t)z * ny + y) * nx + x) * 4;
        return F4(s[0], s[1], s[2], s[3]);
    }
    // trilinear, clamp-to-edge, at cell-space position b
    F4 sample(float bx, float by, float bz) const {
        const float tx = wclamp<float>(bx - 0.5f, -1.f, (float)nx), ty = wclamp<float>(by - 0.5f, -1.f, (float)ny),
                    tz = wclamp<float>(bz - 0.5f, -1.f, (float)nz);
        float fx, fy, fz;
        const int ix = tri_axis<float>(tx, nx, fx), iy = tri_axis<float>(ty, ny, fy), iz = tri_axis<float>(tz, nz, fz);
        F4 c[8];
        for (int k = 0; k < 8; k++) c[k] = ld(ix + (k & 1), iy + ((k >> 1) & 1), iz + (k >> 2));
        float w[8];
        tri_weights<float>(fx, fy, fz, w);
        return tri_blend<float>(c, w);
    }
    // cubic B-spline from 8 trilinear fetches (Sigg & Hadwiger 2005), src/renderer.js:215-233
    F4 cubic(F3 g) const {
        const float c[3] = {g.x - 0.5f, g.y - 0.5f, g.z - 0.5f};
        float h0[3], h1[3], g0[3], g1[3];
        for (int a = 0; a < 3; a++) {
            const float i = std::floor(c[a]), f = c[a] - i;
            const float f2 = f * f, f3 = f2 * f;
            const float w0 = (1.0f - 3.0f * f + 3.0f * f2 - f3) / 6.0f;
            const float w1 = (4.0f - 6.0f * f2 + 3.0f * f3) / 6.0f;
            const float w2 = (1.0f + 3.0f * f + 3.0f * f2 - 3.0f * f3) / 6.0f;
            const float w3 = f3 / 6.0f;
            g0[a] = w0 + w1; g1[a] = w2 + w3;
            h0[a] = i - 0.5f + w1 / g0[a];
            h1[a] = i + 1.5f + w3 / g1[a];
        }
        F4 acc;
        for (int k = 0; k < 8; k++) {
            const bool o0 = k & 1, o1 = k & 2, o2 = k & 4;
            const F4 s = sample(o0 ? h1[0] : h0[0], o1 ? h1[1] : h0[1], o2 ? h1[2] : h0[2]);
            const float wv = (o0 ? g1[0] : g0[0]) * (o1 ? g1[1] : g0[1]) * (o2 ? g1[2] : g0[2]);
            acc = acc + s * wv;
        }
        return acc;
    }
};

// repeat-addressed trilinear sample of an n^3 texture at normalised coordinate q
F4 sample_rep(const float* t, int n, float qx, float qy, float qz) {
    const float tx = qx * (float)n - 0.5f, ty = qy * (float)n - 0.5f, tz = qz * (float)n - 0.5f;
    const float flx = std::floor(tx), fly = std::floor(ty), flz = std::floor(tz);
    const long ix = (long)flx, iy = (long)fly, iz = (long)flz;
    const float fx = tx - flx, fy = ty - fly, fz = tz - flz;
    auto L = [&](long x, long y, long z) {
        x = ((x % n) + n) % n; y = ((y % n) + n) % n; z = ((z % n) + n) % n;
        const float* s = t + (((size_t)z * n + y) * n + x) * 4;
        return F4(s[0], s[1], s[2], s[3]);
    };
    F4 c[8];
    for (int k = 0; k < 8; k++) c[k] = L(ix + (k & 1), iy + ((k >> 1) & 1), iz + (k >> 2));
    float w[8];
    tri_weights<float>(fx, fy, fz, w);
    return tri_blend<float>(c, w);
}

// bbRow (src/renderer.js:74-80)
F3 bb_row(const float* U, const float* lut, float T, int row) {
    const float x = wclamp<float>((T - U[bbTmin]) / (U[bbTmax] - U[bbTmin]), 0.0f, 1.0f) * (U[bbN] - 1.0f);
    const unsigned i = (unsigned)std::floor(x);
    const float f = x - std::floor(x);
    const unsigned n = (unsigned)U[bbN];
    const float* a = lut + ((size_t)row * n + i) * 4;
    const float* b = lut + ((size_t)row * n + std::min(i + 1u, n - 1u)) * 4;
    return F3(mix(a[0], b[0], f), mix(a[1], b[1], f), mix(a[2], b[2], f));
}
float lum(F3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }
F3 normalize(F3 v) { const float l = std::sqrt(dot(v, v)); return F3(v.x / l, v.y / l, v.z / l); }
float fract(float x) { return x - std::floor(x); }
float hg(float c, float g) {   // Henyey-Greenstein
    const float g2 = g * g;
    return (1.0f - g2) / (4.0f * 3.14159265f * fpow(1.0f + g2 - 2.0f * g * c, 1.5f));
}
float f16r(float v) { return f16_to_f32(f32_to_f16(v)); }

}  // namespace

extern "C" {

// One dispatch of the light extraction (8 workgroups of 256): lstate (3 x 8 x 4 floats) is read
// and written, lights (8 x 8 floats) written.
void orc_render_lights(const float* U, const float* S, const float* lut, float* lstate, float* lights) {
    const unsigned nx = (unsigned)U[nxI], ny = (unsigned)U[nyI], nz = (unsigned)U[nzI];
    const unsigned bandH = ny / 8u, hx = nx / 2u, hb = bandH / 2u, total = hx * hb * (nz / 2u);
    for (unsigned wg = 0; wg < 8u; wg++) {
        const unsigned y0 = wg * bandH;
        std::vector<F4> accP(256), accC(256);
        for (unsigned li = 0; li < 256u; li++) {
            F4 P, C;
            for (unsigned i = li; i < total; i += 256u) {
                const unsigned ix = (i % hx) * 2u, iy = ((i / hx) % hb) * 2u + y0, iz = (i / (hx * hb)) * 2u;
                const float* s = S + (((size_t)iz * ny + iy) * nx + ix) * 4;
                const float T = 293.0f + s[0] * U[tempScale];
                if (T < 700.0f) continue;
                const F3 bs = bb_row(U, lut, T, 1);
                const float k = s[2] * U[sootEmit];
                const F3 e(k * bs.x * U[emitK], k * bs.y * U[emitK], k * bs.z * U[emitK]);
                const float w = lum(e);
                const F3 pos(U[ox] + ((float)ix + 1.0f) * U[dxI], U[oy] + ((float)iy + 1.0f) * U[dxI],
                             U[oz] + ((float)iz + 1.0f) * U[dxI]);
                P = P + F4(pos.x * w, pos.y * w, pos.z * w, w);
                C = C + F4(e.x, e.y, e.z, dot(pos, pos) * w);
            }
            accP[li] = P;
            accC[li] = C;
        }
        for (unsigned s = 128u; s > 0u; s >>= 1u)
            for (unsigned li = 0; li < s; li++) { accP[li] = accP[li] + accP[li + s]; accC[li] = accC[li] + accC[li + s]; }
        const float w = accP[0].w, dx = U[dxI];
        const bool some = w > 1e-6f;
        const F3 cen = some ? F3(accP[0].x / w, accP[0].y / w, accP[0].z / w) : F3(0.0f, ((float)y0 + (float)bandH * 0.5f) * dx, 0.0f);
        const float spread = some ? std::fmax(accC[0].w / w - dot(cen, cen), 0.0f) : 0.04f;
        const float scale = (float)(nx * bandH * nz) / (float)total;
        const F3 J(accC[0].x * scale * dx * dx * dx, accC[0].y * scale * dx * dx * dx, accC[0].z * scale * dx * dx * dx);
        float* sp = lstate + wg * 4;
        float* sc = lstate + (8 + wg) * 4;
        float* sm = lstate + (16 + wg) * 4;
        if (U[lightDt] < 0.0f || sm[3] <= 0.0f) {
            sp[0] = cen.x; sp[1] = cen.y; sp[2] = cen.z; sp[3] = spread;
            sc[0] = J.x; sc[1] = J.y; sc[2] = J.z; sc[3] = 0.0f;
            sm[0] = J.x; sm[1] = J.y; sm[2] = J.z; sm[3] = 1.0f;
        } else {
            const float a = 1.0f - fexp(-U[lightDt] / 0.08f), ap = 1.0f - fexp(-U[lightDt] / 0.15f),
                        am = 1.0f - fexp(-U[lightDt] / 1.5f);
            sp[0] = mix(sp[0], cen.x, ap); sp[1] = mix(sp[1], cen.y, ap); sp[2] = mix(sp[2], cen.z, ap); sp[3] = mix(sp[3], spread, ap);
            sc[0] = mix(sc[0], J.x, a); sc[1] = mix(sc[1], J.y, a); sc[2] = mix(sc[2], J.z, a); sc[3] = 0.0f;
            sm[0] = mix(sm[0], J.x, am); sm[1] = mix(sm[1], J.y, am); sm[2] = mix(sm[2], J.z, am); sm[3] = 1.0f;
        }
        const float ratio = lum(F3(sc[0], sc[1], sc[2])) / std::fmax(lum(F3(sm[0], sm[1], sm[2])), 1e-9f);
        const float fk = fpow(std::fmax(ratio, 1e-4f), U[flickerK] - 1.0f);
        const F3 Jc(sc[0] * fk, sc[1] * fk, sc[2] * fk);
        const float wl = lum(Jc);
        const bool on = wl > 1e-3f * lum(F3(sm[0], sm[1], sm[2])) && wl > 0.0f;
        float* L = lights + wg * 8;
        L[0] = sp[0]; L[1] = sp[1]; L[2] = sp[2]; L[3] = on ? wl : 0.0f;
        L[4] = on ? Jc.x : 0.0f; L[5] = on ? Jc.y : 0.0f; L[6] = on ? Jc.z : 0.0f; L[7] = sp[3];
    }
}

// The prefilter; out holds the binary16-rounded result as floats.
void orc_render_sharpen(float k, int nx, int ny, int nz, const float* S, float* out) {
    const Grid G{S, nx, ny, nz};
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const F4 s = G.ld(x, y, z);
                F4 n, lo = s, up = s;
                const int off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
                for (int q = 0; q < 6; q++) {
                    const F4 v = G.ld(x + off[q][0], y + off[q][1], z + off[q][2]);
                    n = n + v;
                    lo = F4(std::fmin(lo.x, v.x), std::fmin(lo.y, v.y), std::fmin(lo.z, v.z), std::fmin(lo.w, v.w));
                    up = F4(std::fmax(up.x, v.x), std::fmax(up.y, v.y), std::fmax(up.z, v.z), std::fmax(up.w, v.w));
                }
                float* o = out + (((size_t)z * ny + y) * nx + x) * 4;
                o[0] = f16r(wclamp<float>(s.x + k * (s.x - n.x / 6.0f), lo.x, up.x));
                o[1] = f16r(wclamp<float>(s.y + k * (s.y - n.y / 6.0f), lo.y, up.y));
                o[2] = f16r(wclamp<float>(s.z + k * (s.z - n.z / 6.0f), lo.z, up.z));
                o[3] = f16r(wclamp<float>(s.w + k * (s.w - n.w / 6.0f), lo.w, up.w));
            }
}

// The volume march over every pixel; vol (vw x vh x 4) holds the binary16-rounded (L, Tr), aux
// (vw x vh x 2) the haze and surface depth.  S is the field the march samples (the sim's S or the
// prefiltered copy), S2 the smoke field, curl the n^3 curl texture.
void orc_render_volume(const float* U, const float* S, const float* S2, const float* lut, const float* lights,
                       const float* depth, int dw, int dh, const float* curl, int nn, float* vol, float* aux) {
    const int nx = (int)U[nxI], ny = (int)U[nyI], nz = (int)U[nzI];
    const Grid G{S, nx, ny, nz}, G2{S2, nx, ny, nz};
    const int VW = (int)U[vwI], VH = (int)U[vhI];
    const float vw = U[vwI], vh = U[vhI];
#pragma omp parallel for schedule(dynamic, 1)
    for (int iy = 0; iy < VH; iy++)
        for (int ix = 0; ix < VW; ix++) {
            const float pxx = (float)ix + 0.5f + U[jitX] * (vw / U[width]);
            const float pxy = (float)iy + 0.5f + U[jitY] * (vw / U[width]);
            const float ndx = pxx / vw * 2.0f - 1.0f, ndy = 1.0f - pxy / vh * 2.0f;
            const F3 fwd(U[fwdX], U[fwdY], U[fwdZ]), rgt(U[rgtX], U[rgtY], U[rgtZ]), up(U[upX], U[upY], U[upZ]);
            const F3 dPin = normalize(fwd + rgt * ndx * U[tanFov] * U[aspect] + up * ndy * U[tanFov]);
            const F3 cam(U[camX], U[camY], U[camZ]);
            F3 ro = cam, rd = dPin;
            if (U[focus] > 0.0f) {   // lensRay (src/renderer.js:87-94)
                ro = cam + rgt * U[lensX] + up * U[lensY];
                const F3 q = cam + dPin * (U[focus] / dot(dPin, fwd));
                rd = normalize(q - ro);
            }
            const int fx = (int)(pxx / vw * (float)dw), fy = (int)(pxy / vh * (float)dh);
            const float sd = depth[(size_t)std::min(std::max(fy, 0), dh - 1) * dw + std::min(std::max(fx, 0), dw - 1)];
            const F3 bmin(U[ox], U[oy], U[oz]);
            const F3 bmax = bmin + F3(U[sx], U[sy], U[sz]);
            const F3 inv(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
            const F3 t0((bmin.x - ro.x) * inv.x, (bmin.y - ro.y) * inv.y, (bmin.z - ro.z) * inv.z);
            const F3 t1((bmax.x - ro.x) * inv.x, (bmax.y - ro.y) * inv.y, (bmax.z - ro.z) * inv.z);
            const float tn = std::fmax(std::fmax(std::fmax(std::fmin(t0.x, t1.x), std::fmin(t0.y, t1.y)), std::fmin(t0.z, t1.z)), 0.0f);
            const float tf = std::fmin(std::fmin(std::fmin(std::fmax(t0.x, t1.x), std::fmax(t0.y, t1.y)), std::fmax(t0.z, t1.z)), sd);
            F3 L;
            float Tr = 1.0f, haze = 0.0f;
            if (tf > tn) {
                const float dx = U[dxI], ds = dx * U[stepMul];
                const float fr = U[frameI] - 64.0f * std::trunc(U[frameI] / 64.0f);   // WGSL % on f32
                const float qx = (float)ix + 5.588238f * fr, qy = (float)iy + 5.588238f * fr;
                const float jit = fract(52.9829189f * fract(qx * 0.06711056f + qy * 0.00583715f));
                float t = tn + ds * jit;
                const int nlit = (int)U[nLights];
                for (int i = 0; i < 1200; i++) {
                    if (t > tf || Tr < 0.003f) break;
                    const F3 p = ro + rd * t;
                    F3 g((p.x - bmin.x) / dx, (p.y - bmin.y) / dx, (p.z - bmin.z) / dx);
                    float mottle = 0.0f;
                    if (U[detailAmt] > 0.0f) {
                        const float rise = U[detailSpeed] * (0.5f + 1.5f * smoothstep(0.0f, 1.0f, p.y));
                        const float ph = U[timeI] / 0.7f;
                        const float f0 = fract(ph), f1 = fract(ph + 0.5f);
                        const float w0 = 1.0f - std::fabs(2.0f * f0 - 1.0f);
                        const F3 o0 = F3(0.37f, 0.0f, 0.61f) * std::floor(ph), o1 = F3(0.71f, 0.0f, 0.29f) * std::floor(ph + 0.5f);
                        const F3 q0 = (p - F3(0.0f, rise * f0 * 0.7f, 0.0f)) * U[detailFreq] * 0.25f + o0;
                        const F3 q1 = (p - F3(0.0f, rise * f1 * 0.7f, 0.0f)) * U[detailFreq] * 0.25f + o1;
                        const F4 c0 = sample_rep(curl, nn, q0.x, q0.y, q0.z), c1 = sample_rep(curl, nn, q1.x, q1.y, q1.z);
                        const F4 cw = c0 * w0 + c1 * (1.0f - w0);
                        const float k12 = U[detailAmt] * 0.12f;
                        g = g + F3(cw.x * k12 * 1.0f, cw.y * k12 * 0.6f, cw.z * k12 * 1.0f);
                        mottle = cw.w * 0.08f - 0.6f;
                    }
                    F4 s = U[cubic] > 0.5f ? G.cubic(g) : G.sample(g.x, g.y, g.z);
                    const float edge = std::fmin(std::fmin(g.x, (float)nx - g.x), std::fmin(g.z, (float)nz - g.z));
                    const float fade = smoothstep(0.02f, 0.15f, edge * dx) * smoothstep(0.0f, 0.4f, ((float)ny - g.y) * dx);
                    s = s * fade;
                    const float T = 293.0f + s.x * U[tempScale] + mottle * 60.0f * U[detailAmt] * (s.x >= 0.3f ? 1.0f : 0.0f);
                    float soot = s.z;
                    if (U[sootEdge] > 0.0f) soot = soot * smoothstep(0.6f * U[sootEdge], 1.4f * U[sootEdge], soot);
                    float smoke = 0.0f;
                    if (U[smokeScat] > 0.0f) smoke = G2.sample(g.x, g.y, g.z).z * fade;
                    if (soot > 1e-4f || s.w > 1e-3f || smoke > 1e-4f) {
                        const float sigSoot = soot * U[sootAbs];
                        const float sigS = sigSoot * U[sootScat] + smoke * U[smokeScat];
                        const float sigT = sigSoot + smoke * U[smokeScat] * 1.1f + 1e-6f;
                        F3 e;
                        if (T > U[emitTempLo]) {
                            const F3 bs = bb_row(U, lut, T, 1);
                            const float k = soot * U[sootEmit];
                            e = F3(k * bs.x * U[emitK], k * bs.y * U[emitK], k * bs.z * U[emitK]);
                        }
                        const float pw = fpow(s.w, 1.6f), bk = U[blueK] / 3.0f, ex = fexp(-soot * 2.0f);
                        e = e + F3(0.15f * pw * bk * ex, 0.3f * pw * bk * ex, 1.0f * pw * bk * ex);
                        F3 ins;
                        if (sigS * Tr > 1e-6f) {
                            for (int k = 0; k < nlit; k++) {
                                const float* Lk = lights + k * 8;
                                const F3 toL = F3(Lk[0], Lk[1], Lk[2]) - p;
                                const float r2 = dot(toL, toL) + Lk[7] * 0.6f + 0.004f;
                                const float ph_ = hg(dot(rd, toL * (1.0f / std::sqrt(r2))), U[smokeG]);
                                ins = ins + F3(Lk[4] * ph_ / r2, Lk[5] * ph_ / r2, Lk[6] * ph_ / r2);
                            }
                            ins = ins * (sigS * U[lightK]);
                        }
                        const float a = fexp(-sigT * ds);
                        const float w = sigT * ds < 1e-4f ? ds : (1.0f - a) / sigT;
                        L = L + F3(Tr * (e.x + ins.x) * w, Tr * (e.y + ins.y) * w, Tr * (e.z + ins.z) * w);
                        Tr *= a;
                    }
                    haze += std::fmax(s.x - 0.05f, 0.0f) * ds;
                    t += ds;
                }
            }
            const size_t o = (size_t)iy * VW + ix;
            vol[o * 4 + 0] = f16r(L.x); vol[o * 4 + 1] = f16r(L.y); vol[o * 4 + 2] = f16r(L.z); vol[o * 4 + 3] = f16r(Tr);
            aux[o * 2 + 0] = haze;
            aux[o * 2 + 1] = sd;
        }
}

}  // extern "C"
