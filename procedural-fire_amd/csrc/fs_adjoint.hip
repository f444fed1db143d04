// Discrete adjoint of one fire-simulation step (float4 storage), stage by stage in
// reverse: projection, the pressure solve (the operator is symmetric, so its adjoint is
// a solve with the same multigrid; the driver runs it), divergence, the source kernel,
// curl, the MacCormack correction and the semi-Lagrangian pass.
//
// Every kernel recomputes the forward quantities of its cell from the stage's recorded
// inputs and accumulates the transposed derivatives: gathers write their own cell,
// scatters (trilinear weights, stencil neighbours) use float atomics.  Branches take
// the forward pass's side (the limiter's active set, clamps, min/max selections), so
// this is the derivative of the discrete step wherever it is differentiable.  The source
// kernel's local Jacobian comes from dual-number evaluations of the same cell function
// the step runs (forces_cell), one per local input and one per spot / floor-flux seed.
#include <hip/hip_runtime.h>

#include "fs_codec.h"
#include "fs_forces_cell.h"
#include "fs_internal.h"
#include "fs_math.h"

namespace {

typedef StF32 S;   // the adjoint runs on float4 storage

__device__ __forceinline__ void aadd(float* p, float v) {
    if (v != 0.f) unsafeAtomicAdd(p, v);
}
__device__ __forceinline__ float* comp(float4* a, int i, int c) { return reinterpret_cast<float*>(a + i) + c; }

// ---- trilinear sampling (sample_clamp) and its derivatives -------------------------
struct Tri {
    int idx[8];           // texel indices, corner k = (k&1, k>>1&1, k>>2&1)
    float w[8];           // weights
    float fx, fy, fz;
    float gx, gy, gz;     // 1 where the coordinate was not clamped (position derivative passes)
};
__device__ __forceinline__ Tri tri_at(const Dims& g, float bx, float by, float bz) {
    Tri t;
    const float tx = wclamp(bx - 0.5f, -1.f, (float)g.nx), ty = wclamp(by - 0.5f, -1.f, (float)g.ny),
                tz = wclamp(bz - 0.5f, -1.f, (float)g.nz);
    // the forward sampler's footprint and fractions; a fraction moves with the position only where the
    // footprint is not clamped (floor in [0, n - 2])
    const int xs = tri_axis(tx, g.nx, t.fx), ys = tri_axis(ty, g.ny, t.fy), zs = tri_axis(tz, g.nz, t.fz);
    const int ix = (int)floorf(tx), iy = (int)floorf(ty), iz = (int)floorf(tz);
    t.gx = (ix >= 0 && ix <= g.nx - 2) ? 1.f : 0.f;
    t.gy = (iy >= 0 && iy <= g.ny - 2) ? 1.f : 0.f;
    t.gz = (iz >= 0 && iz <= g.nz - 2) ? 1.f : 0.f;
#pragma unroll
    for (int k = 0; k < 8; k++) t.idx[k] = ((zs + (k >> 2)) * g.ny + ys + ((k >> 1) & 1)) * g.nx + xs + (k & 1);
    tri_weights(t.fx, t.fy, t.fz, t.w);   // the forward sampler's weights
    return t;
}
// value (the forward sampler's FMA chain) and gradient (d/dx, d/dy, d/dz of the position) of channel values v[8]
__device__ __forceinline__ void tri_eval(const Tri& t, const float* v, float& val, float& dx, float& dy, float& dz) {
    val = v[0] * t.w[0];
#pragma unroll
    for (int k = 1; k < 8; k++) val = __builtin_fmaf(v[k], t.w[k], val);
    const float fx = t.fx, fy = t.fy, fz = t.fz;
    dx = t.gx * ((1.f - fy) * (1.f - fz) * (v[1] - v[0]) + fy * (1.f - fz) * (v[3] - v[2]) +
                 (1.f - fy) * fz * (v[5] - v[4]) + fy * fz * (v[7] - v[6]));
    dy = t.gy * ((1.f - fx) * (1.f - fz) * (v[2] - v[0]) + fx * (1.f - fz) * (v[3] - v[1]) +
                 (1.f - fx) * fz * (v[6] - v[4]) + fx * fz * (v[7] - v[5]));
    dz = t.gz * ((1.f - fx) * (1.f - fy) * (v[4] - v[0]) + fx * (1.f - fy) * (v[5] - v[1]) +
                 (1.f - fx) * fy * (v[6] - v[2]) + fx * fy * (v[7] - v[3]));
}
__device__ __forceinline__ void tri_load(const Tri& t, const float4* f, float4* v) {
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = f[t.idx[k]];
}
__device__ __forceinline__ float chan(const float4& a, int c) { return c == 0 ? a.x : c == 1 ? a.y : c == 2 ? a.z : a.w; }

// outside_fade (sim.js:137-140) and its gradient
__device__ __forceinline__ float fade_grad(const Dims& g, float x, float y, float z, float& gx, float& gy, float& gz) {
    const float lo = fminf(x, z);
    const float hx = x - (float)g.nx, hy = y - (float)g.ny, hz = z - (float)g.nz;
    const float hi = fmaxf(hx, fmaxf(hy, hz));
    const float m = fmaxf(-lo, hi);
    const float r = 1.f - m;
    gx = gy = gz = 0.f;
    if (r > 0.f && r < 1.f) {
        if (-lo >= hi) {
            if (x <= z) gx = 1.f; else gz = 1.f;       // r = 1 + min(x, z)
        } else {
            if (hx >= hy && hx >= hz) gx = -1.f;        // r = 1 - (x - nx), ...
            else if (hy >= hz) gy = -1.f;
            else gz = -1.f;
        }
    }
    return wclamp(r, 0.f, 1.f);
}
__device__ __forceinline__ float dsstep(float e0, float e1, float x) {   // d/dx smoothstep
    const float t = (x - e0) / (e1 - e0);
    if (t <= 0.f || t >= 1.f) return 0.f;
    return 6.f * t * (1.f - t) / (e1 - e0);
}

// ---- adv1^T (src/sim.js:142-157) -------------------------------------------------------
__global__ __launch_bounds__(256) void k_adv1_adj(const DevParams* __restrict__ P, Dims g, const float4* __restrict__ vel,
                                                  const float4* __restrict__ sF, const float4* __restrict__ s2F,
                                                  const float4* __restrict__ lvX, const float4* __restrict__ lsX,
                                                  const float4* __restrict__ ls2X, float4* lvel, float4* ls, float4* ls2) {
    Cell c;
    if (!cell_of(g, c)) return;
    const float dt = (float)P->dt, dx = P->f[P_dx];
    const float cx = (float)c.x + 0.5f, cy = (float)c.y + 0.5f, cz = (float)c.z + 0.5f;
    // forward: h = c - v0 dt/2dx, v1 = vel(h), b = c - v1 dt/dx
    const float4 v0 = vel[c.i];
    const float hx = cx - v0.x * (0.5f * dt) / dx, hy = cy - v0.y * (0.5f * dt) / dx, hz = cz - v0.z * (0.5f * dt) / dx;
    const Tri th = tri_at(g, hx, hy, hz);
    float4 vh[8];
    tri_load(th, vel, vh);
    float v1[3], dv1[3][3];
    for (int ch = 0; ch < 3; ch++) {
        float cv[8];
        for (int k = 0; k < 8; k++) cv[k] = chan(vh[k], ch);
        tri_eval(th, cv, v1[ch], dv1[ch][0], dv1[ch][1], dv1[ch][2]);
    }
    const float bx = cx - v1[0] * dt / dx, by = cy - v1[1] * dt / dx, bz = cz - v1[2] * dt / dx;
    float fgx, fgy, fgz;
    const float f = fade_grad(g, bx, by, bz, fgx, fgy, fgz);
    const float ws = sstep(0.03f, 0.4f, by * dx);
    const float wind[3] = {P->f[P_windX] * ws, 0.f, P->f[P_windZ] * ws};
    const Tri tb = tri_at(g, bx, by, bz);
    const float4 LV = lvX[c.i], LS = lsX[c.i], LS2 = ls2X[c.i];
    float lb[3] = {0.f, 0.f, 0.f}, lf = 0.f;
    // the three sampled fields at b
    const float4* fields[3] = {vel, sF, s2F};
    float4* lfields[3] = {lvel, ls, ls2};
    const float4 lam[3] = {LV, LS, LS2};
    for (int fi = 0; fi < 3; fi++) {
        float4 tv[8];
        tri_load(tb, fields[fi], tv);
        const int nch = fi == 0 ? 3 : 4;
        for (int ch = 0; ch < nch; ch++) {
            float cv[8], val, gxv, gyv, gzv;
            for (int k = 0; k < 8; k++) cv[k] = chan(tv[k], ch);
            tri_eval(tb, cv, val, gxv, gyv, gzv);
            const float L = chan(lam[fi], ch);
            if (L == 0.f) continue;
            // velX = wind (1 - f) + sv f ; scalars = sample f
            lf += fi == 0 ? L * (val - wind[ch]) : L * val;
            const float ls_ = L * f;
            lb[0] += ls_ * gxv; lb[1] += ls_ * gyv; lb[2] += ls_ * gzv;
            for (int k = 0; k < 8; k++) aadd(comp(lfields[fi], tb.idx[k], ch), ls_ * tb.w[k]);
        }
    }
    // wind mix through ws(b.y)
    const float lws = LV.x * (1.f - f) * P->f[P_windX] + LV.z * (1.f - f) * P->f[P_windZ];
    lb[1] += lws * dsstep(0.03f, 0.4f, by * dx) * dx;
    lb[0] += lf * fgx; lb[1] += lf * fgy; lb[2] += lf * fgz;
    // b = c - v1 dt/dx
    float lv1[3];
    for (int ch = 0; ch < 3; ch++) lv1[ch] = -lb[ch] * dt / dx;
    float lh[3] = {0.f, 0.f, 0.f};
    for (int ch = 0; ch < 3; ch++) {
        if (lv1[ch] == 0.f) continue;
        for (int k = 0; k < 8; k++) aadd(comp(lvel, th.idx[k], ch), lv1[ch] * th.w[k]);
        lh[0] += lv1[ch] * dv1[ch][0]; lh[1] += lv1[ch] * dv1[ch][1]; lh[2] += lv1[ch] * dv1[ch][2];
    }
    // h = c - v0 (dt/2)/dx
    for (int ch = 0; ch < 3; ch++) aadd(comp(lvel, c.i, ch), -lh[ch] * (0.5f * dt) / dx);
}

// ---- adv2^T (src/sim.js:159-197): MacCormack correction with the limiter frozen ------------
__global__ __launch_bounds__(256) void k_adv2_adj(const DevParams* __restrict__ P, Dims g, const float4* __restrict__ vel,
                                                  const float4* __restrict__ sF, const float4* __restrict__ s2F,
                                                  const float4* __restrict__ vH, const float4* __restrict__ sH,
                                                  const float4* __restrict__ s2H, const float4* __restrict__ lvY,
                                                  const float4* __restrict__ lsY, const float4* __restrict__ ls2Y,
                                                  float4* lvel, float4* ls, float4* ls2, float4* lvH, float4* lsH,
                                                  float4* ls2H) {
    Cell c;
    if (!cell_of(g, c)) return;
    const float dt = (float)P->dt, dx = P->f[P_dx];
    const float cx = (float)c.x + 0.5f, cy = (float)c.y + 0.5f, cz = (float)c.z + 0.5f;
    const float4 v = vel[c.i];
    const float fwx = cx + dt * v.x / dx, fwy = cy + dt * v.y / dx, fwz = cz + dt * v.z / dx;
    const Tri tf = tri_at(g, fwx, fwy, fwz);
    const float3 bt = backtrace<S>(vel, g, cx, cy, cz, dt, dx);
    const float bx = bt.x - 0.5f, by = bt.y - 0.5f, bz = bt.z - 0.5f;
    const int b0x = (int)floorf(wclamp(bx, -2.f, (float)g.nx + 1.f));
    const int b0y = (int)floorf(wclamp(by, -2.f, (float)g.ny + 1.f));
    const int b0z = (int)floorf(wclamp(bz, -2.f, (float)g.nz + 1.f));
    int lim[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int qx = clampi(b0x + (k & 1), 0, g.nx - 1), qy = clampi(b0y + ((k >> 1) & 1), 0, g.ny - 1);
        const int qz = clampi(b0z + ((k >> 2) & 1), 0, g.nz - 1);
        lim[k] = (qz * g.ny + qy) * g.nx + qx;
    }
    const bool reverted = outside_fade(g, bx + 0.5f, by + 0.5f, bz + 0.5f) < 1.f;
    const float4* orig[3] = {vel, sF, s2F};
    const float4* hf[3] = {vH, sH, s2H};
    float4* lorig[3] = {lvel, ls, ls2};
    float4* lh[3] = {lvH, lsH, ls2H};
    const float4 lam[3] = {lvY[c.i], lsY[c.i], ls2Y[c.i]};
    float lfw[3] = {0.f, 0.f, 0.f};
    for (int fi = 0; fi < 3; fi++) {
        const int nch = fi == 0 ? 3 : 4;
        const float4 hc = hf[fi][c.i], oc = orig[fi][c.i];
        float4 tv[8], lv[8];
        tri_load(tf, hf[fi], tv);
        for (int k = 0; k < 8; k++) lv[k] = orig[fi][lim[k]];
        for (int ch = 0; ch < nch; ch++) {
            float L = chan(lam[fi], ch);
            if (L == 0.f) continue;
            float cv[8], hs, gxv, gyv, gzv;
            for (int k = 0; k < 8; k++) cv[k] = chan(tv[k], ch);
            tri_eval(tf, cv, hs, gxv, gyv, gzv);
            const float h = chan(hc, ch), o = chan(oc, ch);
            float N = h + (o - hs) * 0.5f;
            if (reverted) N = h;
            // limiter bounds over the 8 texels (first texel achieving each), widened by h
            float mn = 1e9f, mx = -1e9f;
            int kmn = -1, kmx = -1;
            for (int k = 0; k < 8; k++) {
                const float t = chan(lv[k], ch);
                if (kmn < 0 || t < mn) { mn = t; kmn = k; }
                if (kmx < 0 || t > mx) { mx = t; kmx = k; }
            }
            const float lo = fminf(mn, h), hi = fmaxf(mx, h);
            const float out = wclamp(N, lo, hi);
            if (fi > 0 && !(out > 0.f)) continue;        // max(., 0) of the scalars
            if (N < lo) {                                  // clamped to the lower bound
                if (mn <= h) aadd(comp(lorig[fi], lim[kmn], ch), L); else aadd(comp(lh[fi], c.i, ch), L);
            } else if (N > hi) {
                if (mx >= h) aadd(comp(lorig[fi], lim[kmx], ch), L); else aadd(comp(lh[fi], c.i, ch), L);
            } else if (reverted) {
                aadd(comp(lh[fi], c.i, ch), L);
            } else {
                aadd(comp(lh[fi], c.i, ch), L);
                aadd(comp(lorig[fi], c.i, ch), 0.5f * L);
                for (int k = 0; k < 8; k++) aadd(comp(lh[fi], tf.idx[k], ch), -0.5f * L * tf.w[k]);
                lfw[0] += -0.5f * L * gxv; lfw[1] += -0.5f * L * gyv; lfw[2] += -0.5f * L * gzv;
            }
        }
    }
    // fw = c + dt v / dx
    for (int ch = 0; ch < 3; ch++) aadd(comp(lvel, c.i, ch), lfw[ch] * dt / dx);
}

// ---- curl^T (src/sim.js:200-211) ---------------------------------------------------------
__global__ __launch_bounds__(256) void k_curl_adj(Dims g, const float4* __restrict__ v, const float4* __restrict__ lvort,
                                                  float4* lv) {
    Cell c;
    if (!cell_of(g, c)) return;
    const int iL = (c.z * g.ny + c.y) * g.nx + clampi(c.x - 1, 0, g.nx - 1), iR = (c.z * g.ny + c.y) * g.nx + clampi(c.x + 1, 0, g.nx - 1);
    const int iB = (c.z * g.ny + clampi(c.y - 1, 0, g.ny - 1)) * g.nx + c.x, iT = (c.z * g.ny + clampi(c.y + 1, 0, g.ny - 1)) * g.nx + c.x;
    const int iK = (clampi(c.z - 1, 0, g.nz - 1) * g.ny + c.y) * g.nx + c.x, iF = (clampi(c.z + 1, 0, g.nz - 1) * g.ny + c.y) * g.nx + c.x;
    const float4 L = v[iL], R = v[iR], B = v[iB], T = v[iT], K = v[iK], F = v[iF];
    const float wx = ((T.z - B.z) - (F.y - K.y)) * 0.5f;
    const float wy = ((F.x - K.x) - (R.z - L.z)) * 0.5f;
    const float wz = ((R.y - L.y) - (T.x - B.x)) * 0.5f;
    const float wl = sqrtf(wx * wx + wy * wy + wz * wz);
    const float4 lw = lvort[c.i];
    float ax = lw.x, ay = lw.y, az = lw.z;
    if (wl > 0.f) { ax += lw.w * wx / wl; ay += lw.w * wy / wl; az += lw.w * wz / wl; }
    ax *= 0.5f; ay *= 0.5f; az *= 0.5f;
    aadd(comp(lv, iT, 2), ax); aadd(comp(lv, iB, 2), -ax); aadd(comp(lv, iF, 1), -ax); aadd(comp(lv, iK, 1), ax);
    aadd(comp(lv, iF, 0), ay); aadd(comp(lv, iK, 0), -ay); aadd(comp(lv, iR, 2), -ay); aadd(comp(lv, iL, 2), ay);
    aadd(comp(lv, iR, 1), az); aadd(comp(lv, iL, 1), -az); aadd(comp(lv, iT, 0), -az); aadd(comp(lv, iB, 0), az);
}

// ---- source kernel^T (src/sim.js:214-396) through forces_cell<Dual> ------------------------
// v(3) s(4) s2(4) w(3) |w| of the 6 neighbours, and with hot-gas viscosity on the velocity Laplacian (3),
// whose adjoint goes to llap for k_visc_adj to spread over the stencil
constexpr int FORCES_IN = 23;
constexpr int FORCES_IN_INVISCID = 20;
__global__ __launch_bounds__(256) void k_forces_adj(const DevParams* __restrict__ P, Dims g, const float4* __restrict__ velY,
                                                    const float4* __restrict__ sY, const float4* __restrict__ s2Y,
                                                    const float4* __restrict__ vort, const float4* __restrict__ geo,
                                                    const half4_t* __restrict__ noise, const half4_t* __restrict__ curlT, int nn,
                                                    const fs_spot* __restrict__ spots, const float* __restrict__ faceFlux,
                                                    const float4* __restrict__ lvOut, const float4* __restrict__ lsOut,
                                                    const float4* __restrict__ ls2Out, float4* lvY, float4* lsY, float4* ls2Y,
                                                    float4* lvort, float4* __restrict__ llap, double* gspot, double* gface) {
    Cell c;
    if (!cell_of(g, c)) return;
    const ForcesCell k = forces_setup<S>(P, g, c, velY, geo, noise, curlT, nn);
    const bool visc = P->f[P_visc] > 0.f;
    const int nin = visc ? FORCES_IN : FORCES_IN_INVISCID;
    const float4 v4 = velY[c.i], s4 = sY[c.i], t4 = s2Y[c.i], w4 = vort[c.i];
    int nb[6];
    nb[0] = (c.z * g.ny + c.y) * g.nx + clampi(c.x + 1, 0, g.nx - 1);
    nb[1] = (c.z * g.ny + c.y) * g.nx + clampi(c.x - 1, 0, g.nx - 1);
    nb[2] = (c.z * g.ny + clampi(c.y + 1, 0, g.ny - 1)) * g.nx + c.x;
    nb[3] = (c.z * g.ny + clampi(c.y - 1, 0, g.ny - 1)) * g.nx + c.x;
    nb[4] = (clampi(c.z + 1, 0, g.nz - 1) * g.ny + c.y) * g.nx + c.x;
    nb[5] = (clampi(c.z - 1, 0, g.nz - 1) * g.ny + c.y) * g.nx + c.x;
    float wn[6];
    for (int j = 0; j < 6; j++) wn[j] = vort[nb[j]].w;
    const float in0[FORCES_IN] = {v4.x, v4.y, v4.z, s4.x, s4.y, s4.z, s4.w, t4.x, t4.y, t4.z, t4.w,
                                  w4.x, w4.y, w4.z, wn[0], wn[1], wn[2], wn[3], wn[4], wn[5], k.lapx, k.lapy, k.lapz};
    const float4 Lv = lvOut[c.i], Ls = lsOut[c.i], Ls2 = ls2Out[c.i];
    const bool face = P->has_face_flux != 0;
    const float fv = (face && c.y == 0) ? faceFlux[c.z * g.nx + c.x] : 0.f;
    // one dual evaluation: seed input `si` (or spot `sp`, or the floor flux), returns lambda . d(out)
    auto pass = [&](int si, int sp, bool seed_face) -> float {
        Dual x[FORCES_IN];
        for (int j = 0; j < FORCES_IN; j++) x[j] = Dual(in0[j], j == si ? 1.f : 0.f);
        Dual vx = x[0], vy = x[1], vz = x[2], phi;
        Dual s[4] = {x[3], x[4], x[5], x[6]}, s2[4] = {x[7], x[8], x[9], x[10]};
        const Dual w[3] = {x[11], x[12], x[13]};
        const Dual gx = x[14] - x[15], gy = x[16] - x[17], gz = x[18] - x[19];
        const Dual lap[3] = {x[20], x[21], x[22]};
        forces_cell<Dual>(P->f, P->d, k, spots, P->n_spots, sp, nullptr, face, Dual(fv, seed_face ? 1.f : 0.f), vx, vy, vz,
                          s, s2, w, gx, gy, gz, lap, phi);
        return Lv.x * vx.d + Lv.y * vy.d + Lv.z * vz.d + Lv.w * phi.d + Ls.x * s[0].d + Ls.y * s[1].d + Ls.z * s[2].d +
               Ls.w * s[3].d + Ls2.x * s2[0].d + Ls2.y * s2[1].d + Ls2.z * s2[2].d + Ls2.w * s2[3].d;
    };
    // the 20 state inputs: four seeded per vector-mode pass (Dual4 gives each column the scalar pass's bits)
    float l[FORCES_IN] = {};
    for (int j0 = 0; j0 < nin; j0 += 4) {
        Dual4 x[FORCES_IN];
#pragma unroll
        for (int j = 0; j < FORCES_IN; j++) {
            x[j] = Dual4(in0[j]);
            FS_D4(x[j].d[u] = j == j0 + u ? 1.f : 0.f);
        }
        Dual4 vx = x[0], vy = x[1], vz = x[2], phi;
        Dual4 s[4] = {x[3], x[4], x[5], x[6]}, s2[4] = {x[7], x[8], x[9], x[10]};
        const Dual4 w[3] = {x[11], x[12], x[13]};
        const Dual4 gx = x[14] - x[15], gy = x[16] - x[17], gz = x[18] - x[19];
        const Dual4 lap[3] = {x[20], x[21], x[22]};
        forces_cell<Dual4>(P->f, P->d, k, spots, P->n_spots, -1, nullptr, face, Dual4(fv), vx, vy, vz, s, s2, w, gx, gy,
                           gz, lap, phi);
        float out[4];
        FS_D4(out[u] = Lv.x * vx.d[u] + Lv.y * vy.d[u] + Lv.z * vz.d[u] + Lv.w * phi.d[u] + Ls.x * s[0].d[u] +
                       Ls.y * s[1].d[u] + Ls.z * s[2].d[u] + Ls.w * s[3].d[u] + Ls2.x * s2[0].d[u] + Ls2.y * s2[1].d[u] +
                       Ls2.z * s2[2].d[u] + Ls2.w * s2[3].d[u]);
#pragma unroll
        for (int j = 0; j < FORCES_IN; j++)
            FS_D4(if (j == j0 + u) l[j] = out[u]);
    }
    lvY[c.i] = make_float4(l[0], l[1], l[2], 0.f);
    lsY[c.i] = make_float4(l[3], l[4], l[5], l[6]);
    ls2Y[c.i] = make_float4(l[7], l[8], l[9], l[10]);
    // own vorticity components are written by this cell only; |w| of the neighbours by atomics
    float* lw = reinterpret_cast<float*>(lvort + c.i);
    lw[0] = l[11]; lw[1] = l[12]; lw[2] = l[13];
    aadd(comp(lvort, nb[0], 3), l[14]); aadd(comp(lvort, nb[1], 3), l[15]);
    aadd(comp(lvort, nb[2], 3), l[16]); aadd(comp(lvort, nb[3], 3), l[17]);
    aadd(comp(lvort, nb[4], 3), l[18]); aadd(comp(lvort, nb[5], 3), l[19]);
    if (visc) llap[c.i] = make_float4(l[20], l[21], l[22], 0.f);
    // parameters: the spots whose sphere holds this cell, the floor flux of a floor cell
    for (int q = 0; q < P->n_spots; q++) {
        const fs_spot sp = spots[q];
        const float ddx = k.px - sp.pos[0], ddy = k.py - sp.pos[1], ddz = k.pz - sp.pos[2];
        if (sqrtf(ddx * ddx + ddy * ddy + ddz * ddz) < sp.radius) {
            const float gq = pass(-1, q, false);
            if (gq != 0.f) unsafeAtomicAdd(gspot + q, (double)gq);
        }
    }
    if (face && c.y == 0) gface[c.z * g.nx + c.x] += (double)pass(-1, -1, true);
}

// ---- the velocity Laplacian of the viscous term, transposed (src/sim.js:339-346) -------------
// lap(c) = sum over the 6 directions d of v(clamp(c + d)) - 6 v(c) (clamped texel loads), so cell m
// receives -6 llap(m), llap(m - d) for every direction d with m - d inside the grid, and llap(m) once
// more for every d with m + d outside it (there the clamped neighbour of m is m itself).  A gather:
// each cell adds to its own adjoint only.
__global__ __launch_bounds__(256) void k_visc_adj(Dims g, const float4* __restrict__ llap, float4* __restrict__ lvY) {
    Cell c;
    if (!cell_of(g, c)) return;
    const float4 me = llap[c.i];
    float ax = -6.f * me.x, ay = -6.f * me.y, az = -6.f * me.z;
    const int sx = 1, sy = g.nx, sz = g.nx * g.ny;
    const int pos[3] = {c.x, c.y, c.z}, dim[3] = {g.nx, g.ny, g.nz}, st[3] = {sx, sy, sz};
#pragma unroll
    for (int a = 0; a < 3; a++) {
#pragma unroll
        for (int sgn = -1; sgn <= 1; sgn += 2) {
            // direction d = sgn e_a: the cell m - d uses m as its neighbour along d
            const int q = pos[a] - sgn;
            if (q >= 0 && q < dim[a]) {
                const float4 o = llap[c.i - sgn * st[a]];
                ax += o.x; ay += o.y; az += o.z;
            }
            const int e = pos[a] + sgn;
            if (e < 0 || e >= dim[a]) { ax += me.x; ay += me.y; az += me.z; }
        }
    }
    float4 v = lvY[c.i];
    v.x += ax; v.y += ay; v.z += az;
    lvY[c.i] = v;
}

// ---- div^T (src/sim.js:399-413) -----------------------------------------------------------
template <class PT>
__global__ __launch_bounds__(256) void k_div_adj(const DevParams* __restrict__ P, Dims g, const PT* __restrict__ lb,
                                                 float4* lvF) {
    Cell c;
    if (!cell_of(g, c)) return;
    const float dx = P->f[P_dx];
    const float L = (float)lb[rb_index(c.x, c.y, c.z, g.nx, g.ny)];
    if (L == 0.f) return;
    const float k = L / (2.f * dx);
    aadd(comp(lvF, (c.z * g.ny + c.y) * g.nx + clampi(c.x + 1, 0, g.nx - 1), 0), k);
    aadd(comp(lvF, (c.z * g.ny + c.y) * g.nx + clampi(c.x - 1, 0, g.nx - 1), 0), -k);
    aadd(comp(lvF, (c.z * g.ny + clampi(c.y + 1, 0, g.ny - 1)) * g.nx + c.x, 1), k);
    if (c.y == 0) aadd(comp(lvF, c.i, 1), k);            // B = -v.y at the floor
    else aadd(comp(lvF, (c.z * g.ny + c.y - 1) * g.nx + c.x, 1), -k);
    aadd(comp(lvF, (clampi(c.z + 1, 0, g.nz - 1) * g.ny + c.y) * g.nx + c.x, 2), k);
    aadd(comp(lvF, (clampi(c.z - 1, 0, g.nz - 1) * g.ny + c.y) * g.nx + c.x, 2), -k);
    aadd(comp(lvF, c.i, 3), -L);
}

// ---- project^T (src/sim.js:526-550) --------------------------------------------------------
template <class PT>
__global__ __launch_bounds__(256) void k_project_adj(const DevParams* __restrict__ P, Dims g, const float4* __restrict__ vIn,
                                                     const PT* __restrict__ p, const float4* __restrict__ geo,
                                                     const float4* __restrict__ lvOut, float4* lvIn, float* lp) {
    Cell c;
    if (!cell_of(g, c)) return;
    const float dx = P->f[P_dx];
    const int io = rb_index(c.x, c.y, c.z, g.nx, g.ny);
    const float me = (float)p[io];
    const float pxp = c.x + 1 < g.nx ? (float)p[rb_index(c.x + 1, c.y, c.z, g.nx, g.ny)] : -me;
    const float pxm = c.x > 0 ? (float)p[rb_index(c.x - 1, c.y, c.z, g.nx, g.ny)] : -me;
    const float pyp = c.y + 1 < g.ny ? (float)p[rb_index(c.x, c.y + 1, c.z, g.nx, g.ny)] : -me;
    const float pym = c.y > 0 ? (float)p[rb_index(c.x, c.y - 1, c.z, g.nx, g.ny)] : me;
    const float pzp = c.z + 1 < g.nz ? (float)p[rb_index(c.x, c.y, c.z + 1, g.nx, g.ny)] : -me;
    const float pzm = c.z > 0 ? (float)p[rb_index(c.x, c.y, c.z - 1, g.nx, g.ny)] : -me;
    const float4 vi = vIn[c.i];
    float vx = vi.x - (pxp - pxm) / (2.f * dx), vy = vi.y - (pyp - pym) / (2.f * dx), vz = vi.z - (pzp - pzm) / (2.f * dx);
    const float4 L4 = lvOut[c.i];
    float lx = L4.x, ly = L4.y, lz = L4.z;
    if (geo[c.i].z > 0.5f) { lvIn[c.i] = make_float4(0.f, 0.f, 0.f, 0.f); return; }
    bool floor_clamped = false;
    if (c.y == 0 && vy < 0.f) { vy = 0.f; floor_clamped = true; }
    const float vl = sqrtf(vx * vx + vy * vy + vz * vz);
    if (vl > 20.f) {                                       // v 20 / |v|
        const float d = (lx * vx + ly * vy + lz * vz) / vl;
        const float s = 20.f / vl;
        lx = s * (lx - d * vx / vl); ly = s * (ly - d * vy / vl); lz = s * (lz - d * vz / vl);
    }
    if (floor_clamped) ly = 0.f;
    lvIn[c.i] = make_float4(lx, ly, lz, 0.f);              // phi's adjoint comes from div^T
    // v = vIn - grad p: d/dp of -(pxp - pxm)/(2dx) etc.
    const float kx = -lx / (2.f * dx), ky = -ly / (2.f * dx), kz = -lz / (2.f * dx);
    if (c.x + 1 < g.nx) aadd(lp + rb_index(c.x + 1, c.y, c.z, g.nx, g.ny), kx); else aadd(lp + io, -kx);
    if (c.x > 0) aadd(lp + rb_index(c.x - 1, c.y, c.z, g.nx, g.ny), -kx); else aadd(lp + io, kx);
    if (c.y + 1 < g.ny) aadd(lp + rb_index(c.x, c.y + 1, c.z, g.nx, g.ny), ky); else aadd(lp + io, -ky);
    if (c.y > 0) aadd(lp + rb_index(c.x, c.y - 1, c.z, g.nx, g.ny), -ky); else aadd(lp + io, -ky);
    if (c.z + 1 < g.nz) aadd(lp + rb_index(c.x, c.y, c.z + 1, g.nx, g.ny), kz); else aadd(lp + io, -kz);
    if (c.z > 0) aadd(lp + rb_index(c.x, c.y, c.z - 1, g.nx, g.ny), -kz); else aadd(lp + io, kz);
}

// ---- the loss and its seed: L = sum (S.x - T*)^2 + (S.z - soot*)^2 ------------------------
__global__ __launch_bounds__(256) void k_loss_seed(Dims g, const float4* __restrict__ Sf, const float* __restrict__ tT,
                                                   const float* __restrict__ tSoot, float4* lS, double* partial) {
    __shared__ double sh[256];
    Cell c;
    double acc = 0.0;
    if (cell_of(g, c)) {
        const float4 s = Sf[c.i];
        const float a = s.x - tT[c.i], b = s.z - tSoot[c.i];
        lS[c.i] = make_float4(2.f * a, 0.f, 2.f * b, 0.f);
        acc = (double)a * a + (double)b * b;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}


// ---- tangent-linear (forward-mode) kernels: the same Jacobians applied to a direction ----
// Each mirrors its adjoint above branch for branch (the same frozen active sets and tie
// rules), so <adjoint(w), d> = <w, tangent(d)> holds to rounding on any run of steps.

// adv1 (src/sim.js:142-157)
__global__ __launch_bounds__(256) void k_adv1_tan(const DevParams* __restrict__ P, Dims g, const float4* __restrict__ vel,
                                                  const float4* __restrict__ sF, const float4* __restrict__ s2F,
                                                  const float4* __restrict__ tvel, const float4* __restrict__ tS,
                                                  const float4* __restrict__ tS2, float4* __restrict__ tvX,
                                                  float4* __restrict__ tsX, float4* __restrict__ ts2X) {
    Cell c;
    if (!cell_of(g, c)) return;
    const float dt = (float)P->dt, dx = P->f[P_dx];
    const float cx = (float)c.x + 0.5f, cy = (float)c.y + 0.5f, cz = (float)c.z + 0.5f;
    const float4 v0 = vel[c.i], tv0 = tvel[c.i];
    const float hx = cx - v0.x * (0.5f * dt) / dx, hy = cy - v0.y * (0.5f * dt) / dx, hz = cz - v0.z * (0.5f * dt) / dx;
    const float th[3] = {-tv0.x * (0.5f * dt) / dx, -tv0.y * (0.5f * dt) / dx, -tv0.z * (0.5f * dt) / dx};
    const Tri t1 = tri_at(g, hx, hy, hz);
    float4 vh[8], tvh[8];
    tri_load(t1, vel, vh);
    tri_load(t1, tvel, tvh);
    float v1[3], tv1[3];
    for (int ch = 0; ch < 3; ch++) {
        float cv[8], ct = 0.f, gx, gy, gz;
        for (int k = 0; k < 8; k++) { cv[k] = chan(vh[k], ch); ct += t1.w[k] * chan(tvh[k], ch); }
        tri_eval(t1, cv, v1[ch], gx, gy, gz);
        tv1[ch] = ct + gx * th[0] + gy * th[1] + gz * th[2];
    }
    const float bx = cx - v1[0] * dt / dx, by = cy - v1[1] * dt / dx, bz = cz - v1[2] * dt / dx;
    const float tb[3] = {-tv1[0] * dt / dx, -tv1[1] * dt / dx, -tv1[2] * dt / dx};
    float fgx, fgy, fgz;
    const float f = fade_grad(g, bx, by, bz, fgx, fgy, fgz);
    const float tf = fgx * tb[0] + fgy * tb[1] + fgz * tb[2];
    const float ws = sstep(0.03f, 0.4f, by * dx);
    const float tws = dsstep(0.03f, 0.4f, by * dx) * dx * tb[1];
    const float W[3] = {P->f[P_windX], 0.f, P->f[P_windZ]};
    const Tri t2 = tri_at(g, bx, by, bz);
    const float4* fields[3] = {vel, sF, s2F};
    const float4* tfields[3] = {tvel, tS, tS2};
    float4* outs[3] = {tvX, tsX, ts2X};
    for (int fi = 0; fi < 3; fi++) {
        float4 tv[8], tt[8];
        tri_load(t2, fields[fi], tv);
        tri_load(t2, tfields[fi], tt);
        float o[4] = {0.f, 0.f, 0.f, 0.f};
        const int nch = fi == 0 ? 3 : 4;
        for (int ch = 0; ch < nch; ch++) {
            float cv[8], ct = 0.f, val, gx, gy, gz;
            for (int k = 0; k < 8; k++) { cv[k] = chan(tv[k], ch); ct += t2.w[k] * chan(tt[k], ch); }
            tri_eval(t2, cv, val, gx, gy, gz);
            const float tsv = ct + gx * tb[0] + gy * tb[1] + gz * tb[2];
            if (fi == 0) o[ch] = (val - W[ch] * ws) * tf + f * tsv + (1.f - f) * W[ch] * tws;
            else o[ch] = val * tf + f * tsv;
        }
        outs[fi][c.i] = make_float4(o[0], o[1], o[2], o[3]);
    }
}

// adv2 (src/sim.js:159-197), limiter frozen
__global__ __launch_bounds__(256) void k_adv2_tan(const DevParams* __restrict__ P, Dims g, const float4* __restrict__ vel,
                                                  const float4* __restrict__ sF, const float4* __restrict__ s2F,
                                                  const float4* __restrict__ vH, const float4* __restrict__ sH,
                                                  const float4* __restrict__ s2H, const float4* __restrict__ tvel,
                                                  const float4* __restrict__ tS, const float4* __restrict__ tS2,
                                                  const float4* __restrict__ tvH, const float4* __restrict__ tsH,
                                                  const float4* __restrict__ ts2H, float4* __restrict__ tvY,
                                                  float4* __restrict__ tsY, float4* __restrict__ ts2Y) {
    Cell c;
    if (!cell_of(g, c)) return;
    const float dt = (float)P->dt, dx = P->f[P_dx];
    const float cx = (float)c.x + 0.5f, cy = (float)c.y + 0.5f, cz = (float)c.z + 0.5f;
    const float4 v = vel[c.i], tv = tvel[c.i];
    const float fwx = cx + dt * v.x / dx, fwy = cy + dt * v.y / dx, fwz = cz + dt * v.z / dx;
    const float tfw[3] = {dt * tv.x / dx, dt * tv.y / dx, dt * tv.z / dx};
    const Tri tf = tri_at(g, fwx, fwy, fwz);
    const float3 bt = backtrace<S>(vel, g, cx, cy, cz, dt, dx);
    const float bx = bt.x - 0.5f, by = bt.y - 0.5f, bz = bt.z - 0.5f;
    const int b0x = (int)floorf(wclamp(bx, -2.f, (float)g.nx + 1.f));
    const int b0y = (int)floorf(wclamp(by, -2.f, (float)g.ny + 1.f));
    const int b0z = (int)floorf(wclamp(bz, -2.f, (float)g.nz + 1.f));
    int lim[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int qx = clampi(b0x + (k & 1), 0, g.nx - 1), qy = clampi(b0y + ((k >> 1) & 1), 0, g.ny - 1);
        const int qz = clampi(b0z + ((k >> 2) & 1), 0, g.nz - 1);
        lim[k] = (qz * g.ny + qy) * g.nx + qx;
    }
    const bool reverted = outside_fade(g, bx + 0.5f, by + 0.5f, bz + 0.5f) < 1.f;
    const float4* orig[3] = {vel, sF, s2F};
    const float4* hf[3] = {vH, sH, s2H};
    const float4* torig[3] = {tvel, tS, tS2};
    const float4* th_[3] = {tvH, tsH, ts2H};
    float4* outs[3] = {tvY, tsY, ts2Y};
    for (int fi = 0; fi < 3; fi++) {
        const int nch = fi == 0 ? 3 : 4;
        const float4 hc = hf[fi][c.i], oc = orig[fi][c.i], thc = th_[fi][c.i], toc = torig[fi][c.i];
        float4 tv8[8], tt8[8], lv[8];
        tri_load(tf, hf[fi], tv8);
        tri_load(tf, th_[fi], tt8);
        for (int k = 0; k < 8; k++) lv[k] = orig[fi][lim[k]];
        float o[4] = {0.f, 0.f, 0.f, 0.f};
        for (int ch = 0; ch < nch; ch++) {
            float cv[8], ct = 0.f, hs, gx, gy, gz;
            for (int k = 0; k < 8; k++) { cv[k] = chan(tv8[k], ch); ct += tf.w[k] * chan(tt8[k], ch); }
            tri_eval(tf, cv, hs, gx, gy, gz);
            const float h = chan(hc, ch), oo = chan(oc, ch), thh = chan(thc, ch);
            float N = h + (oo - hs) * 0.5f;
            if (reverted) N = h;
            float mn = 1e9f, mx = -1e9f;
            int kmn = -1, kmx = -1;
            for (int k = 0; k < 8; k++) {
                const float t = chan(lv[k], ch);
                if (kmn < 0 || t < mn) { mn = t; kmn = k; }
                if (kmx < 0 || t > mx) { mx = t; kmx = k; }
            }
            const float lo = fminf(mn, h), hi = fmaxf(mx, h);
            const float out = wclamp(N, lo, hi);
            if (fi > 0 && !(out > 0.f)) continue;
            float t;
            if (N < lo) t = mn <= h ? chan(torig[fi][lim[kmn]], ch) : thh;
            else if (N > hi) t = mx >= h ? chan(torig[fi][lim[kmx]], ch) : thh;
            else if (reverted) t = thh;
            else t = thh + 0.5f * chan(toc, ch) - 0.5f * (ct + gx * tfw[0] + gy * tfw[1] + gz * tfw[2]);
            o[ch] = t;
        }
        outs[fi][c.i] = make_float4(o[0], o[1], o[2], o[3]);
    }
}

// curl (src/sim.js:200-211)
__global__ __launch_bounds__(256) void k_curl_tan(Dims g, const float4* __restrict__ v, const float4* __restrict__ tv,
                                                  float4* __restrict__ tvort) {
    Cell c;
    if (!cell_of(g, c)) return;
    const int iL = (c.z * g.ny + c.y) * g.nx + clampi(c.x - 1, 0, g.nx - 1), iR = (c.z * g.ny + c.y) * g.nx + clampi(c.x + 1, 0, g.nx - 1);
    const int iB = (c.z * g.ny + clampi(c.y - 1, 0, g.ny - 1)) * g.nx + c.x, iT = (c.z * g.ny + clampi(c.y + 1, 0, g.ny - 1)) * g.nx + c.x;
    const int iK = (clampi(c.z - 1, 0, g.nz - 1) * g.ny + c.y) * g.nx + c.x, iF = (clampi(c.z + 1, 0, g.nz - 1) * g.ny + c.y) * g.nx + c.x;
    const float4 L = v[iL], R = v[iR], B = v[iB], T = v[iT], K = v[iK], F = v[iF];
    const float4 tL = tv[iL], tR = tv[iR], tB = tv[iB], tT = tv[iT], tK = tv[iK], tF = tv[iF];
    const float wx = ((T.z - B.z) - (F.y - K.y)) * 0.5f;
    const float wy = ((F.x - K.x) - (R.z - L.z)) * 0.5f;
    const float wz = ((R.y - L.y) - (T.x - B.x)) * 0.5f;
    const float twx = ((tT.z - tB.z) - (tF.y - tK.y)) * 0.5f;
    const float twy = ((tF.x - tK.x) - (tR.z - tL.z)) * 0.5f;
    const float twz = ((tR.y - tL.y) - (tT.x - tB.x)) * 0.5f;
    const float wl = sqrtf(wx * wx + wy * wy + wz * wz);
    tvort[c.i] = make_float4(twx, twy, twz, wl > 0.f ? (wx * twx + wy * twy + wz * twz) / wl : 0.f);
}

// the source kernel (src/sim.js:214-396): one dual evaluation with the direction as tangents
__global__ __launch_bounds__(256) void k_forces_tan(const DevParams* __restrict__ P, Dims g, const float4* __restrict__ velY,
                                                    const float4* __restrict__ sY, const float4* __restrict__ s2Y,
                                                    const float4* __restrict__ vort, const float4* __restrict__ geo,
                                                    const half4_t* __restrict__ noise, const half4_t* __restrict__ curlT, int nn,
                                                    const fs_spot* __restrict__ spots, const float* __restrict__ faceFlux,
                                                    const float* __restrict__ spot_dir, const float* __restrict__ face_dir,
                                                    const float4* __restrict__ tvY, const float4* __restrict__ tsY,
                                                    const float4* __restrict__ ts2Y, const float4* __restrict__ tvort,
                                                    float4* __restrict__ tvF, float4* __restrict__ tS, float4* __restrict__ tS2) {
    Cell c;
    if (!cell_of(g, c)) return;
    const ForcesCell k = forces_setup<S>(P, g, c, velY, geo, noise, curlT, nn);
    const float4 v4 = velY[c.i], s4 = sY[c.i], t4 = s2Y[c.i], w4 = vort[c.i];
    const float4 dv = tvY[c.i], ds = tsY[c.i], dt2 = ts2Y[c.i], dw = tvort[c.i];
    int nb[6];
    nb[0] = (c.z * g.ny + c.y) * g.nx + clampi(c.x + 1, 0, g.nx - 1);
    nb[1] = (c.z * g.ny + c.y) * g.nx + clampi(c.x - 1, 0, g.nx - 1);
    nb[2] = (c.z * g.ny + clampi(c.y + 1, 0, g.ny - 1)) * g.nx + c.x;
    nb[3] = (c.z * g.ny + clampi(c.y - 1, 0, g.ny - 1)) * g.nx + c.x;
    nb[4] = (clampi(c.z + 1, 0, g.nz - 1) * g.ny + c.y) * g.nx + c.x;
    nb[5] = (clampi(c.z - 1, 0, g.nz - 1) * g.ny + c.y) * g.nx + c.x;
    Dual wn[6];
    for (int j = 0; j < 6; j++) wn[j] = Dual(vort[nb[j]].w, tvort[nb[j]].w);
    Dual vx(v4.x, dv.x), vy(v4.y, dv.y), vz(v4.z, dv.z), phi;
    Dual s[4] = {Dual(s4.x, ds.x), Dual(s4.y, ds.y), Dual(s4.z, ds.z), Dual(s4.w, ds.w)};
    Dual s2[4] = {Dual(t4.x, dt2.x), Dual(t4.y, dt2.y), Dual(t4.z, dt2.z), Dual(t4.w, dt2.w)};
    const Dual w[3] = {Dual(w4.x, dw.x), Dual(w4.y, dw.y), Dual(w4.z, dw.z)};
    const Dual gx = wn[0] - wn[1], gy = wn[2] - wn[3], gz = wn[4] - wn[5];
    // the Laplacian's tangent, in forces_setup's order of terms (+x, -x, +y, -y, +z, -z, then -6 v)
    Dual lap[3] = {Dual(k.lapx), Dual(k.lapy), Dual(k.lapz)};
    if (P->f[P_visc] > 0.f) {
        const float4 a = tvY[nb[0]], bb = tvY[nb[1]], cc = tvY[nb[2]], dd = tvY[nb[3]], ee = tvY[nb[4]], ff = tvY[nb[5]];
        lap[0].d = a.x + bb.x + cc.x + dd.x + ee.x + ff.x - dv.x * 6.f;
        lap[1].d = a.y + bb.y + cc.y + dd.y + ee.y + ff.y - dv.y * 6.f;
        lap[2].d = a.z + bb.z + cc.z + dd.z + ee.z + ff.z - dv.z * 6.f;
    }
    const bool face = P->has_face_flux != 0;
    const bool floor_face = face && c.y == 0;
    const Dual fv(floor_face ? faceFlux[c.z * g.nx + c.x] : 0.f, floor_face ? face_dir[c.z * g.nx + c.x] : 0.f);
    forces_cell<Dual>(P->f, P->d, k, spots, P->n_spots, -1, spot_dir, face, fv, vx, vy, vz, s, s2, w, gx, gy, gz, lap, phi);
    tvF[c.i] = make_float4(vx.d, vy.d, vz.d, phi.d);
    tS[c.i] = make_float4(s[0].d, s[1].d, s[2].d, s[3].d);
    tS2[c.i] = make_float4(s2[0].d, s2[1].d, s2[2].d, s2[3].d);
}

// div (src/sim.js:399-413) into the parity-split right-hand side
__global__ __launch_bounds__(256) void k_div_tan(const DevParams* __restrict__ P, Dims g, const float4* __restrict__ tv,
                                                 float* __restrict__ tb) {
    Cell c;
    if (!cell_of(g, c)) return;
    const float dx = P->f[P_dx];
    const float4 me = tv[c.i];
    const float L = tv[(c.z * g.ny + c.y) * g.nx + clampi(c.x - 1, 0, g.nx - 1)].x;
    const float R = tv[(c.z * g.ny + c.y) * g.nx + clampi(c.x + 1, 0, g.nx - 1)].x;
    const float B = c.y == 0 ? -me.y : tv[(c.z * g.ny + c.y - 1) * g.nx + c.x].y;
    const float T = tv[(c.z * g.ny + clampi(c.y + 1, 0, g.ny - 1)) * g.nx + c.x].y;
    const float K = tv[(clampi(c.z - 1, 0, g.nz - 1) * g.ny + c.y) * g.nx + c.x].z;
    const float F = tv[(clampi(c.z + 1, 0, g.nz - 1) * g.ny + c.y) * g.nx + c.x].z;
    tb[rb_index(c.x, c.y, c.z, g.nx, g.ny)] = ((R - L) + (T - B) + (F - K)) / (2.f * dx) - me.w;
}

// project (src/sim.js:526-550): p the step's pressure, tp its tangent (both FP64, parity-split)
__global__ __launch_bounds__(256) void k_project_tan(const DevParams* __restrict__ P, Dims g, const float4* __restrict__ vIn,
                                                     const double* __restrict__ p, const float4* __restrict__ geo,
                                                     const float4* __restrict__ tvIn, const double* __restrict__ tp,
                                                     float4* __restrict__ tvOut) {
    Cell c;
    if (!cell_of(g, c)) return;
    const float dx = P->f[P_dx];
    if (geo[c.i].z > 0.5f) { tvOut[c.i] = make_float4(0.f, 0.f, 0.f, 0.f); return; }
    auto grad = [&](const double* q, float& gx, float& gy, float& gz) {
        const int io = rb_index(c.x, c.y, c.z, g.nx, g.ny);
        const float me = (float)q[io];
        const float pxp = c.x + 1 < g.nx ? (float)q[rb_index(c.x + 1, c.y, c.z, g.nx, g.ny)] : -me;
        const float pxm = c.x > 0 ? (float)q[rb_index(c.x - 1, c.y, c.z, g.nx, g.ny)] : -me;
        const float pyp = c.y + 1 < g.ny ? (float)q[rb_index(c.x, c.y + 1, c.z, g.nx, g.ny)] : -me;
        const float pym = c.y > 0 ? (float)q[rb_index(c.x, c.y - 1, c.z, g.nx, g.ny)] : me;
        const float pzp = c.z + 1 < g.nz ? (float)q[rb_index(c.x, c.y, c.z + 1, g.nx, g.ny)] : -me;
        const float pzm = c.z > 0 ? (float)q[rb_index(c.x, c.y, c.z - 1, g.nx, g.ny)] : -me;
        gx = (pxp - pxm) / (2.f * dx); gy = (pyp - pym) / (2.f * dx); gz = (pzp - pzm) / (2.f * dx);
    };
    float gx, gy, gz, tgx, tgy, tgz;
    grad(p, gx, gy, gz);
    grad(tp, tgx, tgy, tgz);
    const float4 vi = vIn[c.i], ti = tvIn[c.i];
    float vx = vi.x - gx, vy = vi.y - gy, vz = vi.z - gz;
    float tx = ti.x - tgx, ty = ti.y - tgy, tz = ti.z - tgz;
    if (c.y == 0 && vy < 0.f) { vy = 0.f; ty = 0.f; }
    const float vl = sqrtf(vx * vx + vy * vy + vz * vz);
    if (vl > 20.f) {
        const float d = (tx * vx + ty * vy + tz * vz) / vl;
        const float sc = 20.f / vl;
        tx = sc * (tx - d * vx / vl); ty = sc * (ty - d * vy / vl); tz = sc * (tz - d * vz / vl);
    }
    tvOut[c.i] = make_float4(tx, ty, tz, 0.f);
}
}  // namespace

void launch_adv1_adj(const Dims& g, hipStream_t s, const DevParams* P, const float4* vel, const float4* sF,
                     const float4* s2F, const float4* lvX, const float4* lsX, const float4* ls2X, float4* lvel,
                     float4* ls, float4* ls2) {
    k_adv1_adj<<<g.nblocks, 256, 0, s>>>(P, g, vel, sF, s2F, lvX, lsX, ls2X, lvel, ls, ls2);
}
void launch_adv2_adj(const Dims& g, hipStream_t s, const DevParams* P, const float4* vel, const float4* sF,
                     const float4* s2F, const float4* vH, const float4* sH, const float4* s2H, const float4* lvY,
                     const float4* lsY, const float4* ls2Y, float4* lvel, float4* ls, float4* ls2, float4* lvH, float4* lsH,
                     float4* ls2H) {
    k_adv2_adj<<<g.nblocks, 256, 0, s>>>(P, g, vel, sF, s2F, vH, sH, s2H, lvY, lsY, ls2Y, lvel, ls, ls2, lvH, lsH, ls2H);
}
void launch_curl_adj(const Dims& g, hipStream_t s, const float4* v, const float4* lvort, float4* lv) {
    k_curl_adj<<<g.nblocks, 256, 0, s>>>(g, v, lvort, lv);
}
void launch_forces_adj(const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F, const float4* velY,
                       const float4* sY, const float4* s2Y, const float4* vort, const float4* lvOut, const float4* lsOut,
                       const float4* ls2Out, float4* lvY, float4* lsY, float4* ls2Y, float4* lvort, float4* llap,
                       bool visc, double* gspot, double* gface) {
    k_forces_adj<<<g.nblocks, 256, 0, s>>>(P, g, velY, sY, s2Y, vort, (const float4*)F.geo, F.noise, F.curl, F.noise_n,
                                           F.spots, F.face_flux, lvOut, lsOut, ls2Out, lvY, lsY, ls2Y, lvort, llap, gspot,
                                           gface);
    if (visc) k_visc_adj<<<g.nblocks, 256, 0, s>>>(g, llap, lvY);
}
void launch_div_adj(const Dims& g, hipStream_t s, const DevParams* P, const double* lb, float4* lvF) {
    k_div_adj<double><<<g.nblocks, 256, 0, s>>>(P, g, lb, lvF);
}
void launch_project_adj(const Dims& g, hipStream_t s, const DevParams* P, const float4* vIn, const double* p,
                        const float4* geo, const float4* lvOut, float4* lvIn, float* lp) {
    k_project_adj<double><<<g.nblocks, 256, 0, s>>>(P, g, vIn, p, geo, lvOut, lvIn, lp);
}
void launch_loss_seed(const Dims& g, hipStream_t s, const float4* Sf, const float* tT, const float* tSoot, float4* lS,
                      double* partial) {
    k_loss_seed<<<g.nblocks, 256, 0, s>>>(g, Sf, tT, tSoot, lS, partial);
}

void launch_adv1_tan(const Dims& g, hipStream_t s, const DevParams* P, const float4* vel, const float4* sF,
                     const float4* s2F, const float4* tvel, const float4* tS, const float4* tS2, float4* tvX, float4* tsX,
                     float4* ts2X) {
    k_adv1_tan<<<g.nblocks, 256, 0, s>>>(P, g, vel, sF, s2F, tvel, tS, tS2, tvX, tsX, ts2X);
}
void launch_adv2_tan(const Dims& g, hipStream_t s, const DevParams* P, const float4* vel, const float4* sF,
                     const float4* s2F, const float4* vH, const float4* sH, const float4* s2H, const float4* tvel,
                     const float4* tS, const float4* tS2, const float4* tvH, const float4* tsH, const float4* ts2H,
                     float4* tvY, float4* tsY, float4* ts2Y) {
    k_adv2_tan<<<g.nblocks, 256, 0, s>>>(P, g, vel, sF, s2F, vH, sH, s2H, tvel, tS, tS2, tvH, tsH, ts2H, tvY, tsY, ts2Y);
}
void launch_curl_tan(const Dims& g, hipStream_t s, const float4* v, const float4* tv, float4* tvort) {
    k_curl_tan<<<g.nblocks, 256, 0, s>>>(g, v, tv, tvort);
}
void launch_forces_tan(const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F, const float4* velY,
                       const float4* sY, const float4* s2Y, const float4* vort, const float* spot_dir,
                       const float* face_dir, const float4* tvY, const float4* tsY, const float4* ts2Y,
                       const float4* tvort, float4* tvF, float4* tS, float4* tS2) {
    k_forces_tan<<<g.nblocks, 256, 0, s>>>(P, g, velY, sY, s2Y, vort, (const float4*)F.geo, F.noise, F.curl, F.noise_n,
                                           F.spots, F.face_flux, spot_dir, face_dir, tvY, tsY, ts2Y, tvort, tvF, tS, tS2);
}
void launch_div_tan(const Dims& g, hipStream_t s, const DevParams* P, const float4* tv, float* tb) {
    k_div_tan<<<g.nblocks, 256, 0, s>>>(P, g, tv, tb);
}
void launch_project_tan(const Dims& g, hipStream_t s, const DevParams* P, const float4* vIn, const double* p,
                        const float4* geo, const float4* tvIn, const double* tp, float4* tvOut) {
    k_project_tan<<<g.nblocks, 256, 0, s>>>(P, g, vIn, p, geo, tvIn, tp, tvOut);
}
