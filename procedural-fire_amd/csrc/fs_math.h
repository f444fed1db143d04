// Device helpers shared by the step, diagnostics and adjoint kernels: WGSL built-in
// semantics, the textbook exp/log/pow (bit-identical to the oracle's), the tiling of
// the per-cell kernels, clamped texel loads and trilinear sampling, and the
// semi-Lagrangian backtrace (src/sim.js:123-140).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "fs_codec.h"
#include "fs_divc.h"
#include "fs_internal.h"

// ---- WGSL built-in semantics (clamp = min(max()), smoothstep, mix) -------------
__device__ __forceinline__ float wclamp(float e, float lo, float hi) { return fminf(fmaxf(e, lo), hi); }
__device__ __forceinline__ float sstep(float e0, float e1, float x) {
    float t = wclamp((x - e0) / (e1 - e0), 0.f, 1.f);
    return t * t * (3.f - 2.f * t);
}
// smoothstep whose edges e0 < e1 are literals: bit-identical to sstep.  The numerator is clamped to
// [-d, 2d] first (t is the same for every value outside it, infinities and NaN included), then divided
// by d = e1 - e0 through its reciprocal, which folds at compile time.
__device__ __forceinline__ float sstep_c(float e0, float e1, float x) {
    const float d = e1 - e0;
    const float n = fminf(fmaxf(x - e0, -d), 2.f * d);
    const float t = wclamp(fs_div_recip(n, d, 1.f / d), 0.f, 1.f);
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
//        |r| <= ln2/2 in Horner form with fused multiply-adds, then ldexp(., k).
//   log: frexp to m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
//        log m = 2s(1 + s^2/3 + s^4/5 + s^6/7 + s^8/9) + e ln2, the same way.
//   pow(a, b) = exp(b log a) for finite a > 0.
// exp and log take their special cases as selects after the polynomial (every lane runs the same instructions).
__device__ __forceinline__ float fs_exp(float x) {
    const float k = rintf(x * 1.44269504f);
    const float r = (x - k * 0.693145751953125f) - k * 1.42860682e-6f;
    float p = 1.0f / 5040.0f;
    p = fmaf(r, p, 1.0f / 720.0f);
    p = fmaf(r, p, 1.0f / 120.0f);
    p = fmaf(r, p, 1.0f / 24.0f);
    p = fmaf(r, p, 1.0f / 6.0f);
    p = fmaf(r, p, 0.5f);
    p = fmaf(r, p, 1.0f);
    p = fmaf(r, p, 1.0f);
    const float e = ldexpf(p, (int)fminf(fmaxf(k, -256.0f), 256.0f));   // k is within +-151 wherever e is used
    return x != x ? x : x > 88.72283f ? __int_as_float(0x7f800000) : x < -103.97208f ? 0.0f : e;
}
__device__ __forceinline__ float fs_log(float a) {
    int e;
    float m = frexpf(a, &e);
    if (m < 0.70710678f) { m = m * 2.0f; e = e - 1; }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float q = 1.0f / 9.0f;
    q = fmaf(s2, q, 1.0f / 7.0f);
    q = fmaf(s2, q, 1.0f / 5.0f);
    q = fmaf(s2, q, 1.0f / 3.0f);
    const float t = 2.0f * s;
    const float fe = (float)e;
    const float l = fmaf(fe, 0.693145751953125f, fmaf(fe, 1.42860682e-6f, fmaf(t, s2 * q, t)));
    return (a != a || a < 0.0f) ? __int_as_float(0x7fc00000)
         : a == 0.0f ? -__int_as_float(0x7f800000)
         : a == __int_as_float(0x7f800000) ? a : l;
}
// pow keeps its early exits: where a wave's every lane has a = 1 (the source kernel's emitter terms away from the
// sources) it skips the log and the exp
__device__ __forceinline__ float fs_pow(float a, float b) {
    if (a == 1.0f || b == 0.0f) return 1.0f;
    if (a == 0.0f) return b > 0.0f ? 0.0f : __int_as_float(0x7f800000);
    return fs_exp(b * fs_log(a));
}

// sin / cos (WGSL leaves them to the implementation too): Cody-Waite reduction by
// k = rint(x 2/pi) with pi/2 split into 12 + 12 + 24 significant bits (k*P1, k*P2 exact
// for |k| < 2048), Taylor polynomials of degree 9 (sin) and 10 (cos) on |r| <= pi/4,
// quadrant from k mod 4.  Identical to the oracle's.
__device__ __forceinline__ void fs_sincos_red(float x, float& sr, float& cr, int& q) {
    const float k = rintf(x * 0.636619747f);
    const float r = ((x - k * 1.5703125f) - k * 4.83751297e-4f) - k * 7.54978995e-8f;
    const float r2 = r * r;
    float ps = 1.0f / 362880.0f;
    ps = -1.0f / 5040.0f + r2 * ps;
    ps = 1.0f / 120.0f + r2 * ps;
    ps = -1.0f / 6.0f + r2 * ps;
    sr = r + r * (r2 * ps);
    float pc = -1.0f / 3628800.0f;
    pc = 1.0f / 40320.0f + r2 * pc;
    pc = -1.0f / 720.0f + r2 * pc;
    pc = 1.0f / 24.0f + r2 * pc;
    pc = -0.5f + r2 * pc;
    cr = 1.0f + r2 * pc;
    q = ((int)k) & 3;
}
__device__ __forceinline__ float fs_sin(float x) {
    float s, c;
    int q;
    fs_sincos_red(x, s, c, q);
    return q == 0 ? s : q == 1 ? c : q == 2 ? -s : -c;
}
__device__ __forceinline__ float fs_cos(float x) {
    float s, c;
    int q;
    fs_sincos_red(x, s, c, q);
    return q == 0 ? c : q == 1 ? -s : q == 2 ? -c : s;
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
// ---- field readers: texel i of a field through a global pointer, or through a raw buffer resource.
// FBuf keeps the base and size in scalar registers and each load takes a 32-bit byte offset, so an
// offset formed once per texel position serves every field read there, with no 64-bit address
// arithmetic per load.  A field read this way stays below 4 GiB (checked when a sim is created).
// With 8-byte texels (fp16, bf16) FBuf also reads two x-neighbours in one 16-byte load: the vector
// memory path spends about the same cycles on a wave's 16-byte load as on its 8-byte load, and the
// gather-heavy kernels are bound there.
template <class S> struct FPtr {
    static constexpr bool kPair = false;
    const typename S::raw* p;
    __device__ __forceinline__ FPtr(const typename S::raw* q, const Dims&) : p(q) {}
    __device__ __forceinline__ float4 ld(int i) const { return Codec<S>::ld(p, i); }
};
// ---- binary16 operands straight from the raw words (gfx950 VOP3P) ----------------
// v_fma_mix_f32 with an f16 first operand computes h * s + c in f32 with a single rounding; with c = -0
// that is the rounded product of the (exactly) widened h and s, the same bits as converting h first and
// multiplying (x + -0 = x for every x, -0 included).  So the first products of a blend absorb the
// widening of their texel channels.
__device__ __forceinline__ float hmul_lo(unsigned w, float s) {   // f32(low half of w) * s
    float r;
    asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(w), "v"(s), "v"(__int_as_float(0x80000000)));
    return r;
}
__device__ __forceinline__ float hmul_hi(unsigned w, float s) {   // f32(high half of w) * s
    float r;
    asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(w), "v"(s), "v"(__int_as_float(0x80000000)));
    return r;
}
__device__ __forceinline__ float hfma_lo(unsigned w, float s, float c) {   // fma(f32(low half of w), s, c)
    float r;
    asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(w), "v"(s), "v"(c));
    return r;
}
__device__ __forceinline__ float hfma_hi(unsigned w, float s, float c) {   // fma(f32(high half of w), s, c)
    float r;
    asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(w), "v"(s), "v"(c));
    return r;
}

// ---- trilinear filtering ------------------------------------------------------------
// Corner k = (a, b, c) = (k & 1, (k >> 1) & 1, k >> 2) of the 2x2x2 texel block has the weight
// w_k = wx_a * (wy_b * wz_c), with wx_0 = 1 - fx, wx_1 = fx (and so on), and each channel is the
// chain v = t_0 w_0, then v = fma(t_k, w_k, v) for k = 1..7.  The weights are formed once per sample
// position and serve every field read there; a channel costs 8 VALU instead of the 21 of three
// levels of two-product blends.  The oracle's samplers round the same way (the WebGPU sampler's own
// fixed-point filter is unpinned either way).  At a texel centre the weights are exactly 1 and 0.
__device__ __forceinline__ void tri_weights(float fx, float fy, float fz, float* w) {
    const float ux = 1.f - fx, uy = 1.f - fy, uz = 1.f - fz;
    const float yz[4] = {uy * uz, fy * uz, uy * fz, fy * fz};
#pragma unroll
    for (int j = 0; j < 4; j++) {
        w[2 * j] = ux * yz[j];
        w[2 * j + 1] = fx * yz[j];
    }
}
__device__ __forceinline__ float4 tri_blend(const float4* t, const float* w) {
    float4 r = make_float4(t[0].x * w[0], t[0].y * w[0], t[0].z * w[0], t[0].w * w[0]);
#pragma unroll
    for (int k = 1; k < 8; k++) {
        r.x = __builtin_fmaf(t[k].x, w[k], r.x);
        r.y = __builtin_fmaf(t[k].y, w[k], r.y);
        r.z = __builtin_fmaf(t[k].z, w[k], r.z);
        r.w = __builtin_fmaf(t[k].w, w[k], r.w);
    }
    return r;
}
// the same for binary16 texels as raw words: a[j] is corner 2j, b[j] corner 2j + 1; the widening of each
// half rides in v_fma_mix_f32 (exact, so the bits are tri_blend's of the decoded texels)
__device__ __forceinline__ float4 tri_blend_h(const u32x2_t* a, const u32x2_t* b, const float* w) {
    float4 r = make_float4(hmul_lo(a[0].x, w[0]), hmul_hi(a[0].x, w[0]), hmul_lo(a[0].y, w[0]), hmul_hi(a[0].y, w[0]));
    r.x = hfma_lo(b[0].x, w[1], r.x); r.y = hfma_hi(b[0].x, w[1], r.y);
    r.z = hfma_lo(b[0].y, w[1], r.z); r.w = hfma_hi(b[0].y, w[1], r.w);
#pragma unroll
    for (int j = 1; j < 4; j++) {
        r.x = hfma_lo(a[j].x, w[2 * j], r.x); r.y = hfma_hi(a[j].x, w[2 * j], r.y);
        r.z = hfma_lo(a[j].y, w[2 * j], r.z); r.w = hfma_hi(a[j].y, w[2 * j], r.w);
        r.x = hfma_lo(b[j].x, w[2 * j + 1], r.x); r.y = hfma_hi(b[j].x, w[2 * j + 1], r.y);
        r.z = hfma_lo(b[j].y, w[2 * j + 1], r.z); r.w = hfma_hi(b[j].y, w[2 * j + 1], r.w);
    }
    return r;
}
// Per-half minimum / maximum of two binary16 pairs (IEEE minNum / maxNum, as fminf / fmaxf on the widened
// values; widening is exact and monotonic, so min-then-widen equals widen-then-min).  The operands are
// taken as they are: a signalling NaN would come out quiet here instead of being ignored, and the fields
// never hold one (every binary16 store converts from f32, which only produces quiet NaNs).
__device__ __forceinline__ unsigned pk_min_h(unsigned a, unsigned b) {
    unsigned r;
    asm("v_pk_min_f16 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ unsigned pk_max_h(unsigned a, unsigned b) {
    unsigned r;
    asm("v_pk_max_f16 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

template <class S> struct FBuf {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ FBuf(const typename S::raw* p, const Dims& g)
        : r(__builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0,
                                              (int)((unsigned)g.nx * (unsigned)g.ny * (unsigned)g.nz * (unsigned)sizeof(typename S::raw)),
                                              FS_BUF_WORD3)) {}
    __device__ __forceinline__ float4 ld(int i) const {
        return Codec<S>::ldb(r, (int)((unsigned)i * (unsigned)sizeof(typename S::raw)));
    }
    static constexpr bool kPair = sizeof(typename S::raw) == 8;
    // The clamped texels x0 = clamp(ix, 0, n-1), x1 = clamp(ix+1, 0, n-1) of one row, from texels i and
    // i + 1 where i is the row start + clamp(ix, 0, n-2) (n >= 2): both are the first texel when ix < 0
    // (lo_both) and both the second when ix >= n - 1 (hi_both).  The selects act on the raw words.
    __device__ __forceinline__ void ld_pair_raw(int i, bool lo_both, bool hi_both, u32x2_t& a, u32x2_t& b) const {
        const u32x4_t u = __builtin_amdgcn_raw_buffer_load_b128(r, (int)((unsigned)i * 8u), 0, 0);
        const u32x2_t lo = {u.x, u.y}, hi = {u.z, u.w};
        a = hi_both ? hi : lo;
        b = lo_both ? lo : hi;
    }
    // texels i and i + 1 as they are (raw words)
    __device__ __forceinline__ void ld_pair_raw(int i, u32x2_t& a, u32x2_t& b) const {
        const u32x4_t u = __builtin_amdgcn_raw_buffer_load_b128(r, (int)((unsigned)i * 8u), 0, 0);
        a = u32x2_t{u.x, u.y};
        b = u32x2_t{u.z, u.w};
    }
    __device__ __forceinline__ void ld_pair(int i, float4& a, float4& b) const {
        u32x2_t ra, rb;
        ld_pair_raw(i, ra, rb);
        a = Codec<S>::dec(ra);
        b = Codec<S>::dec(rb);
    }
    __device__ __forceinline__ void ld_pair(int i, bool lo_both, bool hi_both, float4& a, float4& b) const {
        u32x2_t ra, rb;
        ld_pair_raw(i, lo_both, hi_both, ra, rb);
        a = Codec<S>::dec(ra);
        b = Codec<S>::dec(rb);
    }
};

// clamped texel load (ld(), src/sim.js:114)
template <class S, class F>
__device__ __forceinline__ float4 ldc_f(const F& f, const Dims& g, int x, int y, int z) {
    x = clampi(x, 0, g.nx - 1); y = clampi(y, 0, g.ny - 1); z = clampi(z, 0, g.nz - 1);
    return f.ld((z * g.ny + y) * g.nx + x);
}
template <class S>
__device__ __forceinline__ float4 ldc(const typename S::raw* f, const Dims& g, int x, int y, int z) {
    return ldc_f<S>(FPtr<S>(f, g), g, x, y, z);
}
// One axis of the clamp-to-edge footprint at texel-space coordinate t (n >= 2): the texels i0 and i0 + 1 with
// i0 = clamp(floor(t), 0, n - 2), and the weight fraction f of the second.  Where floor(t) lies outside
// [0, n - 2] both clamped texels of the footprint are the same edge texel; the fraction 0 (below) or 1
// (above) then takes that texel exactly, so the footprint never needs two copies of one texel.
__device__ __forceinline__ int tri_axis(float t, int n, float& f) {
    const float fl = floorf(t);
    const int i = (int)fl;
    f = i < 0 ? 0.f : i > n - 2 ? 1.f : t - fl;
    return clampi(i, 0, n - 2);
}
// trilinear, clamp-to-edge, at cell-space position b (texel centres at i + 0.5).
// Clamping the coordinate to [-1, n] first leaves the clamped-texel result unchanged.
template <class S, class F>
__device__ __forceinline__ float4 sample_clamp_f(const F& f, const Dims& g, float bx, float by, float bz) {
    const float tx = wclamp(bx - 0.5f, -1.f, (float)g.nx);
    const float ty = wclamp(by - 0.5f, -1.f, (float)g.ny);
    const float tz = wclamp(bz - 0.5f, -1.f, (float)g.nz);
    float fx, fy, fz;
    const int xs = tri_axis(tx, g.nx, fx), ys = tri_axis(ty, g.ny, fy), zs = tri_axis(tz, g.nz, fz);
    const int r0 = (zs * g.ny + ys) * g.nx + xs;
    const int rows[4] = {r0, r0 + g.nx, r0 + g.ny * g.nx, r0 + (g.ny + 1) * g.nx};
    float w[8];
    tri_weights(fx, fy, fz, w);
    if constexpr (F::kPair && std::is_same<S, StF16>::value) {
        // binary16 pairs: the blend reads the raw halves (tri_blend_h)
        u32x2_t a[4], b[4];
#pragma unroll
        for (int j = 0; j < 4; j++) f.ld_pair_raw(rows[j], a[j], b[j]);
        return tri_blend_h(a, b, w);
    } else {
        float4 c[8];
        if constexpr (F::kPair) {
#pragma unroll
            for (int j = 0; j < 4; j++) f.ld_pair(rows[j], c[2 * j], c[2 * j + 1]);
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                c[2 * j] = f.ld(rows[j]);
                c[2 * j + 1] = f.ld(rows[j] + 1);
            }
        }
        return tri_blend(c, w);
    }
}
// FBuf where it pays in every kernel (8-byte texels, paired); 16-byte texels keep plain pointer reads
template <class S> using FPair = typename std::conditional<sizeof(typename S::raw) == 8, FBuf<S>, FPtr<S>>::type;

// the 8 clamped texels at (b0x + (k & 1), b0y + ((k >> 1) & 1), b0z + (k >> 2)), k = 0..7 (ld(), src/sim.js:114)
template <class S, class F>
__device__ __forceinline__ void corners8_f(const F& f, const Dims& g, int b0x, int b0y, int b0z, float4* c) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int y = clampi(b0y + (j & 1), 0, g.ny - 1), z = clampi(b0z + (j >> 1), 0, g.nz - 1);
        const int row = (z * g.ny + y) * g.nx;
        if constexpr (F::kPair) {
            f.ld_pair(row + clampi(b0x, 0, g.nx - 2), b0x < 0, b0x >= g.nx - 1, c[2 * j], c[2 * j + 1]);
        } else {
            c[2 * j] = f.ld(row + clampi(b0x, 0, g.nx - 1));
            c[2 * j + 1] = f.ld(row + clampi(b0x + 1, 0, g.nx - 1));
        }
    }
}
// The per-channel minimum and maximum over the same 8 texels of a binary16 field read in pairs, folded into
// lo / hi starting values (fminf(lo, .) and fmaxf(hi, .) over the texels in any order give the same value:
// minNum / maxNum of a set; all-NaN texels leave lo / hi).  Packed binary16 min / max, widened once.
template <class F>
__device__ __forceinline__ void minmax8_h(const F& f, const Dims& g, int b0x, int b0y, int b0z, float4 lo, float4 hi,
                                          float4& mn, float4& mx) {
    unsigned m0 = 0, m1 = 0, M0 = 0, M1 = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int y = clampi(b0y + (j & 1), 0, g.ny - 1), z = clampi(b0z + (j >> 1), 0, g.nz - 1);
        const int row = (z * g.ny + y) * g.nx;
        u32x2_t a, b;
        f.ld_pair_raw(row + clampi(b0x, 0, g.nx - 2), b0x < 0, b0x >= g.nx - 1, a, b);
        const unsigned n0 = pk_min_h(a.x, b.x), n1 = pk_min_h(a.y, b.y);
        const unsigned x0 = pk_max_h(a.x, b.x), x1 = pk_max_h(a.y, b.y);
        if (j == 0) { m0 = n0; m1 = n1; M0 = x0; M1 = x1; }
        else { m0 = pk_min_h(m0, n0); m1 = pk_min_h(m1, n1); M0 = pk_max_h(M0, x0); M1 = pk_max_h(M1, x1); }
    }
    const float4 a = Codec<StF16>::dec(u32x2_t{m0, m1}), b = Codec<StF16>::dec(u32x2_t{M0, M1});
    mn = make_float4(fminf(lo.x, a.x), fminf(lo.y, a.y), fminf(lo.z, a.z), fminf(lo.w, a.w));
    mx = make_float4(fmaxf(hi.x, b.x), fmaxf(hi.y, b.y), fmaxf(hi.z, b.z), fmaxf(hi.w, b.w));
}
template <class S>
__device__ __forceinline__ float4 sample_clamp(const typename S::raw* f, const Dims& g, float bx, float by, float bz) {
    return sample_clamp_f<S>(FPtr<S>(f, g), g, bx, by, bz);
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
    const int rows[4] = {(z0 * n + y0) * n, (z0 * n + y1) * n, (z1 * n + y0) * n, (z1 * n + y1) * n};
    float4 c[8];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        c[2 * j] = ld_h4(t, rows[j] + x0);
        c[2 * j + 1] = ld_h4(t, rows[j] + x1);
    }
    float w[8];
    tri_weights(fx, fy, fz, w);
    return tri_blend(c, w);
}
// The same sample from an n^3 binary16 texture read through a buffer resource: each row's texels x0, x0 + 1
// come in one 16-byte load.  Where x0 = n - 1 wraps to x1 = 0, the second half of that load (the next row's
// first texel, or zeros past the end of the buffer, whose range check returns 0) is replaced by a second load,
// behind one branch that waves with no wrapping lane skip.  Bit-identical to sample_rep.
struct RepTex {
    __amdgpu_buffer_rsrc_t r;
    int n;
    __device__ __forceinline__ RepTex(const half4_t* t, int nn)
        : r(__builtin_amdgcn_make_buffer_rsrc((void*)t, (short)0, (int)((unsigned)nn * (unsigned)nn * (unsigned)nn * 8u),
                                              FS_BUF_WORD3)),
          n(nn) {}
};
__device__ __forceinline__ float4 sample_rep_b(const RepTex& t, float qx, float qy, float qz) {
    const int n = t.n;
    const float tx = qx * (float)n - 0.5f, ty = qy * (float)n - 0.5f, tz = qz * (float)n - 0.5f;
    const float flx = floorf(tx), fly = floorf(ty), flz = floorf(tz);
    const int ix = (int)flx, iy = (int)fly, iz = (int)flz;
    const float fx = tx - flx, fy = ty - fly, fz = tz - flz;
    const int m = n - 1;
    const int x0 = ix & m, y0 = iy & m, y1 = (iy + 1) & m, z0 = iz & m, z1 = (iz + 1) & m;
    const int rows[4] = {(z0 * n + y0) * n, (z0 * n + y1) * n, (z1 * n + y0) * n, (z1 * n + y1) * n};
    u32x2_t a[4], b[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const u32x4_t u = __builtin_amdgcn_raw_buffer_load_b128(t.r, (int)((unsigned)(rows[j] + x0) * 8u), 0, 0);
        a[j] = u32x2_t{u.x, u.y};
        b[j] = u32x2_t{u.z, u.w};
    }
    if (x0 == m) {
#pragma unroll
        for (int j = 0; j < 4; j++) b[j] = __builtin_amdgcn_raw_buffer_load_b64(t.r, (int)((unsigned)rows[j] * 8u), 0, 0);
    }
    float w[8];
    tri_weights(fx, fy, fz, w);
    return tri_blend_h(a, b, w);
}

template <class S>
__device__ __forceinline__ float3 backtrace(const typename S::raw* vel, const Dims& g, float cx, float cy, float cz,
                                            float dt, float dx) {   // sim.js:130-135
    const float4 v0 = sample_clamp<S>(vel, g, cx, cy, cz);
    const float hx = cx - v0.x * (0.5f * dt) / dx, hy = cy - v0.y * (0.5f * dt) / dx, hz = cz - v0.z * (0.5f * dt) / dx;
    const float4 v1 = sample_clamp<S>(vel, g, hx, hy, hz);
    return make_float3(cx - v1.x * dt / dx, cy - v1.y * dt / dx, cz - v1.z * dt / dx);
}
// The same backtrace given the cell's own velocity texel v0 (cx, cy, cz the cell centre).  At a centre the
// trilinear weights are exactly 1 and 0, so for finite texels the first sample is v0 up to the sign of a
// zero channel, and a zero only enters as c - (+-0) = c: the result is bit-identical to backtrace() and
// skips eight texel loads and their blends.
template <class S, class F>
__device__ __forceinline__ float3 backtrace_c(const F& vel, const Dims& g, float cx, float cy, float cz,
                                              float dt, float dx, float4 v0) {
    const float hx = cx - v0.x * (0.5f * dt) / dx, hy = cy - v0.y * (0.5f * dt) / dx, hz = cz - v0.z * (0.5f * dt) / dx;
    const float4 v1 = sample_clamp_f<S>(vel, g, hx, hy, hz);
    return make_float3(cx - v1.x * dt / dx, cy - v1.y * dt / dx, cz - v1.z * dt / dx);
}
__device__ __forceinline__ float outside_fade(const Dims& g, float x, float y, float z) {   // sim.js:137-140
    const float lo = fminf(x, z);
    const float hi = fmaxf(x - (float)g.nx, fmaxf(y - (float)g.ny, z - (float)g.nz));
    return wclamp(1.f - fmaxf(-lo, hi), 0.f, 1.f);
}

__device__ __forceinline__ float clamp_lim(float v, float mn, float mx, float h) { return wclamp(v, fminf(mn, h), fmaxf(mx, h)); }
