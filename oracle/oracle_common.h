// TEST INFRASTRUCTURE ONLY — the CPU oracle of the fire-simulation step.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
// and only as the checker / the timed CPU baseline, never as the product path.
//
// Scalar helpers restating WGSL built-in semantics used by src/sim.js, src/noise.js
// and src/shaders/lib.js: clamp = min(max()), smoothstep, mix, RNE stores to
// rgba16float / bf16, PCG3D hashing.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

// ---- storage encodings ------------------------------------------------------
// IEEE binary16 with round-to-nearest-even, overflow to +-inf, NaN kept as NaN
// (what a textureStore to rgba16float does).
static inline uint16_t f32_to_f16(float f) {
    uint32_t x; std::memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ex = (x >> 23) & 0xffu;
    uint32_t man = x & 0x7fffffu;
    if (ex == 0xffu) return (uint16_t)(sign | 0x7c00u | (man ? 0x200u : 0u));
    int e = (int)ex - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        man |= 0x800000u;
        int shift = 14 - e;                 // 24-bit significand -> subnormal
        uint32_t half = man >> shift;
        uint32_t rem = man & ((1u << shift) - 1u);
        uint32_t mid = 1u << (shift - 1);
        if (rem > mid || (rem == mid && (half & 1u))) half++;
        return (uint16_t)(sign | half);
    }
    uint32_t h = sign | ((uint32_t)e << 10) | (man >> 13);
    uint32_t rem = man & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;   // carries into the exponent correctly
    return (uint16_t)h;
}
static inline float f16_to_f32(uint16_t h) {
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t ex = (h >> 10) & 0x1fu, man = h & 0x3ffu;
    uint32_t x;
    if (ex == 0) {
        if (man == 0) x = sign;
        else {  // subnormal
            float v = std::ldexp((float)man, -24);
            std::memcpy(&x, &v, 4); x |= sign;
        }
    } else if (ex == 31) x = sign | 0x7f800000u | (man << 13);
    else x = sign | ((ex - 15 + 127) << 23) | (man << 13);
    float f; std::memcpy(&f, &x, 4); return f;
}
// bfloat16, round-to-nearest-even on the upper 16 bits; NaN -> quiet NaN.
static inline uint16_t f32_to_bf16(float f) {
    uint32_t x; std::memcpy(&x, &f, 4);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
    x += 0x7fffu + ((x >> 16) & 1u);
    return (uint16_t)(x >> 16);
}
static inline float bf16_to_f32(uint16_t h) {
    uint32_t x = (uint32_t)h << 16; float f; std::memcpy(&f, &x, 4); return f;
}

struct half_t { uint16_t u; };
struct bf16_t { uint16_t u; };

template <class S> struct Store;
template <> struct Store<float>  { static double dec(float v) { return v; }
                                   static float enc(double v) { return (float)v; } };
template <> struct Store<double> { static double dec(double v) { return v; }
                                   static double enc(double v) { return v; } };
template <> struct Store<half_t> { static double dec(half_t v) { return f16_to_f32(v.u); }
                                   static half_t enc(double v) { return half_t{f32_to_f16((float)v)}; } };
template <> struct Store<bf16_t> { static double dec(bf16_t v) { return bf16_to_f32(v.u); }
                                   static bf16_t enc(double v) { return bf16_t{f32_to_bf16((float)v)}; } };

// ---- WGSL scalar built-ins ---------------------------------------------------
// clamp(e, lo, hi) = min(max(e, lo), hi) with IEEE maxNum/minNum (a NaN operand
// yields the other operand), as D3D/Vulkan min/max do.
template <class R> static inline R wclamp(R e, R lo, R hi) { return std::fmin(std::fmax(e, lo), hi); }
template <class R> static inline R wmin(R a, R b) { return std::fmin(a, b); }
template <class R> static inline R wmax(R a, R b) { return std::fmax(a, b); }
template <class R> static inline R smoothstep(R e0, R e1, R x) {
    R t = wclamp<R>((x - e0) / (e1 - e0), R(0), R(1));
    return t * t * (R(3) - R(2) * t);
}
// mix(a, b, t) = a * (1 - t) + b * t (WGSL spec form)
template <class R> static inline R mix(R a, R b, R t) { return a * (R(1) - t) + b * t; }

// ---- exp / log / pow in f32 -------------------------------------------------
// WGSL leaves exp and pow to the implementation.  The oracle and the HIP kernels
// both use this textbook construction, so they round identically:
//   exp: Cody-Waite reduction x = k ln2 + r (ln2 split hi/lo, hi with trailing
//        zero bits (9) so k*hi is exact), |r| <= ln2/2, degree-7 Taylor polynomial in
//        Horner form, scaled by ldexp(., k);  error < 1 ulp before scaling.
//   log: frexp to m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
//        log m = 2 s (1 + s^2/3 + s^4/5 + s^6/7 + s^8/9), plus e ln2.
//   pow(a, b) = exp(b log a) for finite a > 0.
// Evaluated in f32: the Horner steps and log's closing sum as fused multiply-adds (std::fma, one rounding each),
// the other operations rounding separately (no contraction).
static inline float fexp(float x) {
    if (x != x) return x;
    if (x > 88.72283f) return INFINITY;
    if (x < -103.97208f) return 0.0f;
    const float k = std::rint(x * 1.44269504f);
    const float r = (x - k * 0.693145751953125f) - k * 1.42860682e-6f;
    float p = 1.0f / 5040.0f;
    p = std::fma(r, p, 1.0f / 720.0f);
    p = std::fma(r, p, 1.0f / 120.0f);
    p = std::fma(r, p, 1.0f / 24.0f);
    p = std::fma(r, p, 1.0f / 6.0f);
    p = std::fma(r, p, 0.5f);
    p = std::fma(r, p, 1.0f);
    p = std::fma(r, p, 1.0f);
    return std::ldexp(p, (int)k);
}
static inline float flog(float a) {
    if (a != a || a < 0.0f) return NAN;
    if (a == 0.0f) return -INFINITY;
    if (a == INFINITY) return INFINITY;
    int e;
    float m = std::frexp(a, &e);
    if (m < 0.70710678f) { m = m * 2.0f; e = e - 1; }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float q = 1.0f / 9.0f;
    q = std::fma(s2, q, 1.0f / 7.0f);
    q = std::fma(s2, q, 1.0f / 5.0f);
    q = std::fma(s2, q, 1.0f / 3.0f);
    const float t = 2.0f * s;
    const float fe = (float)e;
    return std::fma(fe, 0.693145751953125f, std::fma(fe, 1.42860682e-6f, std::fma(t, s2 * q, t)));
}
static inline float fpow(float a, float b) {
    if (a == 1.0f || b == 0.0f) return 1.0f;
    if (a == 0.0f) return b > 0.0f ? 0.0f : INFINITY;
    return fexp(b * flog(a));
}
// sin / cos: Cody-Waite reduction by k = rint(x 2/pi), pi/2 split into 12 + 12 + 24
// significant bits, Taylor polynomials of degree 9 (sin) and 10 (cos) on |r| <= pi/4.
static inline void fsincos_red(float x, float& sr, float& cr, int& q) {
    const float k = std::rint(x * 0.636619747f);
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
static inline float fsin(float x) {
    float s, c;
    int q;
    fsincos_red(x, s, c, q);
    return q == 0 ? s : q == 1 ? c : q == 2 ? -s : -c;
}
static inline float fcos(float x) {
    float s, c;
    int q;
    fsincos_red(x, s, c, q);
    return q == 0 ? c : q == 1 ? -s : q == 2 ? -c : s;
}
template <class R> static inline R wexp(R x);
template <> inline float wexp<float>(float x) { return fexp(x); }
template <> inline double wexp<double>(double x) { return std::exp(x); }
template <class R> static inline R wpow(R a, R b);
template <> inline float wpow<float>(float a, float b) { return fpow(a, b); }
template <> inline double wpow<double>(double a, double b) { return std::pow(a, b); }

template <class R> struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3 operator*(R s) const { return V3(x * s, y * s, z * s); }
    V3 operator/(R s) const { return V3(x / s, y / s, z / s); }
};
// a / b through z = RN(1 / b) and one Markstein correction (q = RN(a z), r = fma(-q, b, a) exact, RN(q + r z)): the
// correctly rounded quotient while nothing underflows (Markstein 1990), the kernels' form of the vorticity
// confinement's three quotients by one length (WGSL's normalize divides with 2.5 ulp of accuracy)
static inline float div_recip(float a, float b, float z) {
    const float q = a * z;
    const float r = std::fma(-q, b, a);
    return std::fma(r, z, q);
}
template <class R> static inline V3<R> unit_dir(const V3<R>& g, R gl);   // g / gl
template <> inline V3<float> unit_dir<float>(const V3<float>& g, float gl) {
    const float z = 1.0f / gl;
    return V3<float>(div_recip(g.x, gl, z), div_recip(g.y, gl, z), div_recip(g.z, gl, z));
}
template <> inline V3<double> unit_dir<double>(const V3<double>& g, double gl) { return g / gl; }
template <class R> static inline R dot(const V3<R>& a, const V3<R>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> static inline R length(const V3<R>& a) { return std::sqrt(dot(a, a)); }
template <class R> static inline V3<R> cross(const V3<R>& a, const V3<R>& b) {
    return V3<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class R> struct V4 {
    R x, y, z, w;
    V4() : x(0), y(0), z(0), w(0) {}
    V4(R a, R b, R c, R d) : x(a), y(b), z(c), w(d) {}
    V4 operator+(const V4& o) const { return V4(x + o.x, y + o.y, z + o.z, w + o.w); }
    V4 operator-(const V4& o) const { return V4(x - o.x, y - o.y, z - o.z, w - o.w); }
    V4 operator*(R s) const { return V4(x * s, y * s, z * s, w * s); }
    V3<R> xyz() const { return V3<R>(x, y, z); }
};
template <class R> static inline V4<R> mix4(const V4<R>& a, const V4<R>& b, R t) {
    return V4<R>(mix(a.x, b.x, t), mix(a.y, b.y, t), mix(a.z, b.z, t), mix(a.w, b.w, t));
}
// Trilinear filter of the 2x2x2 texels c[k] at fractions (fx, fy, fz), corner k = (k & 1, (k >> 1) & 1, k >> 2):
// weight w_k = wx_a * (wy_b * wz_c) (wx_0 = 1 - fx, wx_1 = fx, ...), value = c_0 w_0, then fma(c_k, w_k, .) for
// k = 1..7.  The arithmetic of the texture sampler is the implementation's (src/sim.js:129, 154-156, 176-178,
// 228, 365 call textureSampleLevel); this fixes one, and the HIP kernels round identically.
// One axis of the clamp-to-edge footprint at texel-space coordinate t (n >= 2): texels i0, i0 + 1 with
// i0 = clamp(floor(t), 0, n - 2) and the fraction f of the second.  Where floor(t) is outside [0, n - 2]
// both clamped texels are the same edge texel, and the fraction 0 (below) or 1 (above) takes it exactly.
template <class R> static inline int tri_axis(R t, int n, R& f) {
    const R fl = std::floor(t);
    const int i = (int)fl;
    f = i < 0 ? R(0) : i > n - 2 ? R(1) : t - fl;
    return std::min(std::max(i, 0), n - 2);
}
template <class R> static inline void tri_weights(R fx, R fy, R fz, R* w) {
    const R ux = R(1) - fx, uy = R(1) - fy, uz = R(1) - fz;
    const R yz[4] = {uy * uz, fy * uz, uy * fz, fy * fz};
    for (int j = 0; j < 4; j++) {
        w[2 * j] = ux * yz[j];
        w[2 * j + 1] = fx * yz[j];
    }
}
template <class R> static inline V4<R> tri_blend(const V4<R>* c, const R* w) {
    V4<R> r(c[0].x * w[0], c[0].y * w[0], c[0].z * w[0], c[0].w * w[0]);
    for (int k = 1; k < 8; k++) {
        r.x = std::fma(c[k].x, w[k], r.x);
        r.y = std::fma(c[k].y, w[k], r.y);
        r.z = std::fma(c[k].z, w[k], r.z);
        r.w = std::fma(c[k].w, w[k], r.w);
    }
    return r;
}

// ---- hashing: pcg (src/shaders/lib.js:11-17) --------------------------------
static inline uint32_t pcg1(uint32_t n) {
    const uint32_t s = n * 747796405u + 2891336453u;
    const uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
// rand2(n, m) = f32(pcg(n ^ pcg(m))) * f32(1/4294967295); the constant rounds to 2^-32
static inline float rand2(uint32_t n, uint32_t m) { return (float)pcg1(n ^ pcg1(m)) * 2.3283064365386963e-10f; }

// ---- hashing: pcg3d + hash3f (src/shaders/lib.js:4-16) -----------------------
static inline void pcg3d(uint32_t& x, uint32_t& y, uint32_t& z) {
    x = x * 1664525u + 1013904223u; y = y * 1664525u + 1013904223u; z = z * 1664525u + 1013904223u;
    x += y * z; y += z * x; z += x * y;
    x ^= x >> 16; y ^= y >> 16; z ^= z >> 16;
    x += y * z; y += z * x; z += x * y;
}
// hash3f: f32(u32) * f32(1/4294967295)
static inline void hash3f(int32_t px, int32_t py, int32_t pz, float out[3]) {
    uint32_t x = (uint32_t)px, y = (uint32_t)py, z = (uint32_t)pz;
    pcg3d(x, y, z);
    const float k = (float)(1.0 / 4294967295.0);
    out[0] = (float)x * k; out[1] = (float)y * k; out[2] = (float)z * k;
}

}  // namespace orc
