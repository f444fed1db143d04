// The fused cell update of the source kernel (src/sim.js:230-396), generic in the
// scalar type of the state: `float` for the step itself, `Dual` (value + one
// directional derivative) for the adjoint's local Jacobians.  Quantities that do not
// depend on the state (the emitter terms driven by the noise textures, the curl-noise
// sample, the sponge weights) are computed by the caller as plain floats; everything
// downstream of the state is written once, here, in the reference's order of
// operations, so the float instantiation is the step's arithmetic exactly.
#pragma once
#include <type_traits>

#include "fs_math.h"

// ---- a forward-mode dual number ------------------------------------------------
struct Dual {
    float v, d;
    Dual() = default;
    __device__ __forceinline__ explicit Dual(float v_, float d_ = 0.f) : v(v_), d(d_) {}
};
__device__ __forceinline__ Dual dual(float v, float d = 0.f) { return Dual(v, d); }
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator+(Dual a, float b) { return Dual(a.v + b, a.d); }
__device__ __forceinline__ Dual operator+(float a, Dual b) { return Dual(a + b.v, b.d); }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator-(Dual a, float b) { return Dual(a.v - b, a.d); }
__device__ __forceinline__ Dual operator-(float a, Dual b) { return Dual(a - b.v, -b.d); }
__device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ Dual operator*(Dual a, float b) { return Dual(a.v * b, a.d * b); }
__device__ __forceinline__ Dual operator*(float a, Dual b) { return Dual(a * b.v, a * b.d); }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) { return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
__device__ __forceinline__ Dual operator/(Dual a, float b) { return Dual(a.v / b, a.d / b); }
__device__ __forceinline__ Dual operator/(float a, Dual b) { return Dual(a / b.v, -a * b.d / (b.v * b.v)); }

// ---- the step's scalar functions, for float (exactly the step's) and for Dual ---------
__device__ __forceinline__ float gval(float a) { return a; }
__device__ __forceinline__ float gval(Dual a) { return a.v; }
__device__ __forceinline__ float gmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float gmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ Dual gmin(Dual a, Dual b) { return a.v <= b.v ? a : b; }
__device__ __forceinline__ Dual gmax(Dual a, Dual b) { return a.v >= b.v ? a : b; }
__device__ __forceinline__ Dual gmin(Dual a, float b) { return a.v <= b ? a : dual(b); }
__device__ __forceinline__ Dual gmax(Dual a, float b) { return a.v >= b ? a : dual(b); }
__device__ __forceinline__ Dual gmin(float a, Dual b) { return gmin(b, a); }
__device__ __forceinline__ Dual gmax(float a, Dual b) { return gmax(b, a); }
template <class T> __device__ __forceinline__ T gclamp(T e, float lo, float hi) { return gmin(gmax(e, lo), hi); }
__device__ __forceinline__ float gsstep(float e0, float e1, float x) { return sstep(e0, e1, x); }
__device__ __forceinline__ Dual gsstep(float e0, float e1, Dual x) {
    const Dual t = gclamp((x - e0) / (e1 - e0), 0.f, 1.f);
    return t * t * (3.f - 2.f * t);
}
// literal edges: the float path divides through the reciprocal (sstep_c, bit-identical values); the
// dual path keeps the plain division
__device__ __forceinline__ float gsstep_c(float e0, float e1, float x) { return sstep_c(e0, e1, x); }
__device__ __forceinline__ Dual gsstep_c(float e0, float e1, Dual x) { return gsstep(e0, e1, x); }
template <class A, class B, class C>
__device__ __forceinline__ auto gmix(A a, B b, C t) -> decltype(a * (1.f - t) + b * t) { return a * (1.f - t) + b * t; }
__device__ __forceinline__ float gexp(float x) { return fs_exp(x); }
__device__ __forceinline__ Dual gexp(Dual x) {
    const float e = fs_exp(x.v);
    return Dual(e, e * x.d);
}
__device__ __forceinline__ float gpow(float a, float b) { return fs_pow(a, b); }
__device__ __forceinline__ Dual gpow(Dual a, float b) {   // d/da a^b = b a^(b-1)
    return Dual(fs_pow(a.v, b), a.v > 0.f ? b * fs_pow(a.v, b - 1.f) * a.d : 0.f);
}
__device__ __forceinline__ float gsqrt(float a) { return sqrtf(a); }
__device__ __forceinline__ Dual gsqrt(Dual a) {
    const float r = sqrtf(a.v);
    return Dual(r, r > 0.f ? a.d / (2.f * r) : 0.f);
}
// a spot's mass flux as T: its derivative is spot_dir[q] when a direction is given (tangent
// runs), else 1 for the seeded spot (one column of the adjoint's local Jacobian)
template <class T>
__device__ __forceinline__ T flux_as(float f, int q, int seed_spot, const float* spot_dir);
template <> __device__ __forceinline__ float flux_as<float>(float f, int, int, const float*) { return f; }
template <> __device__ __forceinline__ Dual flux_as<Dual>(float f, int q, int seed_spot, const float* spot_dir) {
    return Dual(f, spot_dir ? spot_dir[q] : (q == seed_spot ? 1.f : 0.f));
}
__device__ __forceinline__ bool gisnan(float a) { return isnan(a); }
__device__ __forceinline__ bool gisnan(Dual a) { return isnan(a.v); }

// ---- four directional derivatives at once (vector forward mode) -----------------------
// Each derivative component goes through exactly the operations of Dual's, in the same order, so a
// Dual4 pass gives four Dual passes' derivatives bit for bit while evaluating the value path once.
struct Dual4 {
    float v, d[4];
    Dual4() = default;
    __device__ __forceinline__ explicit Dual4(float v_) : v(v_), d{0.f, 0.f, 0.f, 0.f} {}
};
#define FS_D4(expr) do { _Pragma("unroll") for (int u = 0; u < 4; u++) { expr; } } while (0)
__device__ __forceinline__ Dual4 operator+(Dual4 a, Dual4 b) { Dual4 r(a.v + b.v); FS_D4(r.d[u] = a.d[u] + b.d[u]); return r; }
__device__ __forceinline__ Dual4 operator+(Dual4 a, float b) { Dual4 r(a.v + b); FS_D4(r.d[u] = a.d[u]); return r; }
__device__ __forceinline__ Dual4 operator+(float a, Dual4 b) { Dual4 r(a + b.v); FS_D4(r.d[u] = b.d[u]); return r; }
__device__ __forceinline__ Dual4 operator-(Dual4 a, Dual4 b) { Dual4 r(a.v - b.v); FS_D4(r.d[u] = a.d[u] - b.d[u]); return r; }
__device__ __forceinline__ Dual4 operator-(Dual4 a, float b) { Dual4 r(a.v - b); FS_D4(r.d[u] = a.d[u]); return r; }
__device__ __forceinline__ Dual4 operator-(float a, Dual4 b) { Dual4 r(a - b.v); FS_D4(r.d[u] = -b.d[u]); return r; }
__device__ __forceinline__ Dual4 operator-(Dual4 a) { Dual4 r(-a.v); FS_D4(r.d[u] = -a.d[u]); return r; }
__device__ __forceinline__ Dual4 operator*(Dual4 a, Dual4 b) {
    Dual4 r(a.v * b.v);
    FS_D4(r.d[u] = a.d[u] * b.v + a.v * b.d[u]);
    return r;
}
__device__ __forceinline__ Dual4 operator*(Dual4 a, float b) { Dual4 r(a.v * b); FS_D4(r.d[u] = a.d[u] * b); return r; }
__device__ __forceinline__ Dual4 operator*(float a, Dual4 b) { Dual4 r(a * b.v); FS_D4(r.d[u] = a * b.d[u]); return r; }
__device__ __forceinline__ Dual4 operator/(Dual4 a, Dual4 b) {
    Dual4 r(a.v / b.v);
    FS_D4(r.d[u] = (a.d[u] * b.v - a.v * b.d[u]) / (b.v * b.v));
    return r;
}
__device__ __forceinline__ Dual4 operator/(Dual4 a, float b) { Dual4 r(a.v / b); FS_D4(r.d[u] = a.d[u] / b); return r; }
__device__ __forceinline__ Dual4 operator/(float a, Dual4 b) {
    Dual4 r(a / b.v);
    FS_D4(r.d[u] = -a * b.d[u] / (b.v * b.v));
    return r;
}
__device__ __forceinline__ float gval(Dual4 a) { return a.v; }
__device__ __forceinline__ Dual4 gmin(Dual4 a, Dual4 b) { return a.v <= b.v ? a : b; }
__device__ __forceinline__ Dual4 gmax(Dual4 a, Dual4 b) { return a.v >= b.v ? a : b; }
__device__ __forceinline__ Dual4 gmin(Dual4 a, float b) { return a.v <= b ? a : Dual4(b); }
__device__ __forceinline__ Dual4 gmax(Dual4 a, float b) { return a.v >= b ? a : Dual4(b); }
__device__ __forceinline__ Dual4 gmin(float a, Dual4 b) { return gmin(b, a); }
__device__ __forceinline__ Dual4 gmax(float a, Dual4 b) { return gmax(b, a); }
__device__ __forceinline__ Dual4 gsstep(float e0, float e1, Dual4 x) {
    const Dual4 t = gclamp((x - e0) / (e1 - e0), 0.f, 1.f);
    return t * t * (3.f - 2.f * t);
}
__device__ __forceinline__ Dual4 gsstep_c(float e0, float e1, Dual4 x) { return gsstep(e0, e1, x); }
__device__ __forceinline__ Dual4 gexp(Dual4 x) {
    const float e = fs_exp(x.v);
    Dual4 r(e);
    FS_D4(r.d[u] = e * x.d[u]);
    return r;
}
__device__ __forceinline__ Dual4 gpow(Dual4 a, float b) {   // d/da a^b = b a^(b-1)
    Dual4 r(fs_pow(a.v, b));
    if (a.v > 0.f) {
        const float k = b * fs_pow(a.v, b - 1.f);
        FS_D4(r.d[u] = k * a.d[u]);
    }
    return r;
}
__device__ __forceinline__ Dual4 gsqrt(Dual4 a) {
    const float rt = sqrtf(a.v);
    Dual4 r(rt);
    if (rt > 0.f) FS_D4(r.d[u] = a.d[u] / (2.f * rt));
    return r;
}
template <> __device__ __forceinline__ Dual4 flux_as<Dual4>(float f, int, int, const float*) { return Dual4(f); }
__device__ __forceinline__ bool gisnan(Dual4 a) { return isnan(a.v); }

// a / b with z = 1 / b.v given: the value through the reciprocal and one Markstein correction (fs_div_recip: the
// correctly rounded quotient while nothing underflows), the derivative that of the plain quotient
__device__ __forceinline__ float gdivr(float a, float b, float z) { return fs_div_recip(a, b, z); }
__device__ __forceinline__ Dual gdivr(Dual a, Dual b, float z) {
    return Dual(fs_div_recip(a.v, b.v, z), (a.d * b.v - a.v * b.d) / (b.v * b.v));
}
__device__ __forceinline__ Dual4 gdivr(Dual4 a, Dual4 b, float z) {
    Dual4 r(fs_div_recip(a.v, b.v, z));
    FS_D4(r.d[u] = (a.d[u] * b.v - a.v * b.d[u]) / (b.v * b.v));
    return r;
}

// ---- the state-independent part of one cell's update -----------------------------
struct ForcesCell {
    float dt, dx, tm, I;
    float px, py, pz;          // cell centre (m)
    int cy;                    // y index (floor fuel flux)
    float4 gg;                 // geometry: band, bed, solid, sdf
    float src, srcW, bed, bedHot, anchor, wsrc;
    float cnx, cny, cnz;       // curl-noise sample
    float lapx, lapy, lapz;    // 7-point Laplacian of the velocity (hot-gas viscosity on; else 0)
    float wSx, wSz, ks, kt;    // sponge targets and weights, top fade
};

// Fuel fraction and temperature at the boundaries of the sub-models of one cell update, for the
// FP64 mass and energy budget (fs_budget.hip).  NoTrace (the step itself) records nothing and
// compiles to nothing.
struct NoTrace {};
struct CellTrace {
    float y_in, th_in;           // after advection (the kernel's inputs)
    float y_src, th_src;         // after the log-band, coal-bed and burst sources
    float y_flux, th_flux;       // after the pyrolysis spots and the floor flux
    float t0;                    // 293 + 1000 theta (K), the burner's input temperature
    float y_burn, tk_burn;       // after mixing-limited burning (fuel; T0 + dT)
    float tk_rad, tk_cool;       // after radiative and after entrainment cooling (K)
    float th_cool;               // theta = (T - 293) / 1000
    float y_diss;                // after fuel dissipation
    float y_out, th_out;         // after the solid cells and the NaN reset
};
#define FS_TRACE(...) do { if constexpr (!std::is_same<Tr, NoTrace>::value) { __VA_ARGS__; } } while (0)

// One cell: v (3), s (4: T, fuel, soot, q), s2 (4: vitiated, age, smoke, blue) in place;
// w = own vorticity (3), gx/gy/gz = differences of |w| of the neighbours, lap = the
// 7-point Laplacian of the velocity (3; read only with hot-gas viscosity on); phi out.
// seed_spot: the spot whose flux carries a unit derivative (Dual only; -1: none), or
// spot_dir: the derivative of every spot's flux (Dual only; null: use seed_spot);
// face: the floor flux of this cell as T; spot_mask: bit q clear only where spot q provably
// misses the cell (spot_reach), so the spots loop visits the set bits in increasing q.
template <class T, class Tr = NoTrace>
__device__ __forceinline__ void forces_cell(const float* U, const float* D, const ForcesCell& k, const fs_spot* __restrict__ spots,
                                            int n_spots, int seed_spot, const float* spot_dir, bool has_face,
                                            T face, T& vx, T& vy,
                                            T& vz, T* s, T* s2, const T* w, T gx, T gy, T gz, const T* lap,
                                            T& phi_out, Tr* tr = nullptr, unsigned long long spot_mask = ~0ull) {
    const float dt = k.dt, I = k.I;
    FS_TRACE(tr->y_in = s[1]; tr->th_in = s[0]);
    // ---- sources (sim.js:246-297)
    s[1] = gmax(s[1], k.src * U[P_fuelSrc]);
    s[1] = gmax(s[1], k.bed * U[P_bedFuel]);
    s[0] = gmax(s[0], k.src * U[P_srcTemp]);
    s[0] = gmax(s[0], k.bedHot * U[P_bedTemp]);
    float srcAll = fmaxf(k.srcW, k.bed);
    s2[3] = gmix(s2[3], 1.f, k.bed * (1.f - k.srcW));
    s2[3] = gmix(s2[3], 0.f, k.srcW);
    s2[1] = s2[1] * (1.f - srcAll);
    s2[0] = gmax(s2[0], (1.f - s[1]) * k.srcW);
    s2[0] = gmax(s2[0], (1.f - s[1]) * k.bed * 0.6f);
    vy = gmix(vy, gmax(vy, U[P_srcVel]), k.wsrc);
    s2[2] = s2[2] + k.gg.x * (1.f - k.anchor) * U[P_smolder] * dt * I;
    if (U[P_burstAmt] > 0.f) {
        const float ddx = k.px - U[P_burstX], ddy = k.py - U[P_burstY], ddz = k.pz - U[P_burstZ];
        const float bd = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
        const float bw = wclamp(fs_exp(-bd * bd / 0.006f) * U[P_burstAmt], 0.f, 1.f);
        s[1] = gmax(s[1], 0.17f * bw);
        s[0] = gmax(s[0], 0.9f * bw);
        s2[0] = s2[0] * (1.f - bw);
        s2[3] = s2[3] * (1.f - bw);
        s2[1] = gmax(s2[1], 0.05f * bw * gclamp(s[1] + s2[0], 0.f, 1.f));
    }
    FS_TRACE(tr->y_src = s[1]; tr->th_src = s[0]);
    // ---- pyrolysis spots and floor fuel flux (build extension: mass-flux sources)
    for (unsigned long long m = spot_mask & (n_spots >= 64 ? ~0ull : (1ull << n_spots) - 1); m != 0; m &= m - 1) {
        const int q = __builtin_ctzll(m);
        const fs_spot sp = spots[q];
        const float ddx = k.px - sp.pos[0], ddy = k.py - sp.pos[1], ddz = k.pz - sp.pos[2];
        const float d = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
        if (d < sp.radius) {
            const float wq = 1.f - sstep(0.5f * sp.radius, sp.radius, d);
            s[1] = gmin(s[1] + flux_as<T>(sp.flux, q, seed_spot, spot_dir) * dt * sp.norm * wq, 1.f);
            s[0] = gmax(s[0], wq * sp.temp);
            s2[3] = gmix(s2[3], 0.f, wq);
            s2[1] = s2[1] * (1.f - wq);
            srcAll = fmaxf(srcAll, wq);
        }
    }
    if (has_face && k.cy == 0) s[1] = gmin(s[1] + face * dt / (FS_RHO0 * k.dx), 1.f);
    FS_TRACE(tr->y_flux = s[1]; tr->th_flux = s[0]);
    T ageF = s2[1] / gmax(s[1] + s2[0], 1e-4f);

    // ---- mixing-limited combustion (sim.js:300-309)
    const T T0 = 293.f + 1000.f * s[0];
    const T o = gclamp(1.f - s[1] - s2[0], 0.f, 1.f);
    const T ign = gsstep(U[P_ignT] - 50.f, U[P_ignT] + 50.f, T0);
    const T b = gmin(s[1], o / U[P_rs]) * (1.f - gexp(-U[P_kMix] * ign * dt));
    s[1] = s[1] - b;
    s2[0] = gmin(s2[0] + (1.f + U[P_rs]) * b, 1.f);
    const T dT = gmin(U[P_heatOut] * b, gmax(2100.f - T0, 0.f));
    T Tk = T0 + dT;
    FS_TRACE(tr->t0 = T0; tr->y_burn = s[1]; tr->tk_burn = Tk);
    const T q = dT / dt;
    const T phi = k.gg.z > 0.5f ? T(0.f) : gmin(U[P_expansion] * q / Tk, 60.f);

    // ---- soot (sim.js:312-318)
    ageF = gmin(ageF + dt, 4.f);
    s[2] = s[2] + U[P_sootForm] * b * gsstep_c(0.015f, 0.05f, ageF) * gsstep_c(1000.f, 1400.f, Tk) * (1.f - s2[3]);
    s[2] = s[2] * gexp(-U[P_sootOx] * (0.25f + o) * gsstep_c(1100.f, 1400.f, Tk) * dt);
    s[2] = s[2] * gexp(-U[P_sootOx] * 0.5f * o * gsstep_c(850.f, 1100.f, Tk) * (1.f - gsstep_c(1100.f, 1300.f, Tk)) * dt);
    s[2] = s[2] * D[D_eSootDissip];                      // exp(-sootDissip dt)

    // ---- cooling (sim.js:321-330)
    const T th = Tk * 0.001f;
    Tk = Tk - gmin(U[P_radCool] * 970.f * s[2] * th * th * (th * th * th * th - 0.00737f), 6000.f) * dt;
    FS_TRACE(tr->tk_rad = Tk);
    const T notBurning = 1.f - gsstep_c(500.f, 5000.f, q);
    Tk = 293.f + gmax(Tk - 293.f, 0.f) * gexp(-U[P_coolRate] * notBurning * dt);
    s[0] = (Tk - 293.f) * 0.001f;
    FS_TRACE(tr->tk_cool = Tk; tr->th_cool = s[0]);
    s[3] = q * 1e-4f;
    s[1] = s[1] * D[D_eFuelDissip];                      // exp(-fuelDissip dt)
    FS_TRACE(tr->y_diss = s[1]);
    s2[0] = s2[0] * D[D_eVitDecay];                      // exp(-vitDecay dt)
    s2[2] = s2[2] * D[D_eSmokeDissip];                   // exp(-smokeDissip dt)
    s2[1] = ageF * gclamp(s[1] + s2[0], 0.f, 1.f);

    // ---- buoyancy (sim.js:333)
    vy = vy + U[P_buoyancy] * s[0] * dt;

    // ---- hot-gas viscosity (sim.js:339-346)
    if (U[P_visc] > 0.f) {
        const T nu = U[P_visc] * 1.5e-5f * gpow(gmax(Tk, 293.f) / 293.f, 1.7f);
        const T kv = gmin(nu * dt / (k.dx * k.dx), 0.15f);
        vx = vx + lap[0] * kv; vy = vy + lap[1] * kv; vz = vz + lap[2] * kv;
    }
    // ---- laminar base gate (sim.js:350)
    const float lamAge = U[P_lamAge];
    const T lam = lamAge > 0.f ? gsstep(0.5f * lamAge, fmaxf(lamAge, 1e-3f), ageF) : T(1.f);

    // ---- vorticity confinement (sim.js:353-361)
    const T gl = gsqrt(gx * gx + gy * gy + gz * gz);
    if (gval(gl) > 1e-6f) {
        const T gate = gsstep_c(450.f, 900.f, Tk) * (1.f - srcAll) * sstep_c(0.02f, 0.08f, k.py) * lam;
        const float zr = 1.f / gval(gl);                // the three quotients by gl through its reciprocal
        const T nx_ = gdivr(gx, gl, zr), ny_ = gdivr(gy, gl, zr), nz_ = gdivr(gz, gl, zr);
        const T kc = U[P_vortEps] * gate * dt;
        vx = vx + (ny_ * w[2] - nz_ * w[1]) * kc;
        vy = vy + (nz_ * w[0] - nx_ * w[2]) * kc;
        vz = vz + (nx_ * w[1] - ny_ * w[0]) * kc;
    }
    // ---- rising curl noise (sim.js:364-368)
    const T tgate = gmax(gsstep_c(500.f, 1100.f, Tk), gsstep_c(0.f, 3.f, s[3])) * (1.f - srcAll) *
                    mixf(0.6f, 1.8f, sstep_c(0.2f, 0.75f, k.py)) * lam;
    {
        const T kn = D[D_turbAmp12] * tgate * dt * 10.f;   // turbAmp / 12
        vx = vx + k.cnx * kn; vy = vy + k.cny * kn; vz = vz + k.cnz * kn;
    }
    // ---- wind (sim.js:372-375)
    const float windK = sstep_c(0.05f, 0.5f, k.py) * (1.f + U[P_gust]);
    vx = vx + (U[P_windX] * windK - vx) * dt * 0.35f * windK;
    vz = vz + (U[P_windZ] * windK - vz) * dt * 0.35f * windK;
    {
        const float e = D[D_eVelDissip];                 // exp(-velDissip dt)
        vx = vx * e; vy = vy * e; vz = vz * e;
    }
    // ---- solids and sponge layers (sim.js:378-388)
    if (k.gg.z > 0.5f) { vx = T(0.f); vy = T(0.f); vz = T(0.f); s[1] = T(0.f); s2[1] = T(0.f); }
    vx = gmix(vx, k.wSx, k.ks);
    vz = gmix(vz, k.wSz, k.ks);
    s[2] = s[2] * k.kt;
    s2[2] = s2[2] * k.kt;
    // ---- clamps and NaN reset (sim.js:390-392)
    const T vl = gsqrt(vx * vx + vy * vy + vz * vz);
    if (gval(vl) > 15.f) { const T kk = 15.f / vl; vx = vx * kk; vy = vy * kk; vz = vz * kk; }
    bool bad = gisnan(vx) || gisnan(vy) || gisnan(vz);
    for (int c = 0; c < 4; c++) bad = bad || gisnan(s[c]) || gisnan(s2[c]);
    if (bad) {
        for (int c = 0; c < 4; c++) { s[c] = T(0.f); s2[c] = T(0.f); }
        vx = T(0.f); vy = T(0.f); vz = T(0.f);
    }
    FS_TRACE(tr->y_out = s[1]; tr->th_out = s[0]);
    phi_out = phi;
}
#undef FS_TRACE

// The spots that can reach a cell of this wave's 64-cell x-row (y and z fixed), as a wave-uniform mask:
// lane q tests spot q against the row's bounding box in float64.  A spot is left out only if the box lies
// farther from its centre than radius / (1 - 1e-5); every cell centre of the row is at least that far, and
// the cell's f32 distance (three rounded differences, squares and sums, a rounded sqrt: within a few units of
// 2^-24 relative) is then still >= radius, so the cell's own test d < radius is false for it.  The mask
// therefore only skips loop trips that change nothing.  All 64 lanes must call it (before any early exit).
__device__ __forceinline__ unsigned long long spot_reach(const float* U, const Dims& g, int cx, int cy, int cz,
                                                         const fs_spot* __restrict__ spots, int n_spots) {
    const int lane = (int)(threadIdx.x & 63);
    const int x0 = cx - lane, x1 = min(x0 + 63, g.nx - 1);
    const float dx = U[P_dx];
    const double lo = (double)(U[P_ox] + ((float)x0 + 0.5f) * dx), hi = (double)(U[P_ox] + ((float)x1 + 0.5f) * dx);
    const double py = (double)(U[P_oy] + ((float)cy + 0.5f) * dx), pz = (double)(U[P_oz] + ((float)cz + 0.5f) * dx);
    bool near = false;
    if (lane < n_spots) {
        const fs_spot sp = spots[lane];
        const double sx = sp.pos[0], ex = fmax(0.0, fmax(lo - sx, sx - hi));
        const double ey = py - (double)sp.pos[1], ez = pz - (double)sp.pos[2];
        const double r = (double)sp.radius / (1.0 - 1e-5);
        near = ex * ex + ey * ey + ez * ez < r * r;
    }
    return __ballot(near);
}

// The per-cell constants of the source kernel: the emitter terms driven by the noise
// textures (sim.js:246-268), the curl-noise sample (sim.js:364-365), the sponge weights
// (sim.js:380-388) and, with hot-gas viscosity on, the velocity Laplacian.  Shared with
// the adjoint, which differentiates only forces_cell.
template <class S>
__device__ __forceinline__ ForcesCell forces_setup(const DevParams* __restrict__ P, const Dims& g, const Cell& c,
                                                   const typename S::raw* __restrict__ velY,
                                                   const typename S::raw* __restrict__ geo, const half4_t* __restrict__ noise,
                                                   const half4_t* __restrict__ curlT, int nn) {
    const float* U = P->f;
    ForcesCell k;
    k.dt = (float)P->dt; k.dx = U[P_dx]; k.tm = (float)P->time; k.I = U[P_intensity];
    const float dx = k.dx, tm = k.tm, I = k.I;
    k.px = U[P_ox] + ((float)c.x + 0.5f) * dx;
    k.py = U[P_oy] + ((float)c.y + 0.5f) * dx;
    k.pz = U[P_oz] + ((float)c.z + 0.5f) * dx;
    k.cy = c.y;
    const float px = k.px, py = k.py, pz = k.pz;
    const float4 gg = Codec<S>::ld(geo, c.i);
    k.gg = gg;
    float src = 0.f, anchor = 0.f;
    if (gg.x > 0.001f) {
        const float af = U[P_anchorFreq];
        const float4 na = sample_rep(noise, nn, px * af + U[P_seed] * 0.37f, py * af + 0.f, pz * af + 0.11f);
        anchor = sstep_c(-0.05f, 0.25f, na.z * 2.f - 1.f);
        const float sf = U[P_srcFreq] * 0.06f;
        const float4 nq = sample_rep(noise, nn, px * sf + U[P_seed] * 0.131f, py * sf + 0.f, pz * sf + tm * U[P_srcSpeed]);
        const float pulse = sstep(U[P_srcThresh] - 0.25f, U[P_srcThresh] + 0.25f, nq.x * 2.f - 1.f);
        src = wclamp(gg.x * U[P_logSrc] * (U[P_srcFloor] + U[P_anchorGain] * anchor) * mixf(0.55f, 1.f, pulse) * 1.2f, 0.f, 1.f) * I;
        src *= 1.f - gg.z;
    }
    const float srcW = fminf(src, 1.f);
    float bed = gg.y * wclamp(I * 1.4f, 0.3f, 1.f);
    if (bed > 0.001f) {
        const float4 nb = sample_rep(noise, nn, px * 2.4f + 0.17f, py * 2.4f + -tm * 0.12f, pz * 2.4f + tm * 0.05f);
        bed *= sstep_c(0.f, 0.3f, nb.y * 2.f - 1.f);
    }
    k.src = src; k.srcW = srcW; k.bed = bed; k.bedHot = gg.y; k.anchor = anchor;
    k.wsrc = 1.f - fs_pow(1.f - wclamp(srcW + bed * 0.4f, 0.f, 1.f) * 0.5f, k.dt * 90.f);
    const float tf = U[P_turbFreq];
    const float tqx = (px - 0.f) * tf + 0.f;
    const float tqy = (py - tm * U[P_turbRise]) * tf + 0.f;
    const float tqz = (pz - 0.f) * tf + tm * U[P_turbEvolve];
    const float4 cn = sample_rep(curlT, nn, tqx, tqy, tqz);
    k.cnx = cn.x; k.cny = cn.y; k.cnz = cn.z;
    k.lapx = k.lapy = k.lapz = 0.f;
    if (U[P_visc] > 0.f) {
        const float4 a = ldc<S>(velY, g, c.x + 1, c.y, c.z), bb = ldc<S>(velY, g, c.x - 1, c.y, c.z);
        const float4 cc = ldc<S>(velY, g, c.x, c.y + 1, c.z), dd = ldc<S>(velY, g, c.x, c.y - 1, c.z);
        const float4 ee = ldc<S>(velY, g, c.x, c.y, c.z + 1), ff = ldc<S>(velY, g, c.x, c.y, c.z - 1);
        const float4 me = Codec<S>::ld(velY, c.i);
        k.lapx = a.x + bb.x + cc.x + dd.x + ee.x + ff.x - me.x * 6.f;
        k.lapy = a.y + bb.y + cc.y + dd.y + ee.y + ff.y - me.y * 6.f;
        k.lapz = a.z + bb.z + cc.z + dd.z + ee.z + ff.z - me.z * 6.f;
    }
    const float fpx = (float)c.x + 0.5f, fpz = (float)c.z + 0.5f;
    const float side = fminf(fminf(fpx, (float)g.nx - fpx), fminf(fpz, (float)g.nz - fpz));
    const float wsg = sstep_c(0.03f, 0.4f, py);
    k.wSx = U[P_windX] * wsg; k.wSz = U[P_windZ] * wsg;
    // side is j + 0.5 (j an integer); from 5 cells in, the sponge weight is 0 and the factor 1 - pow(1, .) = 0
    k.ks = side < 5.f ? P->d[D_ks0 + (int)side] : 0.f;
    // below the top 10 rows the smoothstep is 0 and the factor pow(1, .) = 1
    k.kt = c.y >= g.ny - 10 ? P->d[D_kt0 + c.y - (g.ny - 10)] : 1.f;
    return k;
}

// The per-step derived values (see DevParams::d), from the same expressions the source kernel
// evaluates per cell (src/sim.js:321-375 and the sponge rows of src/sim.js:380-388), so the
// factors it reads are bit-identical to evaluating them in place.
__device__ __forceinline__ void derive_params(DevParams& p) {
    const float* U = p.f;
    const float dt = (float)p.dt;
    p.d[D_eSootDissip] = fs_exp(-U[P_sootDissip] * dt);
    p.d[D_eFuelDissip] = fs_exp(-U[P_fuelDissip] * dt);
    p.d[D_eVitDecay] = fs_exp(-U[P_vitDecay] * dt);
    p.d[D_eSmokeDissip] = fs_exp(-U[P_smokeDissip] * dt);
    p.d[D_eVelDissip] = fs_exp(-U[P_velDissip] * dt);
    p.d[D_turbAmp12] = U[P_turbAmp] / 12.f;
    for (int j = 0; j < 5; j++) {
        const float side = (float)j + 0.5f;
        const float sp = (1.f - sstep(0.f, 5.f, side)) * U[P_sponge];
        p.d[D_ks0 + j] = 1.f - fs_pow(1.f - 0.15f * wclamp(sp, 0.f, 1.f), dt * 90.f);
    }
    const int ny = (int)U[P_ny];
    for (int j = 0; j < 10; j++) {
        const float fpy = (float)(ny - 10 + j) + 0.5f;
        const float top = sstep((float)ny - 10.f, (float)ny, fpy);
        p.d[D_kt0 + j] = fs_pow(1.f - 0.1f * top, dt * 90.f);
    }
}
