// Spark (ember) particles carried by the simulated velocity field: the update kernel of
// src/sparks.js:55-135 on gfx950.
//
// One thread per particle (64-B record: position+age, velocity+burn time, previous
// position+temperature, diameter/seed/alive/flicker).  Dead particles claim spawn slots;
// the slot, not the particle index, drives the randomness, so the spawned set does not
// depend on which particle gets which slot.  The reference takes slots from an atomic
// counter (any order); here the k-th dead particle in index order takes slot k, which is
// one of those orders and makes the result deterministic.  The update kernel also counts,
// per block, the particles dead after it, so the next step only scans those counts: one
// read and one write of each 64-B record per step (a separate count pass, reading the
// alive flags, runs only after the array was written from the host or reset).
#include <hip/hip_runtime.h>

#include "fs_codec.h"
#include "fs_internal.h"
#include "fs_math.h"

namespace {

constexpr int SPARK_BLOCK = 256;

// src/shaders/lib.js:11-17
__device__ __forceinline__ unsigned pcg1(unsigned n) {
    const unsigned s = n * 747796405u + 2891336453u;
    const unsigned w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
__device__ __forceinline__ float rand2(unsigned n, unsigned m) {
    return (float)pcg1(n ^ pcg1(m)) * 2.3283064365386963e-10f;   // 1 / 4294967295 rounded to f32 = 2^-32
}

struct Rng {   // r(slot, k) and gauss(slot, k) of src/sparks.js:66-70
    unsigned slot, seed, frame;
    __device__ float r(unsigned k) const { return rand2(slot * 7919u + k * 104729u + seed, frame); }
    __device__ float gauss(unsigned k) const {
        const float u1 = fmaxf(r(k), 1e-6f);
        const float u2 = r(k + 1u);
        return sqrtf(-2.0f * fs_log(u1)) * fs_cos(6.2831853f * u2);
    }
};

__device__ __forceinline__ float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float3 add3(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 mul3(float3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float len3(float3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ float3 nrm3(float3 a) { const float l = len3(a); return f3(a.x / l, a.y / l, a.z / l); }

// dead particles per block
__global__ __launch_bounds__(SPARK_BLOCK) void k_spark_count(const float4* __restrict__ sp, int count, int* __restrict__ dead) {
    const int i = blockIdx.x * SPARK_BLOCK + threadIdx.x;
    const bool d = i < count && sp[4 * i + 3].z < 0.5f;
    const unsigned long long m = __ballot(d);
    __shared__ int wc[SPARK_BLOCK / 64];
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < SPARK_BLOCK / 64; w++) t += wc[w];
        dead[blockIdx.x] = t;
    }
}

// exclusive scan of the per-block counts (one workgroup, chunks of 1024)
__global__ __launch_bounds__(1024) void k_spark_scan(const int* __restrict__ v, int* __restrict__ out, int n) {
    __shared__ int buf[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += 1024) {
        const int i = base + (int)threadIdx.x;
        const int x = i < n ? v[i] : 0;
        buf[threadIdx.x] = x;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {   // Hillis-Steele inclusive scan
            const int t = (int)threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < n) out[i] = carry + buf[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == 1023) carry += buf[1023];
        __syncthreads();
    }
}

template <class S>
__global__ __launch_bounds__(SPARK_BLOCK) void k_spark_update(SparkUniform U, float4* __restrict__ sp,
                                                              const int* __restrict__ block_off,
                                                              int* __restrict__ block_dead_next,
                                                              const typename S::raw* __restrict__ vel, Dims g,
                                                              const half4_t* __restrict__ curlT, int nn) {
    const int i = blockIdx.x * SPARK_BLOCK + threadIdx.x;
    const int count = (int)U.count;
    const bool in = i < count;
    float4 P = make_float4(0.f, 0.f, 0.f, 0.f), V = P, Q = P, M = P;
    if (in) { P = sp[4 * i]; V = sp[4 * i + 1]; Q = sp[4 * i + 2]; M = sp[4 * i + 3]; }
    const bool dead = in && M.z < 0.5f;
    // the slot of a dead particle: dead particles before it in index order
    const unsigned long long m = __ballot(dead);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __shared__ int wc[SPARK_BLOCK / 64];
    __shared__ int wn[SPARK_BLOCK / 64];
    if (lane == 0) wc[wv] = __popcll(m);
    __syncthreads();
    const float dt = U.dt;
    bool alive_after = false;
    if (dead) {   // src/sparks.js:84-111
        const unsigned want = (unsigned)(U.nBase + U.nBurst);
        int off = block_off[blockIdx.x];
        for (int w = 0; w < wv; w++) off += wc[w];
        const unsigned slot = (unsigned)(off + __popcll(m & ((1ull << lane) - 1ull)));
        if (want != 0u && slot < want) {
            const Rng R{slot, (unsigned)U.seed, (unsigned)U.frame};
            const bool burst = slot >= (unsigned)U.nBase;
            float3 p, v;
            if (burst) {
                p = add3(f3(U.burstX, U.burstY, U.burstZ),
                         mul3(mul3(sub3(f3(R.r(7u), R.r(8u), R.r(9u)), f3(0.5f, 0.5f, 0.5f)), U.burstRadius), 2.0f));
                float3 radial = f3(p.x, 0.0f, p.z);
                radial = len3(radial) > 1e-3f ? nrm3(radial) : f3(0.f, 0.f, 0.f);
                const float3 aim = f3(U.aimX, 0.0f, U.aimZ);
                const float3 dir = nrm3(add3(f3(R.gauss(1u), fabsf(R.gauss(3u)) * 1.2f + 0.7f, R.gauss(5u)),
                                             mul3(add3(radial, aim), U.burstFan)));
                v = mul3(mul3(dir, U.burstSpeed), 0.5f + 0.8f * R.r(10u));
            } else {
                const float a = R.r(1u) * 6.2831853f;
                const float rr = sqrtf(R.r(2u)) * U.bedRadius;
                p = f3(fs_cos(a) * rr, 0.05f + R.r(3u) * 0.12f, fs_sin(a) * rr);
                v = f3(R.gauss(4u) * 0.4f, 0.6f + R.r(6u) * 1.8f, R.gauss(11u) * 0.4f);
            }
            const float dmm =
                wclamp(0.5f * fs_exp(0.55f * R.gauss(13u)) * U.sizeK * (burst ? U.burstSize : 1.0f), 0.15f, 2.5f);
            const float tb = (0.25f + 2.0f * dmm * dmm) * (0.6f + 0.8f * R.r(15u)) * U.lifeK;
            const float T0 = 1230.0f + 70.0f * R.gauss(16u) + (burst ? 80.0f : 0.0f);
            const float fl = R.r(20u) < 0.45f ? 5.0f + 15.0f * R.r(18u) : 0.0f;
            sp[4 * i] = make_float4(p.x, p.y, p.z, 0.0f);
            sp[4 * i + 1] = make_float4(v.x, v.y, v.z, tb);
            sp[4 * i + 2] = make_float4(p.x, p.y, p.z, T0);
            sp[4 * i + 3] = make_float4(dmm, R.r(19u), 1.0f, fl);
            alive_after = true;
        }
    } else if (in) {   // alive: integrate (src/sparks.js:113-134)
        const float dmm = M.x;
        float3 p = f3(P.x, P.y, P.z), v = f3(V.x, V.y, V.z);
        const float tau = fmaxf(0.4f * dmm * dmm, 0.01f);
        // fluidVel (src/sparks.js:71-79): the step's velocity, or a slowing column above the box
        float3 u;
        const float gx = (p.x - U.ox) / U.sx, gy = (p.y - U.oy) / U.sy, gz = (p.z - U.oz) / U.sz;
        if (gx < 0.f || gy < 0.f || gz < 0.f || gx > 1.f || gy > 1.f || gz > 1.f) {
            u = (gy > 1.0f && sqrtf(p.x * p.x + p.z * p.z) < 0.5f) ? f3(0.0f, 1.6f / (1.0f + (gy - 1.0f) * 1.5f), 0.0f)
                                                                  : f3(0.f, 0.f, 0.f);
        } else {
            const float4 s = sample_clamp<S>(vel, g, gx * (float)g.nx, gy * (float)g.ny, gz * (float)g.nz);
            u = f3(s.x, s.y, s.z);
        }
        u = mul3(u, U.upK);
        const float tqx = p.x * 0.9f + 0.0f, tqy = p.y * 0.9f + -U.time * 0.5f;
        const float tqz = p.z * 0.9f + (U.time * 0.13f + M.y * 3.0f);
        const float turbAmp = U.turb * (0.3f + 3.0f * sstep(0.4f, 1.2f, p.y)) / 12.0f;
        const float4 cn = sample_rep(curlT, nn, tqx, tqy, tqz);
        u = add3(u, mul3(f3(cn.x, cn.y, cn.z), turbAmp));
        const float k = 1.0f - fs_exp(-dt / tau);
        v = f3(mixf(v.x, u.x, k), mixf(v.y, u.y, k), mixf(v.z, u.z, k));
        v.y = v.y - 9.77076f * dt * (dmm < 0.3f ? 0.2f : 1.0f);   // 9.81 (1 - 0.004) rounded once
        const float3 prev = p;
        p = add3(p, mul3(v, dt));
        float T = Q.w;
        const float tb = V.w - dt;
        if (tb > 0.0f) {   // burning
            const float rel = len3(sub3(u, v));
            T = mixf(T, 1180.0f + 60.0f * fminf(rel, 3.0f) + 60.0f * (M.y - 0.5f), 1.0f - fs_exp(-dt / 0.3f));
        } else {           // burned out: cools in ~0.1 s
            T = 293.0f + (T - 293.0f) * fs_exp(-dt / (0.06f + 0.08f * dmm));
        }
        const float age = P.w + dt;
        const bool kill = T < 760.0f || p.y < 0.0f || p.y > 4.5f || sqrtf(p.x * p.x + p.z * p.z) > 3.0f || age > 12.0f;
        sp[4 * i] = make_float4(p.x, p.y, p.z, age);
        sp[4 * i + 1] = make_float4(v.x, v.y, v.z, tb);
        sp[4 * i + 2] = make_float4(prev.x, prev.y, prev.z, T);
        sp[4 * i + 3] = make_float4(M.x, M.y, kill ? 0.0f : 1.0f, M.w);
        alive_after = !kill;
    }
    // the particles of this block dead after the update: the next step's slot counts
    const unsigned long long m2 = __ballot(in && !alive_after);
    if (lane == 0) wn[wv] = __popcll(m2);
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < SPARK_BLOCK / 64; w++) t += wn[w];
        block_dead_next[blockIdx.x] = t;
    }
}

}  // namespace

int spark_blocks(int count) { return (count + SPARK_BLOCK - 1) / SPARK_BLOCK; }

void launch_spark_count(hipStream_t s, const float4* sp, int count, int* dead) {
    k_spark_count<<<spark_blocks(count), SPARK_BLOCK, 0, s>>>(sp, count, dead);
}

void launch_spark_step(int prec, hipStream_t s, const SparkUniform& U, float4* sp, int* dead, int* off, const void* vel,
                       const Dims& g, const half4_t* curl, int nn) {
    const int count = (int)U.count;
    const int nb = spark_blocks(count);
    k_spark_scan<<<1, 1024, 0, s>>>(dead, off, nb);
    switch (prec) {
        case FS_PREC_FP32:
            k_spark_update<StF32><<<nb, SPARK_BLOCK, 0, s>>>(U, sp, off, dead, (const float4*)vel, g, curl, nn);
            break;
        case FS_PREC_FP16:
            k_spark_update<StF16><<<nb, SPARK_BLOCK, 0, s>>>(U, sp, off, dead, (const half4_t*)vel, g, curl, nn);
            break;
        default:
            k_spark_update<StBF16><<<nb, SPARK_BLOCK, 0, s>>>(U, sp, off, dead, (const StBF16::raw*)vel, g, curl, nn);
            break;
    }
}
