// Exhaustive check of the literal-edge smoothstep (sstep_c: numerator clamped to [-d, 2d], then divided
// through fs_div_recip of procedural-fire_amd/csrc/fs_divc.h) against the plain one (IEEE division).
// Arguments: pairs "e0 e1" of the literal edges as written in the sources (e.g. 0.03f 0.4f), optionally
// preceded by --plain (divides by RN(x z) alone, without the correction: a planted mistake the check must
// catch).  For each pair, d = e1 - e0 in f32 and z = RN(1/d), and every f32 numerator n in [-d, 2d] goes
// through t = clamp(n / d, 0, 1), t t (3 - 2 t) both ways.  Prints "e0 e1 d checked smoothstep_mismatches
// quotient_mismatches max|n|_of_quotient_mismatches" per pair; exits 1 on any smoothstep mismatch.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#define __host__
#define __device__
#define __forceinline__ inline
#include "fs_divc.h"

static inline float bits_f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
static inline uint32_t f_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static inline float smooth(float q) {          // WGSL smoothstep after the division
    const float t = fminf(fmaxf(q, 0.f), 1.f);
    return t * t * (3.f - 2.f * t);
}

// branch-free pass over one chunk (vectorizes): numerators whose quotient bits differ
template <bool PLAIN>
static int chunk_quotient_diffs(uint32_t u0, int cnt, float d, float z) {
    int bad = 0;
    for (int j = 0; j < cnt; j++) {
        const float n = bits_f(u0 + (uint32_t)j);
        const float want = n / d;
        const float got = PLAIN ? n * z : fs_div_recip(n, d, z);
        bad += f_bits(want) != f_bits(got);
    }
    return bad;
}

int main(int argc, char** argv) {
    int a = 1;
    bool plain = false;
    if (a < argc && std::strcmp(argv[a], "--plain") == 0) { plain = true; a++; }
    int bad_pairs = 0;
    const int CH = 1 << 16;
    for (; a + 1 < argc; a += 2) {
        const float e0 = std::strtof(argv[a], nullptr), e1 = std::strtof(argv[a + 1], nullptr);
        const float d = e1 - e0;
        volatile float one = 1.f;
        const float z = one / d;
        // bit patterns are monotone in magnitude: +0 .. 2d and -0 .. -d
        const uint32_t top[2] = {f_bits(2.f * d), f_bits(d)};
        long long total = 0, bad_s = 0, bad_q = 0;
        float max_q = 0.f;
        for (int sign = 0; sign < 2; sign++) {
            const uint32_t sb = sign ? 0x80000000u : 0u;
            const long long nchunks = ((long long)top[sign] + CH) / CH;
            total += (long long)top[sign] + 1;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : bad_s, bad_q) reduction(max : max_q)
            for (long long c = 0; c < nchunks; c++) {
                const uint32_t u0 = (uint32_t)(c * CH);
                const int cnt = (int)std::min<long long>(CH, (long long)top[sign] + 1 - c * CH);
                const int diffs = plain ? chunk_quotient_diffs<true>(sb | u0, cnt, d, z)
                                        : chunk_quotient_diffs<false>(sb | u0, cnt, d, z);
                if (!diffs) continue;
                for (int j = 0; j < cnt; j++) {          // rare: redo the chunk through the smoothstep
                    const float n = bits_f(sb | (u0 + (uint32_t)j));
                    const float want = n / d;
                    const float got = plain ? n * z : fs_div_recip(n, d, z);
                    if (f_bits(want) == f_bits(got)) continue;
                    if (!(want == 0.f && got == 0.f)) { bad_q++; max_q = fmaxf(max_q, fabsf(n)); }
                    bad_s += f_bits(smooth(want)) != f_bits(smooth(got));
                }
            }
        }
        std::printf("%s %s %.9g %lld %lld %lld %g\n", argv[a], argv[a + 1], (double)d, total, bad_s, bad_q, (double)max_q);
        if (bad_s) bad_pairs++;
    }
    return bad_pairs ? 1 : 0;
}
