// Division by a constant through its reciprocal, for the smoothsteps whose edges are literals.
//
// z = RN(1/d); q = RN(x z) is within an ulp of x / d, and one Markstein correction (the remainder
// r = x - q d is exact under FMA, then RN(q + r z)) gives the correctly rounded x / d while nothing
// underflows (Markstein 1990; Muller et al., Handbook of Floating-Point Arithmetic).  The callers
// (sstep_c) divide numerators clamped to [-d, 2d] and only use t t (3 - 2 t) of the clamped quotient;
// tests/test_const_division.py checks, for every edge pair used and every f32 numerator of that
// interval, that this smoothstep equals the one from IEEE division bit for bit (the quotient itself
// differs only for |x| < 1e-30, where t t underflows to 0 either way).  Three VALU instructions instead
// of the ten of a general correctly rounded division.  Plain C++, so that test compiles this function.
#pragma once

__host__ __device__ __forceinline__ float fs_div_recip(float x, float d, float z) {
    const float q = x * z;
    const float r = fmaf(-q, d, x);
    return fmaf(r, z, q);
}
