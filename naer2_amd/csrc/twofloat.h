// twofloat.h -- error-free transformations of fp32 arithmetic and double-float (hi + lo) values.
//
// two_sum(a, b) returns s = fl(a + b) -- the very value a plain fp32 addition gives -- and the exact
// rounding error e, a + b = s + e (Knuth, TAOCP vol. 2, 4.2.2); two_prod(a, b) returns p = fl(a b) and
// the exact error with one fused multiply-add (a b = p + e).  A computation written with them therefore
// reproduces the plain fp32 results bit for bit in its hi parts while the lo parts carry what rounding
// dropped.  df_add / df_mul are double-float arithmetic on hi + lo pairs (Dekker 1971), about 44 bits.
// The library builds with -ffp-contract=off, so every other operation here rounds on its own.
#pragma once

#include <hip/hip_runtime.h>

namespace n2h_twofloat {

struct df {
    float hi, lo;
};

__device__ __forceinline__ df two_sum(float a, float b) {
    const float s = a + b;
    const float bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}

__device__ __forceinline__ df fast_two_sum(float a, float b) {  // |a| >= |b| or a == 0
    const float s = a + b;
    return {s, b - (s - a)};
}

__device__ __forceinline__ df two_prod(float a, float b) {
    const float p = a * b;
    return {p, __builtin_fmaf(a, b, -p)};
}

__device__ __forceinline__ df df_add(df a, df b) {
    const df s = two_sum(a.hi, b.hi);
    const float e = s.lo + (a.lo + b.lo);
    return fast_two_sum(s.hi, e);
}

__device__ __forceinline__ df df_mul(df a, df b) {
    const df p = two_prod(a.hi, b.hi);
    const float e = p.lo + (a.hi * b.lo + a.lo * b.hi);
    return fast_two_sum(p.hi, e);
}

}  // namespace n2h_twofloat
