// Scalar helpers with the reference's (x86-64 host) semantics, shared by
// the device code of the gridder and of the weighting step.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>

namespace gw {

__device__ __forceinline__ float d_round(float x) { return roundf(x); }
__device__ __forceinline__ double d_round(double x) { return round(x); }
__device__ __forceinline__ float d_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double d_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float d_fabs(float x) { return fabsf(x); }
__device__ __forceinline__ double d_fabs(double x) { return fabs(x); }

// (int)x of an integral-valued x: cvttss2si / cvttsd2si return INT_MIN for
// NaN and for values outside the int range.
__device__ __forceinline__ int to_int_x86(float x)
{
    return (x >= -2147483648.0f && x < 2147483648.0f) ? (int)x : INT_MIN;
}
__device__ __forceinline__ int to_int_x86(double x)
{
    return (x >= -2147483648.0 && x < 2147483648.0) ? (int)x : INT_MIN;
}

// (size_t)x for x >= 0 or NaN, as gcc lowers it on x86-64: values below
// 2^63 convert directly, NaN gives 2^63, larger values go through
// x - 2^63 with bit 63 flipped, and the indefinite result of an out-of-range
// x - 2^63 (e.g. +inf) flips to 0.
template <typename T>
__device__ __forceinline__ unsigned long long to_size_x86(T x)
{
    const T two63 = (T)9223372036854775808.0;
    if (x < two63) return (unsigned long long)(long long)x;
    if (x != x) return 0x8000000000000000ull;
    const T t = x - two63;
    if (t < two63)
        return ((unsigned long long)(long long)t) ^ 0x8000000000000000ull;
    return 0ull;
}

}  // namespace gw
