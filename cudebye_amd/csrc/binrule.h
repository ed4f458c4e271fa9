// The pass-1 bin rule on the device (cuDebye-v15.cu:101-115), shared by the histogram
// kernel and the self-test kernels.  Compile with -ffp-contract=off: the sum of squares
// must stay unfused to match the reference's outputs.
#pragma once
#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

namespace dbx {

// Correctly rounded FP32 square root for s == 0 and 2^-96 <= s <= FLT_MAX.
// r = v_sqrt_f32(s) is within 1 ulp of sqrt(s).  With r- and r+ the float neighbours of
// r, sqrt(s) lies below the midpoint (r- + r)/2 exactly when s < r- * r + (r - r-)^2 / 4.
// s and r- * r are both integer multiples of (r - r-) * ulp(r), a unit four or more times
// larger than (r - r-)^2 / 4, so the test reduces to the sign of s - r- * r, which one fma
// yields exactly; likewise sqrt(s) lies above (r + r+)/2 exactly when s - r+ * r > 0.
// s == 0 gives r = 0 (r- is a NaN pattern, both tests fail).  Smaller nonzero s
// (|a - b| < 2^-48 A) can only fall in bin 0, whatever this returns for them; +inf and
// NaN propagate to the bin-range check.  dbx_selftest_sqrt checks every float exhaustively.
__device__ __forceinline__ float sqrt_cr(float s)
{
    const float r = __builtin_amdgcn_sqrtf(s);
    const float rdn = __int_as_float(__float_as_int(r) - 1);
    const float rup = __int_as_float(__float_as_int(r) + 1);
    const float t = (__builtin_fmaf(-rdn, r, s) <= 0.f) ? rdn : r;
    return (__builtin_fmaf(-rup, r, s) > 0.f) ? rup : t;
}

// round-half-away-from-zero of v >= 0 below 2^31: trunc(v) + (fract(v) >= 0.5), fract exact
__device__ __forceinline__ int round_pos(float v)
{
    return (int)v + (__builtin_amdgcn_fractf(v) >= 0.5f ? 1 : 0);
}

// The same rounding by the hardware's convert-to-nearest-with-ties-up instruction
// (v_cvt_rpi_i32_f32: floor(v + 0.5) as an integer, one instruction).  Used only where
// dbx_selftest_round has shown it equal to round_pos for every float in the range.
__device__ __forceinline__ int round_rpi(float v)
{
    int r;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

// distance of one pair in bin units before rounding: the reference's unfused FP32 chain
__device__ __forceinline__ float pair_scaled(float ax, float ay, float az, float bx, float by, float bz,
                                             float rp)
{
    const float dx = bx - ax, dy = by - ay, dz = bz - az;
    const float s = (dx * dx + dy * dy) + dz * dz;
    return sqrt_cr(s) * rp;
}

// bin of one pair; the scaled distance is below 2^31 for every pair the histograms count
// (larger values land past the last bin either way)
__device__ __forceinline__ int pair_bin(float ax, float ay, float az, float bx, float by, float bz,
                                        float rp)
{
    return round_rpi(pair_scaled(ax, ay, az, bx, by, bz, rp));
}

}  // namespace dbx
