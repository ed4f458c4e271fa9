// FP64 sine for the Debye sum's sinc table.
//
// sin_cw(x), |x| < 2^20: Cody-Waite reduction x = n pi + y with pi split into three parts
// P1 + P2 + P3 (33, 33 and 53 significant bits, so n P1 and n P2 are exact products for
// |n| < 2^19 and the residual of pi is below 1e-36), then sin x = (-1)^n sin y with the
// Taylor series of sin on |y| <= pi/2 through degree 23 (first omitted term below 6e-21).
// One polynomial and no quadrant select keep the instruction count low.  Absolute error
// ~2e-16; the device self test (dbx_selftest_sin) checks it against a 50-digit reference.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace dbx {

__host__ __device__ inline double sin_cw(double x)
{
    const double inv_pi = 0x1.45f306dc9c883p-2;
    const double P1 = 0x1.921fb54400000p+1;
    const double P2 = 0x1.0b4611a600000p-33;
    const double P3 = 0x1.3198a2e037073p-68;
    const double n = rint(x * inv_pi);
    double y = fma(-n, P1, x);
    y = fma(-n, P2, y);
    y = fma(-n, P3, y);
    const double y2 = y * y;
    // sin y = y + y^3 * ps(y^2)
    double ps = -0x1.761b41316381ap-75;        // -1/23!
    ps = fma(ps, y2, 0x1.71b8ef6dcf572p-66);   // 1/21!
    ps = fma(ps, y2, -0x1.2f49b46814157p-57);  // -1/19!
    ps = fma(ps, y2, 0x1.952c77030ad4ap-49);   // 1/17!
    ps = fma(ps, y2, -0x1.ae7f3e733b81fp-41);  // -1/15!
    ps = fma(ps, y2, 0x1.6124613a86d09p-33);   // 1/13!
    ps = fma(ps, y2, -0x1.ae64567f544e4p-26);  // -1/11!
    ps = fma(ps, y2, 0x1.71de3a556c734p-19);   // 1/9!
    ps = fma(ps, y2, -0x1.a01a01a01a01ap-13);  // -1/7!
    ps = fma(ps, y2, 0x1.1111111111111p-7);    // 1/5!
    ps = fma(ps, y2, -0x1.5555555555555p-3);   // -1/3!
    const double s = fma(y * y2, ps, y);
    return ((int)n & 1) ? -s : s;
}

}  // namespace dbx
