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

// the arguments sin_cw and sincos_cw hold for: |x| below this
constexpr double kSinArgMax = 0x1p20;

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

// sin and cos of x, |x| < 2^20, from one reduction: x = n pi + y, sin x = (-1)^n sin y,
// cos x = (-1)^n cos y, Taylor series on |y| <= pi/2 through degree 23 (sin) and 24 (cos).
__host__ __device__ inline void sincos_cw(double x, double &s_out, double &c_out)
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
    const double sy = fma(y * y2, ps, y);
    double pc = 0x1.f2cf01972f578p-80;         // 1/24!
    pc = fma(pc, y2, -0x1.0ce396db7f853p-70);  // -1/22!
    pc = fma(pc, y2, 0x1.e542ba4020225p-62);   // 1/20!
    pc = fma(pc, y2, -0x1.6827863b97d97p-53);  // -1/18!
    pc = fma(pc, y2, 0x1.ae7f3e733b81fp-45);   // 1/16!
    pc = fma(pc, y2, -0x1.93974a8c07c9dp-37);  // -1/14!
    pc = fma(pc, y2, 0x1.1eed8eff8d898p-29);   // 1/12!
    pc = fma(pc, y2, -0x1.27e4fb7789f5cp-22);  // -1/10!
    pc = fma(pc, y2, 0x1.a01a01a01a01ap-16);   // 1/8!
    pc = fma(pc, y2, -0x1.6c16c16c16c17p-10);  // -1/6!
    pc = fma(pc, y2, 0x1.5555555555555p-5);    // 1/4!
    pc = fma(pc, y2, -0x1.0000000000000p-1);   // -1/2!
    const double cy = fma(y2, pc, 1.0);
    const bool odd = ((int)n & 1) != 0;
    s_out = odd ? -sy : sy;
    c_out = odd ? -cy : cy;
}

}  // namespace dbx
