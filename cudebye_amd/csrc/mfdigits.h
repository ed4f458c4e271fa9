// The digit encoding of the matrix-core Debye sum (pass 2), shared by the device kernels and
// the host-side check dbx_debug_sinc_digits_host.
//
//   A = 2 N = sum_i e_i 256^i,  e_i in [-128, 127]   (balanced count digits, I of them:
//                                                     2N < 127 256^(I-1))
//   T = sinc(q r) = 3/8 + V / c + err,  c = 3 2^46,  V = rint(c (T - 3/8)) = sum_j d_j 256^j,
//       d_j in [-128, 127], six of them
// T in [-0.21724, 1] puts V in [-0.4442, 0.4688] 2^48, inside the range of six signed digits,
// [-0.5020, 0.4981] 2^48; |err| <= 1 / (2c) < 2^-48, below the error of the FP64 sine that T
// comes from (a few 1e-15), so the integer sum is as accurate as an FP64 one.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace dbx {

constexpr int kMfImax = 5;           // count digits (2N below 127 2^32)
constexpr int kMfJ = 6;              // sinc digits
constexpr double kMfC = 0x1.8p47;    // c = 3 2^46
constexpr double kMfShift = 0.375;   // the digits hold T - 3/8
// 1.5 2^52 (an integer sum lands in the low mantissa bits) + 0x808080808080 (128 in each of the
// six bytes: W = V + 0x808080808080 has the bytes d_j + 128) - c 3/8 (= 9 2^43, the shift)
constexpr double kMfMagic = 0x1.8p52 + (double)0x808080808080ll - 0x1.2p46;

// largest 2N + 1 that I count digits hold
__host__ __device__ constexpr double mf_count_limit(int I)
{
    return 127.0 * (double)(1ull << (8 * (I - 1)));
}

// v_perm_b32: byte b of the result is byte sel.b of the 64-bit value {hi:lo} (selectors 0..7),
// or 0 (selector 0x0C); the host version covers exactly those selectors
__host__ __device__ inline uint32_t mf_perm(uint32_t hi, uint32_t lo, uint32_t sel)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_perm(hi, lo, sel);
#else
    const uint64_t x = ((uint64_t)hi << 32) | lo;
    uint32_t r = 0;
    for (int b = 0; b < 4; b++) {
        const uint32_t s = (sel >> (8 * b)) & 0xFFu;
        const uint32_t v = s < 8 ? (uint32_t)(x >> (8 * s)) & 0xFFu : 0u;
        r |= v << (8 * b);
    }
    return r;
#endif
}

// The sinc word of one bin from its sine sv = sin(Q r), cQ = c / Q and ir = 1 / r:
// fma(sv cQ, ir, magic) rounds c T + magic to an integer whose low 48 mantissa bits are
// W = V + 0x808080808080.  lo: bytes 0..3 of W, each xor 0x80 (= the signed digits d_0..d_3);
// hi: the upper word of the double, bytes 0 and 1 = bytes 4 and 5 of W.
// svc = sv cQ: the sine already scaled (the kernel's recurrence runs on scaled sines)
__host__ __device__ inline void sinc_word_scaled(double svc, double ir, uint32_t &lo, uint32_t &hi)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, fma(svc, ir, kMfMagic));
    lo = (uint32_t)u ^ 0x80808080u;
    hi = (uint32_t)(u >> 32);
}
__host__ __device__ inline void sinc_word(double sv, double cQ, double ir, uint32_t &lo, uint32_t &hi)
{
    sinc_word_scaled(sv * cQ, ir, lo, hi);
}

// Four bins' words -> six digit dwords (dword j: byte b = digit j of bin b): a 4 x 4 byte
// transpose of the low words and a 4 x 2 one of the top digits.
__host__ __device__ inline void pack4(const uint32_t lo[4], const uint32_t hi[4], int dj[kMfJ])
{
    const uint32_t t0 = mf_perm(lo[1], lo[0], 0x06020400u);  // 0b0 1b0 0b2 1b2
    const uint32_t t1 = mf_perm(lo[1], lo[0], 0x07030501u);  // 0b1 1b1 0b3 1b3
    const uint32_t t2 = mf_perm(lo[3], lo[2], 0x06020400u);  // 2b0 3b0 2b2 3b2
    const uint32_t t3 = mf_perm(lo[3], lo[2], 0x07030501u);  // 2b1 3b1 2b3 3b3
    dj[0] = (int)mf_perm(t2, t0, 0x05040100u);
    dj[1] = (int)mf_perm(t3, t1, 0x05040100u);
    dj[2] = (int)mf_perm(t2, t0, 0x07060302u);
    dj[3] = (int)mf_perm(t3, t1, 0x07060302u);
    const uint32_t e01 = mf_perm(hi[1], hi[0], 0x05010400u);  // 0b0 1b0 0b1 1b1
    const uint32_t e23 = mf_perm(hi[3], hi[2], 0x05010400u);  // 2b0 3b0 2b1 3b1
    dj[4] = (int)(mf_perm(e23, e01, 0x05040100u) ^ 0x80808080u);
    dj[5] = (int)(mf_perm(e23, e01, 0x07060302u) ^ 0x80808080u);
}

}  // namespace dbx
