// Self tests of the device arithmetic behind pass 1, callable through the C ABI:
//  - dbx_selftest_sqrt: sqrt_cr against the exact rounding criterion, evaluated in FP64
//    (exact there: every float and every midpoint between neighbouring floats squares
//    exactly into 53 bits), for every float bit pattern in a range;
//  - dbx_selftest_bins: the device bin rule on caller-given pairs.
#include <hip/hip_runtime.h>

#include "binrule.h"
#include "fastmath.h"
#include "internal.h"

#pragma clang fp contract(off)

namespace dbx {

__global__ void st_sqrt_kernel(uint32_t lo, uint32_t hi, unsigned long long *bad)
{
    unsigned long long nbad = 0, nmis = 0, nlo = 0, nhi = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t u = (uint64_t)lo + blockIdx.x * blockDim.x + threadIdx.x; u < hi; u += stride) {
        const float s = __uint_as_float((uint32_t)u);
        if (s == 0.f) {
            nbad += __float_as_uint(sqrt_cr(s)) != 0u;
            continue;
        }
        if (!(s >= 0x1p-96f) || !(s <= 3.4028235e38f)) continue;
        const float r = sqrt_cr(s);
        const float raw = __builtin_amdgcn_sqrtf(s);
        nlo += __float_as_uint(raw) < __float_as_uint(r);
        nhi += __float_as_uint(raw) > __float_as_uint(r);
        const double R = r;
        const double Rd = __uint_as_float(__float_as_uint(r) - 1);
        const double Ru = __uint_as_float(__float_as_uint(r) + 1);
        const double mdn = 0.5 * (R + Rd), mup = 0.5 * (R + Ru);
        const double sd = s;
        if (!(mdn * mdn < sd && sd < mup * mup)) nbad++;
        if (__float_as_uint(r) != __float_as_uint(__builtin_sqrtf(s))) nmis++;
    }
    if (nbad) atomicAdd(&bad[0], nbad);
    if (nmis) atomicAdd(&bad[1], nmis);
    if (nlo) atomicAdd(&bad[2], nlo);
    if (nhi) atomicAdd(&bad[3], nhi);
}

// n_bad[0]: floats v in [lo, hi) (bit patterns, v >= 0) where round_rpi(v) != round_pos(v)
__global__ void st_round_kernel(uint32_t lo, uint32_t hi, unsigned long long *bad)
{
    unsigned long long nbad = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t u = (uint64_t)lo + blockIdx.x * blockDim.x + threadIdx.x; u < hi; u += stride) {
        const float v = __uint_as_float((uint32_t)u);
        if (!(v >= 0.f) || !(v < 2147483520.f)) continue;
        if (round_rpi(v) != round_pos(v)) nbad++;
    }
    if (nbad) atomicAdd(&bad[0], nbad);
}

// LDS atomic throughput: every lane adds 1 to a u32 of a 32768-word window, lanes of a wave on
// consecutive words (no bank conflict), `iters` times at moving offsets
__global__ __launch_bounds__(1024) void st_lds_atomic_kernel(int iters, unsigned int *out)
{
    __shared__ uint32_t h[32768];
    for (int i = threadIdx.x; i < 32768; i += 1024) h[i] = 0u;
    __syncthreads();
    const unsigned base = (threadIdx.x >> 6) * 2048u + (threadIdx.x & 63);
    for (int it = 0; it < iters; it++) atomicAdd(&h[(base + (unsigned)it * 64u) & 32767u], 1u);
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, h[blockIdx.x & 32767u]);
}

__global__ void st_bins_kernel(const float *__restrict__ a, const float *__restrict__ b, int n,
                               float rp, int *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = pair_bin(a[3 * i], a[3 * i + 1], a[3 * i + 2], b[3 * i], b[3 * i + 1], b[3 * i + 2], rp);
}

__global__ void st_sin_kernel(const double *__restrict__ x, int n, double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = sin_cw(x[i]);
}

}  // namespace dbx

using namespace dbx;

// out[i] = the device sine of the Debye sum (sin_cw) at x[i], |x| < 2^20.
extern "C" int dbx_selftest_sin(const double *x, int n, double *out)
{
    if (!x || !out || n < 0) return DBX_ERR_ARG;
    if (n == 0) return DBX_OK;
    double *d = nullptr;
    if (hipMalloc(&d, sizeof(double) * 2 * (size_t)n) != hipSuccess) return DBX_ERR_HIP;
    (void)hipMemcpy(d, x, sizeof(double) * n, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(st_sin_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, d, n, d + n);
    hipError_t e = hipMemcpy(out, d + n, sizeof(double) * n, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return e == hipSuccess ? DBX_OK : DBX_ERR_HIP;
}

// Checks sqrt_cr for every float bit pattern in [lo, hi) with 2^-96 <= s <= FLT_MAX
// (plus s == 0 when lo == 0, which must give +0).  n_bad[0]: results that are not correctly rounded;
// n_bad[1]: results that differ from the compiler's correctly rounded sqrtf; n_bad[2] and
// n_bad[3]: inputs where the bare v_sqrt_f32 result lies below / above the correctly rounded one.
extern "C" int dbx_selftest_sqrt(uint32_t lo, uint32_t hi, uint64_t *n_bad)
{
    if (!n_bad || hi < lo) return DBX_ERR_ARG;
    unsigned long long *d = nullptr;
    if (hipMalloc(&d, 4 * sizeof(*d)) != hipSuccess) return DBX_ERR_HIP;
    (void)hipMemset(d, 0, 4 * sizeof(*d));
    hipLaunchKernelGGL(st_sqrt_kernel, dim3(8192), dim3(256), 0, 0, lo, hi, d);
    unsigned long long h[4] = {0, 0, 0, 0};
    hipError_t e = hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return DBX_ERR_HIP;
    n_bad[0] = h[0];
    n_bad[1] = h[1];
    n_bad[2] = h[2];
    n_bad[3] = h[3];
    return DBX_OK;
}

// out[i] = bin of the pair (a[3i..3i+2], b[3i..3i+2]) under the device bin rule.
extern "C" int dbx_selftest_bins(const float *a, const float *b, int n, float rp, int32_t *out)
{
    if (!a || !b || !out || n < 0) return DBX_ERR_ARG;
    if (n == 0) return DBX_OK;
    float *da = nullptr, *db = nullptr;
    int *dout = nullptr;
    if (hipMalloc(&da, sizeof(float) * 3 * n) != hipSuccess) return DBX_ERR_HIP;
    if (hipMalloc(&db, sizeof(float) * 3 * n) != hipSuccess) { (void)hipFree(da); return DBX_ERR_HIP; }
    if (hipMalloc(&dout, sizeof(int) * n) != hipSuccess) { (void)hipFree(da); (void)hipFree(db); return DBX_ERR_HIP; }
    (void)hipMemcpy(da, a, sizeof(float) * 3 * n, hipMemcpyHostToDevice);
    (void)hipMemcpy(db, b, sizeof(float) * 3 * n, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(st_bins_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, da, db, n, rp, dout);
    hipError_t e = hipMemcpy(out, dout, sizeof(int) * n, hipMemcpyDeviceToHost);
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dout);
    return e == hipSuccess ? DBX_OK : DBX_ERR_HIP;
}

// Counts the floats v in [lo, hi) (bit patterns; 0 <= v < 2^31) where the hardware's
// round-half-up conversion differs from the reference's roundf (half away from zero).
extern "C" int dbx_selftest_round(uint32_t lo, uint32_t hi, uint64_t *n_bad)
{
    if (!n_bad || hi < lo) return DBX_ERR_ARG;
    unsigned long long *d = nullptr;
    if (hipMalloc(&d, sizeof(*d)) != hipSuccess) return DBX_ERR_HIP;
    (void)hipMemset(d, 0, sizeof(*d));
    hipLaunchKernelGGL(st_round_kernel, dim3(8192), dim3(256), 0, 0, lo, hi, d);
    unsigned long long h = 0;
    hipError_t e = hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return DBX_ERR_HIP;
    *n_bad = h;
    return DBX_OK;
}

// Measured LDS atomic rate of the device (ds_add_u32 per second over all CUs), from one launch of
// 4 workgroups per CU, 1024 lanes each, 4096 adds per lane, timed with HIP events.
extern "C" int dbx_selftest_lds_atomic_rate(double *atomics_per_s)
{
    if (!atomics_per_s) return DBX_ERR_ARG;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return DBX_ERR_HIP;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return DBX_ERR_HIP;
    unsigned int *d = nullptr;
    if (hipMalloc(&d, sizeof(unsigned int)) != hipSuccess) return DBX_ERR_HIP;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const int iters = 4096, blocks = 4 * cus;
    hipLaunchKernelGGL(st_lds_atomic_kernel, dim3(blocks), dim3(1024), 0, 0, 64, d);  // warm-up
    (void)hipEventRecord(e0, 0);
    hipLaunchKernelGGL(st_lds_atomic_kernel, dim3(blocks), dim3(1024), 0, 0, iters, d);
    (void)hipEventRecord(e1, 0);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(d);
    if (e != hipSuccess || ms <= 0.f) return DBX_ERR_HIP;
    *atomics_per_s = (double)blocks * 1024.0 * iters / (ms * 1e-3);
    return DBX_OK;
}
