// The allocation-free pass 2 (dbx_debye_noalloc): the Debye sum of dbx_debye on device
// histograms, with every buffer its kernels use taken from one fixed device region, so that
// the call allocates nothing (no hipMalloc, hipFree or pinned host allocation).
//
// The region is a file-scope device array.  Each device holds its own copy for as long as the
// library stays loaded; its address on a device is looked up once, with that device current
// (noalloc_init, called by dbx_open), and only that address is ever handed to kernels or
// copies.  Every call on one device shares that device's region: calls that overlap on one
// device (two contexts, two threads) race; one context serialises its own calls.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "internal.h"

// 64 MiB: the C2 sphere (3 partials, 106k bins, 4096 Q) takes about 10 MiB, the reference's
// Gamma example (6 partials, 3.56M bins of which 120k are occupied, 6501 2-theta points) about
// 52 MiB in FP64 mode and 32 MiB on the matrix cores.
constexpr size_t kNoAllocBytes = (size_t)64 << 20;

// the fixed region: in host code this name stands for the device array, not its address
static __device__ __align__(256) char scratch[kNoAllocBytes];

namespace dbx {

constexpr int kNoAllocDevices = 64;

static std::mutex g_mu;
static char *g_scratch[kNoAllocDevices];  // the array's device address on each device (null: not looked up)

void noalloc_init(int device)
{
    DBX_REQUIRE(device >= 0 && device < kNoAllocDevices, DBX_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return;
    void *p = nullptr;
    DBX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));
    DBX_REQUIRE(p != nullptr, DBX_ERR_HIP, "no device address for the fixed scratch region");
    g_scratch[device] = (char *)p;
}

namespace {

// pass 2's buffers carved out of the fixed region in order, 256-byte aligned.  A problem whose
// buffers do not fit is refused with DBX_ERR_BUDGET at the first carve that exceeds the region
// (pass 2 sizes its largest buffers from upper bounds when they fit the room left, else from
// the compacted-bin count it reads back first).
struct FixedMem : P2Mem {
    char *base;
    size_t used = 0;
    unsigned long long *rb;
    FixedMem(char *b, unsigned long long *r) : base(b), rb(r) {}
    void *dev(int, size_t bytes) override
    {
        const size_t a = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
        if (a > kNoAllocBytes - used)
            throw Error(DBX_ERR_BUDGET, "dbx_debye_noalloc: the problem needs more than the " +
                                            std::to_string(kNoAllocBytes >> 20) +
                                            " MiB of the fixed scratch region (use dbx_debye)");
        void *p = base + used;
        used += a;
        return p;
    }
    size_t room() const override { return kNoAllocBytes - used; }
    void *staging(size_t bytes) override { return bytes <= kReadbackBytes ? rb : nullptr; }
};

}  // namespace

}  // namespace dbx

using namespace dbx;

extern "C" {

int dbx_debye_noalloc(dbx_ctx *ctx, const uint64_t *d_hist, int P, int64_t n_bins, double rstep, int first_n,
                      const double *d_q, int nq, const double *d_ffoobb, double *d_I, double *d_S,
                      double *d_rel_err)
{
    if (!d_hist || !d_q || !d_ffoobb || !d_I || P <= 0 || nq <= 0 || n_bins <= 0 || first_n < 0 ||
        !(rstep > 0))
        return DBX_ERR_ARG;
    return guarded(ctx, [&] {
        char *base = g_scratch[ctx->device];
        DBX_REQUIRE(base != nullptr, DBX_ERR_STATE, "fixed scratch region not set up on this device");
        FixedMem mem(base, (unsigned long long *)ctx->p2_readback_h);
        // the contraction profile (a debugging aid) reads back through the context's growing
        // slots: off for this call; and its waits free no buffer that an earlier call retired
        struct NoAllocScope {
            dbx_ctx *c;
            bool was;
            ~NoAllocScope()
            {
                c->profiling = was;
                c->no_alloc = false;
            }
        } np{ctx, ctx->profiling};
        ctx->profiling = false;
        ctx->no_alloc = true;
        // the structure sums S[p][q], their bounds and the bound of I: the caller's buffers, or
        // the region's
        double *S = d_S ? d_S : (double *)mem.dev(kSlotQ, sizeof(double) * (size_t)P * (size_t)nq);
        double *rel = d_rel_err ? d_rel_err : (double *)mem.dev(kSlotQ, sizeof(double) * (size_t)nq);
        double *B = (double *)mem.dev(kSlotQ, sizeof(double) * (size_t)P);
        unsigned long long *worst = (unsigned long long *)mem.dev(kSlotQ, sizeof(unsigned long long));
        const size_t kept = mem.used;  // the pass-2 buffers above this mark are reused below
        ctx->last_p2_refined = 0;
        debye_structure_mem(ctx, mem, d_hist, P, n_bins, rstep, first_n, d_q, nq, S, B);
        launch_combine(ctx->stream, d_ffoobb, S, B, nullptr, 1, nq, d_I, rel, make_int4(0, P, 0, 0));
        if (ctx->pass2_mode == DBX_PASS2_MFMA && ctx->rel_err_target > 0) {
            // the largest bound, read back; past the target the partials are redone in FP64
            launch_max_bits(ctx->stream, rel, nq, worst);
            DBX_HIP(hipMemcpyAsync(ctx->p2_readback_h, worst, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                   ctx->stream));
            sync_pass2(ctx);
            double w;
            std::memcpy(&w, ctx->p2_readback_h, sizeof(w));
            if (exceeds_target(w, ctx->rel_err_target)) {
                mem.used = kept;
                P2Refining mark(ctx);
                debye_structure_mem(ctx, mem, d_hist, P, n_bins, rstep, first_n, d_q, nq, S, B, DBX_PASS2_FP64);
                launch_combine(ctx->stream, d_ffoobb, S, B, nullptr, 1, nq, d_I, rel, make_int4(0, P, 0, 0));
                ctx->last_p2_refined = 1;
            }
        }
        sync_pass2(ctx);
    });
}

}  // extern "C"
