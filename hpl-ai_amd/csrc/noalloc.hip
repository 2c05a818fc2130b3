// noalloc.hip — the residual GEMV for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).
//
// hplai_otf_gemv_f64 takes its n * ceil(n / 2048) FP64 partial sums from a
// hipMallocAsync per call.  hplai_otf_gemv_f64_noalloc runs the same two
// kernels (otf_gemv_partial_kernel, otf_gemv_reduce_kernel) with the partial
// sums in `scratch`, a file-scope device array of kNoAllocScratchBytes that
// every device the library runs on holds until the process ends.  The
// partial kernel writes every element the reduce kernel reads, so nothing
// has to start at zero.  Calls that overlap on one device (two streams, two
// host threads) share the array and race.
#include <mutex>

#include "internal.hpp"

namespace hplai {

constexpr int64_t kNoAllocScratchBytes = int64_t(32) << 20;   // 32 MiB
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

// device address of `scratch` per device, filled once per device by
// noalloc_setup (the array's name in host code is a host-side stand-in)
static void *g_scratch_ptr[kNoAllocMaxDevices] = {};
static std::mutex g_scratch_mu;

int noalloc_setup(int device, void **out)
{
    *out = nullptr;
    HPLAI_CHECK(device >= 0 && device < kNoAllocMaxDevices, "noalloc: device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device] == nullptr) {
        int cur = -1;
        HPLAI_HIP(hipGetDevice(&cur));
        HPLAI_CHECK(cur == device, "noalloc: device %d is not current (%d)", device, cur);
        void *p = nullptr;
        HPLAI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        HPLAI_CHECK(p != nullptr, "noalloc: no device address for the scratch array");
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return HPLAI_OK;
}

}  // namespace hplai

using namespace hplai;

extern "C" int hplai_otf_gemv_f64_noalloc(hplai_ctx ctx, int64_t n, uint64_t seed, const double *d,
                                          double alpha, const double *x, double *y, void *stream)
{
    HPLAI_CHECK(ctx != nullptr, "null hplai_ctx");
    HPLAI_HIP(hipSetDevice(ctx->device));
    HPLAI_CHECK(n > 0 && d != nullptr && x != nullptr && y != nullptr, "otf_gemv_noalloc: bad args");
    HPLAI_CHECK(ctx->scratch != nullptr, "otf_gemv_noalloc: context has no scratch address");
    const int64_t need = static_cast<int64_t>(sizeof(double)) * otf_gemv_work_doubles(n);
    HPLAI_CHECK(need <= kNoAllocScratchBytes,
                "otf_gemv_noalloc: n=%lld needs %lld bytes of partial sums, more than the %lld-byte "
                "scratch array (n <= 92160 fits); use hplai_otf_gemv_f64",
                (long long)n, (long long)need, (long long)kNoAllocScratchBytes);
    return launch_otf_gemv(n, seed, d, alpha, x, y, static_cast<double *>(ctx->scratch),
                           as_stream(stream));
}
