// Integral rotation that allocates nothing: for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops, harnesses that
// forbid allocation).  It runs the same quarter-transform kernels as
// cf_rotate_integrals; the one buffer those kernels need besides the inputs
// and outputs -- the ping-pong work array of the four quarter transforms,
// batch*L^4 doubles -- lives in a file-scope device array instead of the
// stream-ordered pool.
//
// The array is a host-side stand-in in host code: its device address differs
// per GPU and is looked up once per device with hipGetSymbolAddress when a
// context is created on that device (cf_ctx_create calls noalloc_setup with
// the device current); only that address is ever passed to kernels.
//
// Every call on one GPU shares that GPU's array, so calls that overlap on one
// GPU (two contexts, or two streams, rotating at the same time) race on it:
// callers serialise them.  The work array is written in full before it is
// read, so nothing has to be cleared between calls.
#include "../../include/ciflow_amd.h"

#include <mutex>
#include <string>

#include "transform.h"

namespace cf {
// Defined with the context handle in capi.cpp.
hipStream_t ctx_stream(const cf_ctx* c);
int ctx_device(const cf_ctx* c);
}  // namespace cf

using namespace cf;

namespace {
// 16 MiB per GPU: batch * L^4 <= 2,097,152 (L = 20: 13 Hamiltonians per call;
// L = 12: 101; L = 24: 6).
constexpr uint64_t kScratchDoubles = uint64_t(2) << 20;
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;
}  // namespace

namespace cf {
// Looks up this device's copy of the work array once; the device must be current.
void noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error(ERR_ARG, "device index beyond the no-alloc table");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw Error(ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = static_cast<double*>(p);
}

uint64_t noalloc_scratch_doubles() { return kScratchDoubles; }
}  // namespace cf

extern "C" int cf_rotate_integrals_noalloc(cf_ctx* ctx, int L, int batch, const double* dh, const double* dV,
                                           const double* dU, double* dh_out, double* dV_out) {
    try {
        CF_REQUIRE(ctx && dh && dV && dU && dh_out && dV_out, "NULL argument");
        CF_REQUIRE(L >= 1 && L <= kMaxL && batch >= 1, "bad L / batch");
        const uint64_t L2 = (uint64_t)L * L;
        const uint64_t need = (uint64_t)batch * L2 * L2;
        if (need > kScratchDoubles)
            throw Error(ERR_ARG, "cf_rotate_integrals_noalloc: batch*L^4 = " + std::to_string(need) +
                                     " doubles exceeds the " + std::to_string(kScratchDoubles) +
                                     "-double (16 MiB) work array; split the batch");
        const int dev = ctx_device(ctx);
        CF_HIP(hipSetDevice(dev));
        double* tmp = nullptr;
        {
            std::lock_guard<std::mutex> g(g_scratch_mu);
            tmp = (dev >= 0 && dev < kMaxDevices) ? g_scratch[dev] : nullptr;
        }
        if (!tmp) throw Error(ERR_STATE, "no-alloc work array not set up on this device (create a context first)");
        rotate_integrals_dev(L, batch, dh, dV, dU, dh_out, dV_out, tmp, ctx_stream(ctx));
    } catch (const Error& e) {
        return set_last_error(e.code, e.what());
    } catch (const std::exception& e) {
        return set_last_error(ERR_HIP, e.what());
    }
    return CF_OK;
}
