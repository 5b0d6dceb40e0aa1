// Shared definitions for the ciflow_amd library (host + device).
#pragma once

#include <hip/hip_runtime.h>

#include "errors.h"

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace cf {

typedef uint64_t str_t;  // one spin string: bit k = spatial orbital k occupied

#define CF_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw ::cf::Error(e_ == hipErrorOutOfMemory ? ::cf::ERR_NOMEM : ::cf::ERR_HIP,    \
                              std::string(#call) + ": " + hipGetErrorString(e_));             \
    } while (0)

// Stream that device buffers allocated on this host thread are ordered on:
// a context makes its stream current here, and its buffers then come from the
// device's stream-ordered pool (hipMallocAsync / hipFreeAsync), so allocating
// and freeing neither blocks the host nor waits for other streams.  Null:
// plain hipMalloc / hipFree.
inline thread_local hipStream_t tls_alloc_stream = nullptr;

// Owning device buffer.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    hipStream_t s = nullptr;  // allocation stream (null: hipMalloc)
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n), s(o.s) {
        o.p = nullptr;
        o.n = 0;
        o.s = nullptr;
    }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            s = o.s;
            o.p = nullptr;
            o.n = 0;
            o.s = nullptr;
        }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count) {
            if (tls_alloc_stream) {
                CF_HIP(hipMallocAsync((void**)&p, count * sizeof(T), tls_alloc_stream));
                s = tls_alloc_stream;
            } else {
                CF_HIP(hipMalloc((void**)&p, count * sizeof(T)));
            }
        }
        n = count;
    }
    void release() {
        if (p) {
            if (s) (void)hipFreeAsync(p, s);
            else (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
        s = nullptr;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Binomial coefficients C(k, m) for 0 <= k <= 64, 0 <= m <= 64 (exact while < 2^64).
inline uint64_t binom_host(int n, int k) {
    if (k < 0 || k > n) return 0;
    unsigned __int128 r = 1;
    for (int i = 1; i <= k; ++i) r = r * (unsigned)(n - k + i) / (unsigned)i;
    return (uint64_t)r;
}


}  // namespace cf
