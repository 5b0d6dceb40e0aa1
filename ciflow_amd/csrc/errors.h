// Status codes and the error type carried across the C ABI; no HIP
// dependency, so host-only sources (parsers, the reverse-communication
// Lanczos) build with a plain C++ compiler.
#pragma once

#include <stdexcept>
#include <string>

namespace cf {

// Error carrying a status code across the C ABI.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Records msg as the calling thread's cf_last_error() text; returns code.
int set_last_error(int code, const std::string& msg);

enum Status {
    OK = 0,
    ERR_ARG = 1,
    ERR_HIP = 2,
    ERR_NOMEM = 3,
    ERR_NOCONV = 4,
    ERR_UNSUPPORTED = 5,
    ERR_STATE = 6,
};

#define CF_REQUIRE(cond, msg)                                  \
    do {                                                       \
        if (!(cond)) throw ::cf::Error(::cf::ERR_ARG, (msg));  \
    } while (0)

constexpr int kMaxL = 64;
constexpr int kBinomCols = 33;  // m <= 32 electrons per spin string

}  // namespace cf
