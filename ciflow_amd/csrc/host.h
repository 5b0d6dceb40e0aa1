// Host-only pieces of the library shared with capi.cpp (no HIP dependency).
#pragma once

#include <vector>

namespace cf {
// Sign convention of returned eigenvectors: largest-magnitude component positive.
void sign_fix(std::vector<double>& x);
}  // namespace cf
