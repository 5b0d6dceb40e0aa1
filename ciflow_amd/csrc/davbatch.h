// Batched Davidson solves (solve_batch.cpp): the C ABI entry cf_solve_batch and
// what the server asks of it.
#pragma once

#include <string>

#include "../../include/ciflow_amd.h"

namespace cf {

// Whether cf_solve_batch takes r; `why` gets the reason when it does not.
bool bd_accepts(const cf_request& r, std::string* why);
// Requests that may share a batch have equal keys (kind, L, nup, ndown and
// basis size).
uint64_t bd_class_key(const cf_request& r);

}  // namespace cf
