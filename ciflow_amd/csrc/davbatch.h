// Batched Davidson solves (solve_batch.cpp): the C ABI entry cf_solve_batch and
// what the server asks of it.
#pragma once

#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ciflow_amd.h"

namespace cf {

// Whether cf_solve_batch takes r; `why` gets the reason when it does not.
bool bd_accepts(const cf_request& r, std::string* why);
// Requests that may share a batch have equal keys (kind, L, nup, ndown and
// basis size).
uint64_t bd_class_key(const cf_request& r);
// How a running batch gets more requests and hands back results.
struct BatchFeed {
    // Up to `max` more requests of the batch's shape for its free slots, as
    // (id, request); each request's arrays stay valid until done(id) (may be empty).
    std::function<void(int max, std::vector<std::pair<int, cf_request>>& out)> more;
    // Request `id` has finished; its result is final (called before the rest
    // of the batch ends).
    std::function<void(int id, const cf_result& r)> done;
};
// One lock-step batch of `capacity` slots (<= 32) of requests of one shape:
// `first` at once, then whatever feed.more adds as slots free up.  Results go
// to feed.done; a non-OK status means the batch stopped (requests not yet
// reported have no result).
int solve_batch_feed(cf_ctx* ctx, int capacity, const std::vector<std::pair<int, cf_request>>& first,
                     const BatchFeed& feed);

}  // namespace cf
