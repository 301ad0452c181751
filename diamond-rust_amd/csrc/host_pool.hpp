// A fixed set of host worker threads for the CPU-side loops of a search
// (exact E-values, outfmt-6 rendering). Created on first use and kept for the
// life of the process, so a loop costs a wake-up instead of thread starts.
#pragma once

#include <cstdint>
#include <functional>

namespace dmr {

// Runs f(begin, end) over [0, n) split into contiguous parts of at least
// min_per_part items, on the pool's workers and the calling thread; returns
// when every part is done. Small n runs inline. Calls from several threads are
// served one at a time.
void host_parallel_for(uint64_t n, uint64_t min_per_part,
                       const std::function<void(uint64_t, uint64_t)>& f);

// Threads that host_parallel_for runs on (workers and the caller).
unsigned host_threads();

}  // namespace dmr
