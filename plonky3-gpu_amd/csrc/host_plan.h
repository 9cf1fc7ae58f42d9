// Host-only accounting of the runtime: allocator size classes, the device
// footprint of a served commit request, and the serving queue's admission
// rule.  No HIP dependency, so the sanitizer build of the host planners
// (tests/cpp/planner_sanitize.cpp) compiles the same code the runtime runs.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace p3h {

// Allocator size class: powers of two from 256 bytes up to 2 MiB, then
// multiples of 2 MiB.
inline size_t size_class(size_t bytes) {
    if (bytes == 0) bytes = 1;
    if (bytes <= (size_t(2) << 20)) {
        size_t c = 256;
        while (c < bytes) c <<= 1;
        return c;
    }
    const size_t g = size_t(2) << 20;
    return (bytes + g - 1) / g * g;
}

// Device bytes a served BabyBear commit of an h x w trace with k added bits
// holds until it is released, in allocator size classes: the trace, its LDE
// (h << k rows), every digest layer (32 bytes per node), the leaf-row table
// and a small fixed allowance for per-request tables.
inline uint64_t serve_footprint(uint64_t h, uint32_t w, uint32_t k) {
    const uint64_t N = h << k;
    uint64_t tree = 0;
    for (uint64_t c = N; c >= 1; c >>= 1) {
        tree += size_class(c * 32);
        if (c == 1) break;
    }
    return size_class(h * w * 4) + size_class(N * w * 4) + tree + size_class(N * 4) + 4096;
}

// Admission: a worker takes queued requests in arrival order while the next
// one fits the budget next to what is already reserved, at most max_batch;
// the first request that does not fit stops the batch (no overtaking).
// footprints[0..n) are the queued requests' footprints, oldest first; returns
// how many to admit.  reserved <= budget.
inline size_t serve_admit(uint64_t reserved, uint64_t budget, size_t max_batch, const uint64_t* footprints,
                          size_t n) {
    size_t taken = 0;
    while (taken < n && taken < max_batch && footprints[taken] <= budget - reserved) {
        reserved += footprints[taken];
        ++taken;
    }
    return taken;
}

}  // namespace p3h
