// lbvh.h — standalone self-contact broadphase (linear BVH with a Karras
// hierarchy) over device arrays.
#pragma once

#include "common.h"

namespace cloth {

size_t lbvh_scratch_bytes(uint32_t m);

// selfPairs over device arrays x0, x1 (3n FP64), idx (3m u32): writes up to
// `cap` sorted pairs (tA, tB) to `pairs` (2 u32 each) and the pair count to
// *count.  Returns false when the count exceeds `cap` (nothing written).
// Throws std::invalid_argument on a corner index >= n.  Synchronous on the
// null stream.  Allocates its scratch for the call and optionally reports the
// hierarchy build and the query times (HIP events).
bool lbvh_self_pairs(const double* x0, const double* x1, uint32_t n, const uint32_t* idx, uint32_t m, double pad,
                     uint32_t* pairs, uint64_t cap, uint64_t* count, double* build_ms, double* query_ms);

}  // namespace cloth
