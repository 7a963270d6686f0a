// lbvh.h — standalone self-contact broadphase (linear BVH with a Karras
// hierarchy) over device arrays.
#pragma once

#include "common.h"

namespace cloth {

// Largest triangle count of the allocation-free variant (its scratch is a
// fixed device array of lbvh_noalloc_bytes(), about 69 MB per GPU).
constexpr uint32_t kLbvhMaxTris = 1u << 19;

size_t lbvh_scratch_bytes(uint32_t m);
size_t lbvh_noalloc_bytes();

// selfPairs over device arrays x0, x1 (3n FP64), idx (3m u32): writes up to
// `cap` sorted pairs (tA, tB) to `pairs` (2 u32 each) and the pair count to
// *count.  Returns false when the count exceeds `cap` (nothing written).
// Throws std::invalid_argument on a corner index >= n.  Synchronous on the
// null stream.  Allocates its scratch for the call and optionally reports the
// hierarchy build and the query times (HIP events).
bool lbvh_self_pairs(const double* x0, const double* x1, uint32_t n, const uint32_t* idx, uint32_t m, double pad,
                     uint32_t* pairs, uint64_t cap, uint64_t* count, double* build_ms, double* query_ms);
// The same without allocating: the scratch is the library's fixed device
// array (one copy per GPU), so m <= kLbvhMaxTris; calls that overlap on one
// GPU share it and race.  Must be called with `device` current.
bool lbvh_self_pairs_noalloc(int device, const double* x0, const double* x1, uint32_t n, const uint32_t* idx,
                             uint32_t m, double pad, uint32_t* pairs, uint64_t cap, uint64_t* count);

}  // namespace cloth
