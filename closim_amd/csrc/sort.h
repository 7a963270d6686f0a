// sort.h — LDS-ranked LSD radix sort of (u64 key, u32 value) pairs and a
// device exclusive scan, used for Morton codes, the active contact list and
// the per-vertex contact incidence.
#pragma once

#include "common.h"

namespace cloth {

struct SortScratch {
  uint32_t* hist = nullptr;  // 256 * max_blocks
  size_t hist_cap = 0;
};

// Stable sort of n pairs on key bits [begin_bit, end_bit).  Sorts in place
// between (keys, vals) and (keys_tmp, vals_tmp); returns true when the result
// ended up in the *_tmp buffers.
bool radix_sort_pairs(hipStream_t st, uint64_t* keys, uint32_t* vals, uint64_t* keys_tmp, uint32_t* vals_tmp,
                      size_t n, int begin_bit, int end_bit, SortScratch& scratch);

size_t radix_hist_entries(size_t n);

}  // namespace cloth
