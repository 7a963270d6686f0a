// sort.hip — stable LSD radix sort (8-bit digits) with LDS ranking.
//
// Per pass: (1) per-tile digit histogram in LDS, written digit-major;
// (2) one-block exclusive scan of the histogram; (3) each tile re-reads its
// keys and ranks them stably with wave64 ballots (match mask per digit) plus
// per-wave LDS digit counts, then scatters.
#include "sort.h"

namespace cloth {

constexpr int kSortBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kSortBlock * kItems;  // 4096 keys per tile
constexpr int kRadix = 256;
constexpr int kWaves = kSortBlock / kWave;

size_t radix_hist_entries(size_t n) { return (size_t)kRadix * ((n + kTile - 1) / kTile); }

__global__ void __launch_bounds__(kSortBlock) k_radix_hist(const uint64_t* __restrict__ keys, size_t n, int shift,
                                                           uint32_t* __restrict__ hist, unsigned nblocks) {
  __shared__ uint32_t h[kRadix];
  for (int i = threadIdx.x; i < kRadix; i += kSortBlock) h[i] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * kTile;
  for (int it = 0; it < kItems; it++) {
    const size_t i = base + (size_t)it * kSortBlock + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xff], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kRadix; d += kSortBlock) hist[(size_t)d * nblocks + blockIdx.x] = h[d];
}

// Exclusive scan of `cnt` u32 in place, one block of 1024 threads.
__global__ void __launch_bounds__(1024) k_scan_one_block(uint32_t* a, size_t cnt) {
  __shared__ uint32_t sums[1024];
  const size_t per = (cnt + 1023) / 1024;
  const size_t lo = threadIdx.x * per, hi = lo + per < cnt ? lo + per : cnt;
  uint32_t s = 0;
  for (size_t i = lo; i < hi; i++) s += a[i];
  sums[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = threadIdx.x >= (unsigned)off ? sums[threadIdx.x - off] : 0;
    __syncthreads();
    sums[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = threadIdx.x ? sums[threadIdx.x - 1] : 0;
  for (size_t i = lo; i < hi; i++) {
    const uint32_t v = a[i];
    a[i] = run;
    run += v;
  }
}

__global__ void __launch_bounds__(kSortBlock) k_radix_scatter(const uint64_t* __restrict__ kin,
                                                              const uint32_t* __restrict__ vin, size_t n, int shift,
                                                              const uint32_t* __restrict__ hist, unsigned nblocks,
                                                              uint64_t* __restrict__ kout, uint32_t* __restrict__ vout) {
  __shared__ uint32_t base_off[kRadix];
  __shared__ uint32_t running[kRadix];
  __shared__ uint32_t wave_cnt[kWaves][kRadix];
  for (int d = threadIdx.x; d < kRadix; d += kSortBlock) {
    base_off[d] = hist[(size_t)d * nblocks + blockIdx.x];
    running[d] = 0;
  }
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  const uint64_t lt_mask = lane ? ((~0ull) >> (64 - lane)) : 0ull;
  const size_t base = (size_t)blockIdx.x * kTile;
  for (int it = 0; it < kItems; it++) {
    for (int d = threadIdx.x; d < kWaves * kRadix; d += kSortBlock) (&wave_cnt[0][0])[d] = 0;
    __syncthreads();
    const size_t i = base + (size_t)it * kSortBlock + threadIdx.x;
    const bool valid = i < n;
    uint64_t key = 0;
    uint32_t val = 0, digit = 0;
    if (valid) {
      key = kin[i];
      val = vin[i];
      digit = (uint32_t)((key >> shift) & 0xff);
    }
    // lanes holding the same digit (valid lanes only)
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; bit++) {
      const uint64_t bset = __ballot(valid && ((digit >> bit) & 1));
      match &= ((digit >> bit) & 1) ? bset : ~bset;
    }
    const uint32_t rank_in_wave = __popcll(match & lt_mask);
    if (valid && rank_in_wave == 0) wave_cnt[wid][digit] = __popcll(match);
    __syncthreads();
    if (valid) {
      uint32_t off = running[digit] + rank_in_wave;
      for (int w = 0; w < wid; w++) off += wave_cnt[w][digit];
      const size_t dst = (size_t)base_off[digit] + off;
      kout[dst] = key;
      vout[dst] = val;
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kRadix; d += kSortBlock) {
      uint32_t t = 0;
      for (int w = 0; w < kWaves; w++) t += wave_cnt[w][d];
      running[d] += t;
    }
    __syncthreads();
  }
}

bool radix_sort_pairs(hipStream_t st, uint64_t* keys, uint32_t* vals, uint64_t* keys_tmp, uint32_t* vals_tmp,
                      size_t n, int begin_bit, int end_bit, SortScratch& scratch) {
  if (n <= 1 || end_bit <= begin_bit) return false;
  const unsigned nblocks = (unsigned)((n + kTile - 1) / kTile);
  const size_t entries = (size_t)kRadix * nblocks;
  if (scratch.hist_cap < entries) throw HipError("radix sort scratch too small");
  bool in_tmp = false;
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    const uint64_t* ki = in_tmp ? keys_tmp : keys;
    const uint32_t* vi = in_tmp ? vals_tmp : vals;
    uint64_t* ko = in_tmp ? keys : keys_tmp;
    uint32_t* vo = in_tmp ? vals : vals_tmp;
    hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(kSortBlock), 0, st, ki, n, shift, scratch.hist, nblocks);
    hipLaunchKernelGGL(k_scan_one_block, dim3(1), dim3(1024), 0, st, scratch.hist, entries);
    hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(kSortBlock), 0, st, ki, vi, n, shift, scratch.hist,
                       nblocks, ko, vo);
    CS_LAUNCH_CHECK();
    in_tmp = !in_tmp;
  }
  return in_tmp;
}

}  // namespace cloth
