// common.h — shared HIP utilities for the cloth step library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

#include "jsmath.h"

namespace cloth {

struct HipError : std::runtime_error {
  explicit HipError(const std::string& s) : std::runtime_error(s) {}
};

#define CS_HIP(call)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      throw ::cloth::HipError(std::string(#call) + ": " + hipGetErrorString(e_) + " at " + \
                              __FILE__ + ":" + std::to_string(__LINE__));                  \
  } while (0)

#define CS_LAUNCH_CHECK() CS_HIP(hipGetLastError())

// Device buffer that grows on demand (contents are not preserved on growth).
template <typename T>
struct DevBuf {
  T* ptr = nullptr;
  size_t count = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    count = 0;
  }
  void alloc(size_t n) {
    if (n <= count && ptr) return;
    release();
    if (n == 0) n = 1;
    CS_HIP(hipMalloc(&ptr, n * sizeof(T)));
    count = n;
  }
  void upload(const T* h, size_t n, hipStream_t st) {
    alloc(n);
    if (n) CS_HIP(hipMemcpyAsync(ptr, h, n * sizeof(T), hipMemcpyHostToDevice, st));
  }
  void zero(hipStream_t st) { CS_HIP(hipMemsetAsync(ptr, 0, count * sizeof(T), st)); }
};

constexpr int kBlock = 256;       // threads per block for 1-D kernels (4 wave64)
constexpr int kWave = 64;

inline unsigned grid_for(size_t n, int block = kBlock) {
  size_t g = (n + block - 1) / block;
  return (unsigned)(g ? g : 1);
}

// Bounded grid for reduction kernels: a fixed count per size keeps the
// partial layout (and thus the reduction tree) deterministic.
constexpr unsigned kMaxRedBlocks = 1024;  // partial slots per reduction
constexpr unsigned kMaxHvpBlocks = 4096;  // HVP kernels: more blocks to hide gather latency
constexpr unsigned kXcds = 8;             // MI355X: 8 XCDs, work-groups dispatched round-robin
static_assert(kMaxHvpBlocks % kXcds == 0, "HVP grids are whole multiples of the XCD count");
inline unsigned red_grid(size_t n, int block = kBlock) {
  size_t g = (n + block - 1) / block;
  if (g > kMaxRedBlocks) g = kMaxRedBlocks;
  return (unsigned)(g ? g : 1);
}

// ---- double-double reductions (order-independent, correctly rounded) -------

__device__ __forceinline__ jsm::dd shfl_dd(jsm::dd v, int off) {
  jsm::dd r;
  r.hi = __shfl_down(v.hi, off, kWave);
  r.lo = __shfl_down(v.lo, off, kWave);
  return r;
}

// Reduce one dd per thread to lane 0 of thread 0 (block-wide); fixed tree.
template <int BLOCK>
__device__ __forceinline__ jsm::dd block_reduce_dd(jsm::dd v) {
  __shared__ double s_hi[BLOCK / kWave], s_lo[BLOCK / kWave];
  for (int off = kWave / 2; off > 0; off >>= 1) v = jsm::dd_merge(v, shfl_dd(v, off));
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  if (lane == 0) { s_hi[wid] = v.hi; s_lo[wid] = v.lo; }
  __syncthreads();
  jsm::dd r = jsm::dd_zero();
  if (threadIdx.x == 0) {
    for (int w = 0; w < BLOCK / kWave; w++) r = jsm::dd_merge(r, jsm::dd{s_hi[w], s_lo[w]});
  }
  __syncthreads();
  return r;
}

// Grid reductions are two-phase and fence-free: each block writes its
// partial to partials[blockIdx.x]; a later kernel (the consumer, or a
// one-block finish kernel when the host reads the value) merges the
// partials in a fixed order.  On gfx950 a per-block release fence costs an
// L2 write-back each, which serialises the blocks; the kernel boundary
// provides the same visibility once per launch.
template <int BLOCK>
__device__ __forceinline__ void block_partial_dd(jsm::dd v, jsm::dd* partials) {
  const jsm::dd b = block_reduce_dd<BLOCK>(v);
  if (threadIdx.x == 0) partials[blockIdx.x] = b;
}

// Merge partials[0..count) (fixed tree); every thread of the block gets the total.
template <int BLOCK>
__device__ __forceinline__ jsm::dd merge_partials_dd(const jsm::dd* partials, unsigned count) {
  __shared__ double s_tot[2];
  jsm::dd acc = jsm::dd_zero();
  for (unsigned i = threadIdx.x; i < count; i += BLOCK) acc = jsm::dd_merge(acc, partials[i]);
  acc = block_reduce_dd<BLOCK>(acc);
  if (threadIdx.x == 0) { s_tot[0] = acc.hi; s_tot[1] = acc.lo; }
  __syncthreads();
  const jsm::dd r{s_tot[0], s_tot[1]};
  __syncthreads();
  return r;
}

// Max / min reductions (exact, order independent).  Op picks the JS
// semantics: JsMax propagates NaN like Math.max; GtMax / LtMin ignore NaN
// like `if (v > best) best = v` loops (their inputs are never NaN here).
struct JsMax {
  __device__ static double ident() { return -__builtin_inf(); }
  __device__ static double op(double a, double b) { return jsm::jmax(a, b); }
};
struct GtMax {
  __device__ static double ident() { return -__builtin_inf(); }
  __device__ static double op(double a, double b) { return b > a ? b : a; }
};
struct LtMin {
  __device__ static double ident() { return __builtin_inf(); }
  __device__ static double op(double a, double b) { return b < a ? b : a; }
};

// Block-wide Op reduction; every thread gets the result.
template <int BLOCK, typename Op>
__device__ __forceinline__ double block_reduce_ext(double v) {
  __shared__ double s[BLOCK / kWave];
  __shared__ double s_res;
  for (int off = kWave / 2; off > 0; off >>= 1) v = Op::op(v, __shfl_down(v, off, kWave));
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  if (lane == 0) s[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = s[0];
    for (int w = 1; w < BLOCK / kWave; w++) b = Op::op(b, s[w]);
    s_res = b;
  }
  __syncthreads();
  const double r = s_res;
  __syncthreads();
  return r;
}

template <int BLOCK, typename Op>
__device__ __forceinline__ void block_partial_ext(double v, double* partials) {
  const double b = block_reduce_ext<BLOCK, Op>(v);
  if (threadIdx.x == 0) partials[blockIdx.x] = b;
}

template <int BLOCK, typename Op>
__device__ __forceinline__ double merge_partials_ext(const double* partials, unsigned count) {
  double acc = Op::ident();
  for (unsigned i = threadIdx.x; i < count; i += BLOCK) acc = Op::op(acc, partials[i]);
  return block_reduce_ext<BLOCK, Op>(acc);
}

// acc + v[b*S] + v[(b+1)*S] + ... + v[(e-1)*S], added in that order (two
// values per entry, v[k*S] then v[k*S + D], when D > 0).  Loads are issued
// 32 entries ahead of the dependent additions.
template <int S, int D>
__device__ __forceinline__ double ordered_sum(const double* __restrict__ v, uint32_t b, uint32_t e, double acc) {
  constexpr int U = 32;
  uint32_t k = b;
  for (; k + U <= e; k += U) {
    double t0[U], t1[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      t0[j] = v[(size_t)(k + j) * S];
      if (D > 0) t1[j] = v[(size_t)(k + j) * S + D];
    }
#pragma unroll
    for (int j = 0; j < U; j++) {
      acc += t0[j];
      if (D > 0) acc += t1[j];
    }
  }
  for (; k < e; k++) {
    acc += v[(size_t)k * S];
    if (D > 0) acc += v[(size_t)k * S + D];
  }
  return acc;
}

// A latch read once per block: thread 0 loads it and the block shares the
// value, so all threads of a block take the same early exit even when
// block 0 of the same launch sets the latch meanwhile.
__device__ __forceinline__ bool block_latched(const int32_t* flag) {
  __shared__ int32_t s_flag;
  if (threadIdx.x == 0) s_flag = __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  return s_flag != 0;
}

// Host-read reductions: one block per job merges its partials.
struct FinishJob {
  const void* partials;
  unsigned count;
  int kind;  // FIN_DD, FIN_JSMAX, FIN_GTMAX, FIN_LTMIN
  int clamp0;  // result = Op(0, result) (max(0, .) for the max kinds)
  double* out;
};
enum { FIN_DD = 0, FIN_JSMAX = 1, FIN_GTMAX = 2, FIN_LTMIN = 3 };
struct FinishJobs {
  FinishJob j[4];
};
void launch_finish(hipStream_t st, const FinishJobs& jobs, int njobs);

}  // namespace cloth
