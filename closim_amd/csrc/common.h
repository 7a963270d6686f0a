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
inline unsigned red_grid(size_t n, int block = kBlock) {
  size_t g = (n + block - 1) / block;
  if (g > 1024) g = 1024;
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

// Grid reduction with the "last block finishes" ticket: every block writes its
// partial, the block that takes the last ticket merges all partials in block
// order and returns true (in thread 0) with the total in `out`.
// `ticket` must be zero on entry; it is reset by the last block.
template <int BLOCK>
__device__ __forceinline__ bool grid_reduce_dd(jsm::dd v, jsm::dd* partials, unsigned* ticket, jsm::dd& out) {
  jsm::dd b = block_reduce_dd<BLOCK>(v);
  __shared__ bool s_last;
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = b;
    __threadfence();
    const unsigned t = atomicAdd(ticket, 1u);
    s_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
  jsm::dd acc = jsm::dd_zero();
  for (unsigned i = threadIdx.x; i < gridDim.x; i += BLOCK) {
    acc = jsm::dd_merge(acc, partials[i]);
  }
  acc = block_reduce_dd<BLOCK>(acc);
  if (threadIdx.x == 0) {
    *ticket = 0;
    out = acc;
  }
  return threadIdx.x == 0;
}

// Max / min reductions (exact, order independent) with the same ticket
// pattern.  Op picks the JS semantics: JsMax propagates NaN like Math.max;
// GtMax / LtMin ignore NaN like `if (v > best) best = v` loops.
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

template <int BLOCK, typename Op>
__device__ __forceinline__ bool grid_reduce_ext(double v, double* partials, unsigned* ticket, double& out) {
  __shared__ double s[BLOCK / kWave];
  __shared__ bool s_last;
  for (int off = kWave / 2; off > 0; off >>= 1) v = Op::op(v, __shfl_down(v, off, kWave));
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  if (lane == 0) s[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = s[0];
    for (int w = 1; w < BLOCK / kWave; w++) b = Op::op(b, s[w]);
    partials[blockIdx.x] = b;
    __threadfence();
    const unsigned t = atomicAdd(ticket, 1u);
    s_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
  if (threadIdx.x == 0) {
    double m = Op::ident();
    for (unsigned i = 0; i < gridDim.x; i++) m = Op::op(m, partials[i]);
    *ticket = 0;
    out = m;
    return true;
  }
  return false;
}

}  // namespace cloth
