// Shared helpers for the jrx HIP library (gfx950, FP64 throughout).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/jrx.h"

namespace jrx {

// Last error message of the calling host thread (returned by
// jrx_last_error()).
void set_error(const std::string& msg);

#define JRX_CHECK_HIP(expr)                                                   \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      ::jrx::set_error(std::string("HIP error in ") + #expr + ": " +          \
                       hipGetErrorString(_e));                                \
      return JRX_ERR_HIP;                                                     \
    }                                                                         \
  } while (0)

#define JRX_REQUIRE(cond, msg)                                                \
  do {                                                                        \
    if (!(cond)) {                                                            \
      ::jrx::set_error(msg);                                                  \
      return JRX_ERR_INVALID;                                                 \
    }                                                                         \
  } while (0)

#define JRX_LAUNCH_CHECK()                                                    \
  do {                                                                        \
    hipError_t _e = hipGetLastError();                                        \
    if (_e != hipSuccess) {                                                   \
      ::jrx::set_error(std::string("kernel launch failed: ") +                \
                       hipGetErrorString(_e));                                \
      return JRX_ERR_HIP;                                                     \
    }                                                                         \
  } while (0)

// ---------------------------------------------------------------- complex --
__host__ __device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__host__ __device__ inline double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__host__ __device__ inline double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__host__ __device__ inline double2 csub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
__host__ __device__ inline double2 cscale(double2 a, double s) {
  return make_double2(a.x * s, a.y * s);
}
__host__ __device__ inline double2 cconj(double2 a) {
  return make_double2(a.x, -a.y);
}
__host__ __device__ inline double2 czero() { return make_double2(0.0, 0.0); }

// Sum over the 64 lanes of a wavefront (xor butterfly: the same order on
// every run); every lane receives the sum.  No LDS and no block barrier.
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Sum of the values each thread of a block stored in sh[0, blockDim.x),
// after the barrier that follows those stores; wave 0 reduces (in a fixed
// order) and lane 0 of wave 0 gets the sum.  Needs only the caller's one
// barrier.
__device__ inline double wave0_sum(const double* sh) {
  double a = 0.0;
  for (int i = threadIdx.x; i < (int)blockDim.x; i += 64) a += sh[i];
  return wave_sum(a);
}

// Fixed-order block reduction of one double (blockDim.x a multiple of 64,
// at most 1024; sh holds blockDim.x / 64 doubles).  Deterministic: the
// order depends only on blockDim.  Every thread receives the sum.
__device__ inline double block_sum(double v, double* sh) {
  // wave partials by shuffles, then every thread sums them in wave order
  v = wave_sum(v);
  const int nw = ((int)blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  for (int w = 0; w < nw; ++w) r += sh[w];
  __syncthreads();  // sh may be reused right after
  return r;
}

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

}  // namespace jrx
