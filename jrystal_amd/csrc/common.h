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

// Fixed-order block reduction of one double (blockDim.x a multiple of 64,
// at most 1024).  Deterministic: the tree shape depends only on blockDim.
__device__ inline double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

}  // namespace jrx
