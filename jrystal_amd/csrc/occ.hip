// The simplex projector of parameterised occupations on the device.
//
// occupation.py:295-337 (proj) of the reference maps x (n entries in [0, 1])
// to the point of {0 <= y <= 1, sum y = total} reached by one uniform
// shift and clipping: when sum x > total the entries are sorted ascending,
// cs = cumulative sums, rho_k = (s - total - cs_k) / (n - k - 1), k = the
// first k with xs_k > rho_k (0 if none), lam = (s - total - cs_{k-1}) /
// (n - k) (k = 0: (s - total) / n) and y = max(x - lam, 0); otherwise the
// same with the entries sorted descending, rho_k = (s - total + (k + 1 -
// cs_k)) / (n - k - 1), the first k with xs_k - rho_k < 1, lam = (total -
// s - (k - cs_{k-1})) / (n - k) and y = min(x + lam, 1).  Ties sort by
// index (a stable sort).  The divisions follow IEEE semantics, so the rule
// behaves as the reference's at n - k - 1 = 0.
// The vector-Jacobian product keeps the active set fixed: lam depends on
// the entries not among the k left out, d lam / d x_j = [j kept] / (n - k)
// (up to sign), and clipped entries have zero derivative:
//   xbar = ybar * live - (sum ybar * live) * [kept] / (n - k).
#include "common.h"

#include <cmath>

namespace jrx {

namespace {

constexpr int kProjBlock = 1024;
constexpr int kProjMax = 8192;  // entries per channel

__device__ inline bool proj_less(double ka, int ia, double kb, int ib,
                                 bool desc) {
  if (ka == kb) return ia < ib;
  return desc ? ka > kb : ka < kb;
}

// One block per channel c: in/out at c * n.
__global__ __launch_bounds__(kProjBlock) void simplex_proj_kernel(
    const double* __restrict__ in, int n, int N2, const double* __restrict__ totals,
    int sigmoid, double* __restrict__ y, const double* __restrict__ ybar,
    double* __restrict__ in_bar) {
  extern __shared__ double sh[];
  double* key = sh;                                      // [N2]
  int* idx = reinterpret_cast<int*>(key + N2);           // [N2]
  double* psum = reinterpret_cast<double*>(idx + N2);    // [kProjBlock]
  unsigned char* out_k = reinterpret_cast<unsigned char*>(psum + kProjBlock);
  __shared__ double red[kProjBlock / 64];
  __shared__ int kmin;
  __shared__ double cs_prev;
  const int tid = threadIdx.x, c = blockIdx.x;
  const double total = totals[c];
  in += (int64_t)c * n;
  y += (int64_t)c * n;
  double s = 0.0;
  for (int i = tid; i < n; i += kProjBlock) {
    const double t = in[i];
    const double x = sigmoid ? 1.0 / (1.0 + exp(-t)) : t;
    key[i] = x;
    idx[i] = i;
    s += x;
  }
  s = block_sum(s, red);
  const bool up = s <= total;  // uniform
  const double pad = up ? -INFINITY : INFINITY;
  for (int i = n + tid; i < N2; i += kProjBlock) {
    key[i] = pad;
    idx[i] = i;
  }
  __syncthreads();
  // bitonic sort of (key, index): ascending (down) or descending (up)
  for (int size = 2; size <= N2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = tid; t < N2 / 2; t += kProjBlock) {
        const int i = 2 * stride * (t / stride) + (t % stride);
        const int j = i + stride;
        const bool asc = (i & size) == 0;
        const double ki = key[i], kj = key[j];
        const int ii = idx[i], ij = idx[j];
        const bool swap = asc ? proj_less(kj, ij, ki, ii, up)
                              : proj_less(ki, ii, kj, ij, up);
        if (swap) {
          key[i] = kj; key[j] = ki;
          idx[i] = ij; idx[j] = ii;
        }
      }
      __syncthreads();
    }
  }
  // cumulative sums: thread t owns sorted positions [t E, t E + E)
  const int E = (n + kProjBlock - 1) / kProjBlock;
  const int p0 = tid * E;
  double loc = 0.0;
  for (int q = 0; q < E; ++q)
    if (p0 + q < n) loc += key[p0 + q];
  psum[tid] = loc;
  __syncthreads();
  for (int d = 1; d < kProjBlock; d <<= 1) {  // inclusive scan, fixed order
    const double v = tid >= d ? psum[tid - d] : 0.0;
    __syncthreads();
    psum[tid] += v;
    __syncthreads();
  }
  if (tid == 0) kmin = n;
  __syncthreads();
  double run = psum[tid] - loc;  // exclusive prefix of this thread's range
  int kfirst = n;
  for (int q = 0; q < E; ++q) {
    const int k = p0 + q;
    if (k >= n) break;
    run += key[k];  // cs_k
    const double den = (double)(n - k - 1);
    bool cond;
    if (!up) {
      const double rho = (s - total - run) / den;
      cond = key[k] > rho;
    } else {
      const double rho = (s - total + ((double)(k + 1) - run)) / den;
      cond = key[k] - rho < 1.0;
    }
    if (cond && kfirst == n) kfirst = k;
  }
  if (kfirst < n) atomicMin(&kmin, kfirst);
  __syncthreads();
  const int k = kmin < n ? kmin : 0;
  // cs_{k-1}, from the thread that owns sorted position k - 1
  if (k > 0 && tid == (k - 1) / E) {
    double r2 = psum[tid] - loc;
    for (int q = p0; q <= k - 1; ++q) r2 += key[q];
    cs_prev = r2;
  }
  __syncthreads();
  double lam;
  if (!up) lam = k == 0 ? (s - total) / n : (s - total - cs_prev) / (n - k);
  else lam = k == 0 ? (total - s) / n : (total - s - ((double)k - cs_prev)) / (n - k);
  if (ybar) {
    for (int q = tid; q < n; q += kProjBlock) out_k[idx[q]] = q < k ? 1 : 0;
  }
  // y, and the live entries' part of the product
  double g = 0.0;
  for (int i = tid; i < n; i += kProjBlock) {
    const double t = in[i];
    const double x = sigmoid ? 1.0 / (1.0 + exp(-t)) : t;
    double v;
    bool live;
    if (!up) {
      v = x - lam;
      v = v < 0.0 ? 0.0 : v;
      live = v > 0.0;
    } else {
      v = x + lam;
      v = v > 1.0 ? 1.0 : v;
      live = v < 1.0;
    }
    y[i] = v;
    if (ybar && live) g += ybar[(int64_t)c * n + i];
  }
  if (!ybar) return;  // uniform
  g = block_sum(g, red);
  __syncthreads();  // out_k complete
  const double inv = 1.0 / (double)(n - k);
  for (int i = tid; i < n; i += kProjBlock) {
    const double t = in[i];
    const double x = sigmoid ? 1.0 / (1.0 + exp(-t)) : t;
    const double v = y[i];
    const bool live = !up ? v > 0.0 : v < 1.0;
    double xb = (live ? ybar[(int64_t)c * n + i] : 0.0) -
                (out_k[i] ? 0.0 : g * inv);
    if (sigmoid) xb *= x * (1.0 - x);
    in_bar[(int64_t)c * n + i] = xb;
  }
}

}  // namespace

}  // namespace jrx

using namespace jrx;

extern "C" int jrx_simplex_projection(const double* in, int64_t n,
                                      int channels, const double* totals,
                                      int sigmoid, double* y,
                                      const double* ybar, double* in_bar,
                                      void* stream) {
  JRX_REQUIRE(in && totals && y && n >= 1 && channels >= 1,
              "invalid simplex projection arguments");
  JRX_REQUIRE(!ybar || in_bar, "ybar requires in_bar");
  if (n > kProjMax) {
    set_error("simplex projection of more than 8192 entries per channel is "
              "not supported by this build");
    return JRX_ERR_UNSUPPORTED;
  }
  int N2 = 2;
  while (N2 < n) N2 <<= 1;
  const size_t shm = (size_t)N2 * (sizeof(double) + sizeof(int)) +
                     kProjBlock * sizeof(double) + (size_t)N2;
  hipLaunchKernelGGL(simplex_proj_kernel, dim3(channels), dim3(kProjBlock),
                     shm, (hipStream_t)stream, in, (int)n, N2, totals,
                     sigmoid, y, ybar, in_bar);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}
