// Training and inference kernels of the learned-preconditioner network
// (see mlp_core.h for the architecture).  All reductions run in a fixed
// order, so a training run is bitwise reproducible.
#include "common.h"
#include "mlp_core.h"

namespace jrx {
namespace {

constexpr int kFwdBlock = 256;
constexpr int kBwdTile = 128;      // samples per backward tile (one per thread)
constexpr int kBwdStride = kBwdTile + 1;  // padded LDS rows
constexpr int kBwdMaxBlocks = 1024;
constexpr int kLossBlock = 256;

__global__ void __launch_bounds__(kFwdBlock)
mlp_forward_kernel(const float* __restrict__ params,
                   const float* __restrict__ norm,
                   const float* __restrict__ x, int64_t n,
                   float* __restrict__ y) {
  __shared__ float w[kMlpParams];
  mlp_stage(params, w);
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * kFwdBlock + threadIdx.x;
  if (i >= n) return;
  float xi[kMlpIn];
#pragma unroll
  for (int k = 0; k < kMlpIn; ++k) xi[k] = x[i * kMlpIn + k];
  mlp_normalise(norm, xi);
  float yi[kMlpOut];
  mlp_eval(w, xi, yi);
  y[i * kMlpOut] = yi[0];
  y[i * kMlpOut + 1] = yi[1];
}

// Parameter gradient of sum_i dy_i . y(x_i).  Block b takes tiles b,
// b + gridDim.x, ... of 128 samples; thread t owns sample t of a tile in
// the forward/backward sweep, then parameters t, t + 128, ... in the
// reduction over the tile's samples, added to the block's own row
// gpart[b][:] (no other block touches it).
__global__ void __launch_bounds__(kBwdTile)
mlp_backward_kernel(const float* __restrict__ params,
                    const float* __restrict__ norm,
                    const float* __restrict__ x,
                    const float* __restrict__ dy, int64_t n,
                    float* __restrict__ gpart) {
  __shared__ float w[kMlpParams];
  __shared__ float H1[kMlpHidden * kBwdStride];
  __shared__ float H2[kMlpHidden * kBwdStride];
  __shared__ float D1[kMlpHidden * kBwdStride];
  __shared__ float D2[kMlpHidden * kBwdStride];
  __shared__ float X[kMlpIn * kBwdStride];
  __shared__ float DY[kMlpOut * kBwdStride];
  mlp_stage(params, w);
  const int t = threadIdx.x;
  float* grow = gpart + (int64_t)blockIdx.x * kMlpParams;
  for (int p = t; p < kMlpParams; p += kBwdTile) grow[p] = 0.f;
  const int64_t ntiles = (n + kBwdTile - 1) / kBwdTile;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    __syncthreads();  // LDS of the previous tile fully consumed
    const int64_t i = tile * kBwdTile + t;
    const bool valid = i < n;
    float xi[kMlpIn];
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) xi[k] = valid ? x[i * kMlpIn + k] : 0.f;
    mlp_normalise(norm, xi);
    const float dy0 = valid ? dy[i * kMlpOut] : 0.f;
    const float dy1 = valid ? dy[i * kMlpOut + 1] : 0.f;
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) X[k * kBwdStride + t] = xi[k];
    DY[t] = dy0;
    DY[kBwdStride + t] = dy1;
    // forward, keeping h1 and h2 in LDS; dZ2 on the fly
    for (int j = 0; j < kMlpHidden; ++j) {
      float z = w[kMlpB1 + j];
#pragma unroll
      for (int k = 0; k < kMlpIn; ++k) z += w[kMlpW1 + j * kMlpIn + k] * xi[k];
      H1[j * kBwdStride + t] = tanhf(z);
    }
    for (int j = 0; j < kMlpHidden; ++j) {
      float z = w[kMlpB2 + j];
#pragma unroll 16
      for (int k = 0; k < kMlpHidden; ++k)
        z += w[kMlpW2 + j * kMlpHidden + k] * H1[k * kBwdStride + t];
      const float h = tanhf(z);
      H2[j * kBwdStride + t] = h;
      const float dh = w[kMlpW3 + j] * dy0 + w[kMlpW3 + kMlpHidden + j] * dy1;
      D2[j * kBwdStride + t] = dh * (1.f - h * h);
    }
    // dZ1 = (W2^T dZ2) * (1 - h1^2), own sample only
    for (int k = 0; k < kMlpHidden; ++k) {
      float s = 0.f;
#pragma unroll 16
      for (int j = 0; j < kMlpHidden; ++j)
        s += w[kMlpW2 + j * kMlpHidden + k] * D2[j * kBwdStride + t];
      const float h = H1[k * kBwdStride + t];
      D1[k * kBwdStride + t] = s * (1.f - h * h);
    }
    __syncthreads();
    // reduction over the tile's samples, fixed order
    for (int p = t; p < kMlpParams; p += kBwdTile) {
      const float* a;
      const float* b = nullptr;
      if (p < kMlpB1) {
        a = D1 + (p / kMlpIn) * kBwdStride;
        b = X + (p % kMlpIn) * kBwdStride;
      } else if (p < kMlpW2) {
        a = D1 + (p - kMlpB1) * kBwdStride;
      } else if (p < kMlpB2) {
        const int q = p - kMlpW2;
        a = D2 + (q / kMlpHidden) * kBwdStride;
        b = H1 + (q % kMlpHidden) * kBwdStride;
      } else if (p < kMlpW3) {
        a = D2 + (p - kMlpB2) * kBwdStride;
      } else if (p < kMlpB3) {
        const int q = p - kMlpW3;
        a = DY + (q / kMlpHidden) * kBwdStride;
        b = H2 + (q % kMlpHidden) * kBwdStride;
      } else {
        a = DY + (p - kMlpB3) * kBwdStride;
      }
      float s = 0.f;
      if (b) {
        for (int k = 0; k < kBwdTile; ++k) s += a[k] * b[k];
      } else {
        for (int k = 0; k < kBwdTile; ++k) s += a[k];
      }
      grow[p] += s;
    }
  }
}

// grad[p] (+)= sum_b gpart[b][p], blocks in order.
__global__ void mlp_grad_reduce_kernel(const float* __restrict__ gpart,
                                       int nblk, int accumulate,
                                       float* __restrict__ grad) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= kMlpParams) return;
  float s = 0.f;
  for (int b = 0; b < nblk; ++b) s += gpart[(int64_t)b * kMlpParams + p];
  grad[p] = accumulate ? grad[p] + s : s;
}

__device__ inline float block_sum_f(float v, float* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  const float r = sh[0];
  __syncthreads();
  return r;
}

// One block per snapshot of m samples.  With p_i = exp(clamp(y_i0)):
//   c = sum p a,  U = sum p^2 b,  L = 1 - c^2 / (T U)
// (one minus the squared cosine between the preconditioned step and the
// direction to the minimum); dy_i0 = w dL/dy_i0, dy_i1 = 0, w = 1/nsnap.
__global__ void __launch_bounds__(kLossBlock)
cosine_loss_kernel(const float* __restrict__ y, const float* __restrict__ a,
                   const float* __restrict__ b, const float* __restrict__ T,
                   int64_t m, float wgt, float* __restrict__ dy,
                   float* __restrict__ loss) {
  __shared__ float red[kLossBlock];
  const int64_t s = blockIdx.x;
  const int64_t base = s * m;
  float c = 0.f, U = 0.f;
  for (int64_t i = threadIdx.x; i < m; i += kLossBlock) {
    const float yi = y[(base + i) * kMlpOut];
    const float p = expf(fminf(fmaxf(yi, kLogFactorMin), kLogFactorMax));
    c += p * a[base + i];
    U += p * p * b[base + i];
  }
  c = block_sum_f(c, red);
  U = block_sum_f(U, red);
  const float Ts = T[s];
  const float TU = Ts * U > 1e-30f ? Ts * U : 1e-30f;
  const float Uc = U > 1e-30f ? U : 1e-30f;
  if (threadIdx.x == 0) loss[s] = 1.f - c * c / TU;
  for (int64_t i = threadIdx.x; i < m; i += kLossBlock) {
    const float yi = y[(base + i) * kMlpOut];
    const bool inside = yi > kLogFactorMin && yi < kLogFactorMax;
    const float p = expf(fminf(fmaxf(yi, kLogFactorMin), kLogFactorMax));
    const float dLdp = -2.f * c / TU * (a[base + i] - c / Uc * p * b[base + i]);
    dy[(base + i) * kMlpOut] = inside ? wgt * p * dLdp : 0.f;
    dy[(base + i) * kMlpOut + 1] = 0.f;
  }
}

// Mean squared error of output 1: L = (1/n) sum (y_i1 - t_i)^2, per-block
// partial losses; dy_i1 = 2 (y_i1 - t_i) / n, dy_i0 = 0.
__global__ void __launch_bounds__(kLossBlock)
mse_loss_kernel(const float* __restrict__ y, const float* __restrict__ tgt,
                int64_t n, float* __restrict__ dy, float* __restrict__ part) {
  __shared__ float red[kLossBlock];
  const int64_t i = (int64_t)blockIdx.x * kLossBlock + threadIdx.x;
  float e = 0.f;
  if (i < n) {
    const float r = y[i * kMlpOut + 1] - tgt[i];
    e = r * r;
    dy[i * kMlpOut] = 0.f;
    dy[i * kMlpOut + 1] = 2.f * r / (float)n;
  }
  e = block_sum_f(e, red);
  if (threadIdx.x == 0) part[blockIdx.x] = e / (float)n;
}

__global__ void sum_kernel_f(const float* __restrict__ v, int64_t n,
                             float* __restrict__ out) {
  __shared__ float red[kLossBlock];
  float a = 0.f;
  for (int64_t i = threadIdx.x; i < n; i += kLossBlock) a += v[i];
  a = block_sum_f(a, red);
  if (threadIdx.x == 0) out[0] = a;
}

// optax.adam rule in FP32 (eps_root = 0); t is the step count after this
// update (1 for the first step).
__global__ void adam_f32_kernel(float* __restrict__ p,
                                const float* __restrict__ g,
                                float* __restrict__ m, float* __restrict__ v,
                                int64_t n, float lr, float b1, float b2,
                                float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace
}  // namespace jrx

using namespace jrx;

extern "C" {

size_t jrx_mlp_num_params(void) { return kMlpParams; }

int jrx_mlp_forward(const float* params, const float* norm, const float* x,
                    int64_t n, float* y, void* stream) {
  JRX_REQUIRE(params && norm && x && y && n >= 0, "invalid mlp arguments");
  if (n == 0) return JRX_OK;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(ceil_div(n, kFwdBlock)),
                     dim3(kFwdBlock), 0, (hipStream_t)stream, params, norm, x,
                     n, y);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

size_t jrx_mlp_backward_workspace_bytes(int64_t n) {
  int64_t nblk = (n + kBwdTile - 1) / kBwdTile;
  if (nblk < 1) nblk = 1;
  if (nblk > kBwdMaxBlocks) nblk = kBwdMaxBlocks;
  return (size_t)nblk * kMlpParams * sizeof(float);
}

int jrx_mlp_backward(const float* params, const float* norm, const float* x,
                     const float* dy, int64_t n, float* grad, int accumulate,
                     void* workspace, size_t workspace_bytes, void* stream) {
  JRX_REQUIRE(params && norm && x && dy && grad && workspace && n >= 0,
              "invalid mlp arguments");
  JRX_REQUIRE(workspace_bytes >= jrx_mlp_backward_workspace_bytes(n),
              "mlp workspace too small");
  int64_t nblk = (n + kBwdTile - 1) / kBwdTile;
  if (nblk < 1) nblk = 1;
  if (nblk > kBwdMaxBlocks) nblk = kBwdMaxBlocks;
  hipStream_t s = (hipStream_t)stream;
  float* gpart = static_cast<float*>(workspace);
  hipLaunchKernelGGL(mlp_backward_kernel, dim3((unsigned)nblk),
                     dim3(kBwdTile), 0, s, params, norm, x, dy, n, gpart);
  JRX_LAUNCH_CHECK();
  hipLaunchKernelGGL(mlp_grad_reduce_kernel, dim3(ceil_div(kMlpParams, 256)),
                     dim3(256), 0, s, gpart, (int)nblk, accumulate, grad);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int jrx_mlp_cosine_loss(const float* y, const float* a, const float* b,
                        const float* t2, int64_t snap_size, int64_t nsnap,
                        float* dy, float* loss, void* stream) {
  JRX_REQUIRE(y && a && b && t2 && dy && loss && snap_size > 0 && nsnap > 0,
              "invalid loss arguments");
  JRX_REQUIRE(nsnap <= 2147483647, "too many snapshots");
  hipLaunchKernelGGL(cosine_loss_kernel, dim3((unsigned)nsnap),
                     dim3(kLossBlock), 0, (hipStream_t)stream, y, a, b, t2,
                     snap_size, 1.f / (float)nsnap, dy, loss);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

size_t jrx_mlp_mse_workspace_bytes(int64_t n) {
  return ((size_t)ceil_div(n > 0 ? n : 1, kLossBlock) + 1) * sizeof(float);
}

int jrx_mlp_mse_loss(const float* y, const float* target, int64_t n,
                     float* dy, float* loss, void* workspace,
                     size_t workspace_bytes, void* stream) {
  JRX_REQUIRE(y && target && dy && loss && workspace && n > 0,
              "invalid loss arguments");
  JRX_REQUIRE(workspace_bytes >= jrx_mlp_mse_workspace_bytes(n),
              "loss workspace too small");
  const int nblk = ceil_div(n, kLossBlock);
  hipStream_t s = (hipStream_t)stream;
  float* part = static_cast<float*>(workspace);
  hipLaunchKernelGGL(mse_loss_kernel, dim3(nblk), dim3(kLossBlock), 0, s, y,
                     target, n, dy, part);
  JRX_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum_kernel_f, dim3(1), dim3(kLossBlock), 0, s, part,
                     (int64_t)nblk, loss);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int jrx_mlp_adam(float* params, const float* grad, float* m, float* v,
                 int64_t n, double lr, double b1, double b2, double eps,
                 int64_t count, void* stream) {
  JRX_REQUIRE(params && grad && m && v && n >= 0 && count >= 1,
              "invalid adam arguments");
  if (n == 0) return JRX_OK;
  const double c1 = 1.0 - pow(b1, (double)count);
  const double c2 = 1.0 - pow(b2, (double)count);
  hipLaunchKernelGGL(adam_f32_kernel, dim3(ceil_div(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, params, grad, m, v, n, (float)lr,
                     (float)b1, (float)b2, (float)eps, (float)c1, (float)c2);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

}  // extern "C"
