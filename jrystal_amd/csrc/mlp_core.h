// Small FP32 multilayer perceptron of the learned preconditioner:
//   x (kMlpIn features, normalised (x - mu) * inv_sd)
//   -> tanh(W1 x + b1) (64) -> tanh(W2 h1 + b2) (64) -> W3 h2 + b3 (2).
// Output 0 is the log of the per-coefficient preconditioner factor, output
// 1 the log amplitude of the warm-start coefficients.  Parameters live in
// one flat FP32 vector (offsets below); every kernel stages it in LDS.
#pragma once
#include <hip/hip_runtime.h>

namespace jrx {

constexpr int kMlpIn = 8;  // feature slots; unused slots are zero
constexpr int kMlpHidden = 64;
constexpr int kMlpOut = 2;
constexpr int kMlpW1 = 0;  // [hidden][in]
constexpr int kMlpB1 = kMlpW1 + kMlpHidden * kMlpIn;
constexpr int kMlpW2 = kMlpB1 + kMlpHidden;  // [hidden][hidden]
constexpr int kMlpB2 = kMlpW2 + kMlpHidden * kMlpHidden;
constexpr int kMlpW3 = kMlpB2 + kMlpHidden;  // [out][hidden]
constexpr int kMlpB3 = kMlpW3 + kMlpOut * kMlpHidden;
constexpr int kMlpParams = kMlpB3 + kMlpOut;  // 4866

// Feature normalisation: norm[0..7] = mu, norm[8..15] = 1 / sd.
constexpr int kMlpNorm = 2 * kMlpIn;

__device__ inline void mlp_stage(const float* __restrict__ params,
                                 float* __restrict__ lds) {
  for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) lds[i] = params[i];
}

__device__ inline void mlp_normalise(const float* __restrict__ norm,
                                     float x[kMlpIn]) {
#pragma unroll
  for (int i = 0; i < kMlpIn; ++i) x[i] = (x[i] - norm[i]) * norm[kMlpIn + i];
}

// Forward pass of one sample with the parameters in LDS (w).
__device__ inline void mlp_eval(const float* __restrict__ w,
                                const float x[kMlpIn], float y[kMlpOut]) {
  float h1[kMlpHidden];
#pragma unroll
  for (int j = 0; j < kMlpHidden; ++j) {
    float z = w[kMlpB1 + j];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) z += w[kMlpW1 + j * kMlpIn + i] * x[i];
    h1[j] = tanhf(z);
  }
  float y0 = w[kMlpB3], y1 = w[kMlpB3 + 1];
  for (int j = 0; j < kMlpHidden; ++j) {
    float z = w[kMlpB2 + j];
#pragma unroll
    for (int i = 0; i < kMlpHidden; ++i)
      z += w[kMlpW2 + j * kMlpHidden + i] * h1[i];
    const float h = tanhf(z);
    y0 += w[kMlpW3 + j] * h;
    y1 += w[kMlpW3 + kMlpHidden + j] * h;
  }
  y[0] = y0;
  y[1] = y1;
}

// Raw features of one coefficient: |G+k|^2, the band's tb (twice its
// kinetic energy), its energy gap below the highest occupied band of its
// (s,k) block (Rayleigh quotients; 0 for empty bands) and the density mean /
// variance of the current iterate.
__device__ inline void mlp_features(double gk2, double tb, double gap,
                                    double rho_mean, double rho_var,
                                    float x[kMlpIn]) {
  const double t = tb > 1e-300 ? tb : 1e-300;
  const double m = rho_mean > 1e-300 ? rho_mean : 1e-300;
  x[0] = (float)log1p(gk2);
  x[1] = (float)log1p(gk2 / t);
  x[2] = (float)log(t);
  x[3] = (float)log(m);
  x[4] = (float)log1p((rho_var > 0.0 ? rho_var : 0.0) / (m * m));
  x[5] = (float)log1p(gap > 0.0 ? gap : 0.0);
  x[6] = 0.f;
  x[7] = 0.f;
}

// Clamp ranges of the two outputs before exponentiation.
constexpr float kLogFactorMin = -8.f, kLogFactorMax = 8.f;
constexpr float kLogAmpMin = -30.f, kLogAmpMax = 8.f;

}  // namespace jrx
