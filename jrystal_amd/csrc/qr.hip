// Batched CholeskyQR2 and the QR adjoint (FP64).
//
// The reference orthonormalises W = w_re + i w_im with a Householder QR
// (jrystal/_src/unitary_module.py:74-75).  The total energy is invariant
// under Q -> Q D (D diagonal unitary) and W -> W T (T invertible upper
// triangular), so any QR of W gives the same energy and gradient.
// CholeskyQR2 (Fukaya et al. 2014) maps the work onto Gram and small
// triangular products:
//   S1 = W^H W = R1^H R1,  S2 = (W R1^-1)^H (W R1^-1) = R2^H R2,
//   Q = W (R2 R1)^-1 = W R1^-1 R2^-1.
// Adjoint (PyTorch's form of Walter & Lehmann 2018 for a loss of Q only):
//   Wbar = (Qbar + Q copyltu(X)) R^-H,  X = triu(-Q^H Qbar),
//   copyltu(X) = X + X^H - Re diag(X).
#include "qr_kernels.h"

namespace jrx {

GramSplit gram_split(int64_t P, int ng) {
  const int tiles = (ng + kQrRows - 1) / kQrRows;
  int64_t want = (2048 + P - 1) / P;  // chunks per problem
  if (want < 1) want = 1;
  if (want > tiles) want = tiles;
  const int tiles_per_chunk = (int)((tiles + want - 1) / want);
  GramSplit sp;
  sp.rows_per_chunk = tiles_per_chunk * kQrRows;
  sp.nchunk = (ng + sp.rows_per_chunk - 1) / sp.rows_per_chunk;
  return sp;
}

constexpr int kMaxOut = kMaxBands * kMaxBands / kQrBlock;    // 16
constexpr int kMaxTile = kQrRows * kMaxBands / kQrBlock;     // 16

template <bool PRE>
__global__ __launch_bounds__(kQrBlock) void gram_gmajor_kernel(
    const double* __restrict__ wr, const double* __restrict__ wi,
    const double2* __restrict__ pre, double2* __restrict__ partial, int ng,
    int nb, int rows_per_chunk, int nchunk) {
  extern __shared__ double2 sm[];
  const int ld = nb + 1;
  double2* tile = sm;                   // [kQrRows][nb+1]
  double2* pm = sm + kQrRows * ld;      // [nb][nb]
  const int64_t p = blockIdx.y;
  const int chunk = blockIdx.x;
  const int tid = threadIdx.x;
  if (PRE)
    for (int i = tid; i < nb * nb; i += kQrBlock) pm[i] = pre[p * nb * nb + i];
  double2 acc[kMaxOut];
#pragma unroll
  for (int k = 0; k < kMaxOut; ++k) acc[k] = czero();
  const int nout = nb * nb;
  const int gb = chunk * rows_per_chunk;
  const int ge = min(ng, gb + rows_per_chunk);
  for (int gt = gb; gt < ge; gt += kQrRows) {
    const int nr = min(kQrRows, ge - gt);
    __syncthreads();
    const double* r0 = wr + (p * ng + gt) * nb;
    const double* i0 = wi + (p * ng + gt) * nb;
    for (int e = tid; e < nr * nb; e += kQrBlock) {
      const int r = e / nb, c = e - r * nb;
      tile[r * ld + c] = make_double2(r0[e], i0[e]);
    }
    __syncthreads();
    if (PRE) {
      double2 tmp[kMaxTile];
#pragma unroll
      for (int k = 0; k < kMaxTile; ++k) {
        const int e = tid + k * kQrBlock;
        double2 a = czero();
        if (e < nr * nb) {
          const int r = e / nb, c = e - r * nb;
          for (int i = 0; i <= c; ++i)
            a = cadd(a, cmul(tile[r * ld + i], pm[i * nb + c]));
        }
        tmp[k] = a;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kMaxTile; ++k) {
        const int e = tid + k * kQrBlock;
        if (e < nr * nb) {
          const int r = e / nb, c = e - r * nb;
          tile[r * ld + c] = tmp[k];
        }
      }
      __syncthreads();
    }
    for (int r = 0; r < nr; ++r) {
#pragma unroll
      for (int k = 0; k < kMaxOut; ++k) {
        const int o = tid + k * kQrBlock;
        if (o < nout) {
          const int i = o / nb, j = o - i * nb;
          acc[k] = cadd(acc[k], cmulc(tile[r * ld + i], tile[r * ld + j]));
        }
      }
    }
  }
  double2* out = partial + (p * nchunk + chunk) * nout;
#pragma unroll
  for (int k = 0; k < kMaxOut; ++k) {
    const int o = tid + k * kQrBlock;
    if (o < nout) out[o] = acc[k];
  }
}

int launch_gram_gmajor(const double* wr, const double* wi, const double2* pre,
                       double2* partial, int64_t P, int ng, int nb,
                       const GramSplit& sp, hipStream_t s) {
  size_t shm = (size_t)kQrRows * (nb + 1) * sizeof(double2);
  if (pre) shm += (size_t)nb * nb * sizeof(double2);
  dim3 grid(sp.nchunk, (unsigned)P), block(kQrBlock);
  if (pre)
    hipLaunchKernelGGL((gram_gmajor_kernel<true>), grid, block, shm, s, wr, wi,
                       pre, partial, ng, nb, sp.rows_per_chunk, sp.nchunk);
  else
    hipLaunchKernelGGL((gram_gmajor_kernel<false>), grid, block, shm, s, wr,
                       wi, pre, partial, ng, nb, sp.rows_per_chunk, sp.nchunk);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

__global__ __launch_bounds__(kQrBlock) void gram_bmajor_kernel(
    const double2* __restrict__ q, const double2* __restrict__ qb,
    double2* __restrict__ partial, int ng, int nb, int rows_per_chunk,
    int nchunk) {
  extern __shared__ double2 sm[];
  const int ld = kQrRows + 1;
  double2* tq = sm;              // [nb][kQrRows+1]
  double2* tb = sm + nb * ld;    // [nb][kQrRows+1]
  const int64_t p = blockIdx.y;
  const int chunk = blockIdx.x;
  const int tid = threadIdx.x;
  double2 acc[kMaxOut];
#pragma unroll
  for (int k = 0; k < kMaxOut; ++k) acc[k] = czero();
  const int nout = nb * nb;
  const int gb = chunk * rows_per_chunk;
  const int ge = min(ng, gb + rows_per_chunk);
  for (int gt = gb; gt < ge; gt += kQrRows) {
    const int nr = min(kQrRows, ge - gt);
    __syncthreads();
    for (int e = tid; e < nb * nr; e += kQrBlock) {
      const int i = e / nr, r = e - i * nr;
      const int64_t src = (p * nb + i) * ng + gt + r;
      tq[i * ld + r] = q[src];
      tb[i * ld + r] = qb[src];
    }
    __syncthreads();
    for (int r = 0; r < nr; ++r) {
#pragma unroll
      for (int k = 0; k < kMaxOut; ++k) {
        const int o = tid + k * kQrBlock;
        if (o < nout) {
          const int i = o / nb, j = o - i * nb;
          acc[k] = cadd(acc[k], cmulc(tq[i * ld + r], tb[j * ld + r]));
        }
      }
    }
  }
  double2* out = partial + (p * nchunk + chunk) * nout;
#pragma unroll
  for (int k = 0; k < kMaxOut; ++k) {
    const int o = tid + k * kQrBlock;
    if (o < nout) out[o] = acc[k];
  }
}

int launch_gram_bmajor(const double2* q, const double2* qb, double2* partial,
                       int64_t P, int ng, int nb, const GramSplit& sp,
                       hipStream_t s) {
  size_t shm = 2 * (size_t)nb * (kQrRows + 1) * sizeof(double2);
  dim3 grid(sp.nchunk, (unsigned)P), block(kQrBlock);
  hipLaunchKernelGGL(gram_bmajor_kernel, grid, block, shm, s, q, qb, partial,
                     ng, nb, sp.rows_per_chunk, sp.nchunk);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

__global__ __launch_bounds__(kQrBlock) void reduce_partials_kernel(
    const double2* __restrict__ partial, double2* __restrict__ out,
    int nchunk, int nout) {
  const int64_t p = blockIdx.y;
  const int o = blockIdx.x * kQrBlock + threadIdx.x;
  if (o >= nout) return;
  double2 a = czero();
  for (int c = 0; c < nchunk; ++c) a = cadd(a, partial[(p * nchunk + c) * nout + o]);
  out[p * nout + o] = a;
}

int launch_reduce_partials(const double2* partial, double2* out, int64_t P,
                           int nchunk, int nb, hipStream_t s) {
  const int nout = nb * nb;
  dim3 grid(ceil_div(nout, kQrBlock), (unsigned)P), block(kQrBlock);
  hipLaunchKernelGGL(reduce_partials_kernel, grid, block, 0, s, partial, out,
                     nchunk, nout);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

// One block per problem; A and X live in LDS ([nb][nb+1] each).
__global__ __launch_bounds__(kQrBlock) void chol_inv_kernel(
    const double2* __restrict__ S, double2* __restrict__ rinv, int nb) {
  extern __shared__ double2 sm[];
  const int ld = nb + 1;
  double2* A = sm;
  double2* X = sm + nb * ld;
  const int64_t p = blockIdx.x;
  const int tid = threadIdx.x;
  const double2* Sp = S + p * nb * nb;
  for (int e = tid; e < nb * nb; e += kQrBlock) {
    const int i = e / nb, j = e - i * nb;
    A[i * ld + j] = Sp[e];
  }
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    if (tid == 0) {
      const double d = A[j * ld + j].x;
      // NaN for a non-positive pivot: sqrt of a negative number.
      A[j * ld + j] = make_double2(d > 0.0 ? sqrt(d) : sqrt(-1.0 - d * 0.0), 0.0);
    }
    __syncthreads();
    const double d = A[j * ld + j].x;
    for (int c = j + 1 + tid; c < nb; c += kQrBlock)
      A[j * ld + c] = cscale(A[j * ld + c], 1.0 / d);
    __syncthreads();
    const int m = nb - j - 1;
    for (int e = tid; e < m * m; e += kQrBlock) {
      const int r = j + 1 + e / m, c = j + 1 + e % m;
      if (c >= r)
        A[r * ld + c] = csub(A[r * ld + c], cmulc(A[j * ld + r], A[j * ld + c]));
    }
    __syncthreads();
  }
  // X = R^{-1}, column c by thread c: back substitution.
  for (int c = tid; c < nb; c += kQrBlock) {
    for (int i = c; i >= 0; --i) {
      double2 s = (i == c) ? make_double2(1.0, 0.0) : czero();
      for (int k = i + 1; k <= c; ++k) s = csub(s, cmul(A[i * ld + k], X[k * ld + c]));
      X[i * ld + c] = cscale(s, 1.0 / A[i * ld + i].x);
    }
  }
  __syncthreads();
  double2* out = rinv + p * nb * nb;
  for (int e = tid; e < nb * nb; e += kQrBlock) {
    const int i = e / nb, j = e - i * nb;
    out[e] = (j >= i) ? X[i * ld + j] : czero();
  }
}

int launch_chol_inv(const double2* S, double2* rinv, int64_t P, int nb,
                    hipStream_t s) {
  size_t shm = 2 * (size_t)nb * (nb + 1) * sizeof(double2);
  hipLaunchKernelGGL(chol_inv_kernel, dim3((unsigned)P), dim3(kQrBlock), shm,
                     s, S, rinv, nb);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

__global__ __launch_bounds__(kQrBlock) void trmm_uu_kernel(
    const double2* __restrict__ a, const double2* __restrict__ b,
    double2* __restrict__ c, int nb) {
  const int64_t p = blockIdx.y;
  const int e = blockIdx.x * kQrBlock + threadIdx.x;
  if (e >= nb * nb) return;
  const int i = e / nb, j = e - i * nb;
  const double2* ap = a + p * nb * nb;
  const double2* bp = b + p * nb * nb;
  double2 acc = czero();
  for (int k = i; k <= j; ++k) acc = cadd(acc, cmul(ap[i * nb + k], bp[k * nb + j]));
  c[p * nb * nb + e] = acc;
}

int launch_trmm_uu(const double2* a, const double2* b, double2* c, int64_t P,
                   int nb, hipStream_t s) {
  dim3 grid(ceil_div(nb * nb, kQrBlock), (unsigned)P), block(kQrBlock);
  hipLaunchKernelGGL(trmm_uu_kernel, grid, block, 0, s, a, b, c, nb);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

__global__ __launch_bounds__(kQrBlock) void apply_fwd_kernel(
    const double* __restrict__ wr, const double* __restrict__ wi,
    const double2* __restrict__ rinv, double2* __restrict__ q, int ng,
    int nb) {
  extern __shared__ double2 sm[];
  const int ldw = nb + 1, ldo = kQrRows + 1;
  double2* tile = sm;                  // [kQrRows][nb+1]
  double2* otile = sm + kQrRows * ldw;  // [nb][kQrRows+1]
  const int64_t p = blockIdx.y;
  const int gt = blockIdx.x * kQrRows;
  const int nr = min(kQrRows, ng - gt);
  const int tid = threadIdx.x;
  const double* r0 = wr + (p * ng + gt) * nb;
  const double* i0 = wi + (p * ng + gt) * nb;
  for (int e = tid; e < nr * nb; e += kQrBlock) {
    const int r = e / nb, c = e - r * nb;
    tile[r * ldw + c] = make_double2(r0[e], i0[e]);
  }
  __syncthreads();
  const double2* R = rinv + p * nb * nb;
  const int nrg = kQrBlock / nb;
  const int j = tid % nb, rg = tid / nb;
  if (rg < nrg) {
    for (int r = rg; r < nr; r += nrg) {
      double2 acc = czero();
      for (int i = 0; i <= j; ++i) acc = cadd(acc, cmul(tile[r * ldw + i], R[i * nb + j]));
      otile[j * ldo + r] = acc;
    }
  }
  __syncthreads();
  for (int e = tid; e < nb * nr; e += kQrBlock) {
    const int b = e / nr, r = e - b * nr;
    q[(p * nb + b) * ng + gt + r] = otile[b * ldo + r];
  }
}

int launch_apply_fwd(const double* wr, const double* wi, const double2* rinv,
                     double2* q, int64_t P, int ng, int nb, hipStream_t s) {
  size_t shm = ((size_t)kQrRows * (nb + 1) + (size_t)nb * (kQrRows + 1)) *
               sizeof(double2);
  dim3 grid(ceil_div(ng, kQrRows), (unsigned)P), block(kQrBlock);
  hipLaunchKernelGGL(apply_fwd_kernel, grid, block, shm, s, wr, wi, rinv, q,
                     ng, nb);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

__global__ __launch_bounds__(kQrBlock) void qr_bwd_small_kernel(
    const double2* __restrict__ G, const double2* __restrict__ rinv,
    double2* __restrict__ A, double2* __restrict__ Bm, int nb) {
  extern __shared__ double2 sm[];
  const int ld = nb + 1;
  double2* M = sm;             // copyltu(triu(-G))
  double2* Ri = sm + nb * ld;  // R^{-1}
  const int64_t p = blockIdx.x;
  const int tid = threadIdx.x;
  const double2* Gp = G + p * nb * nb;
  const double2* Rp = rinv + p * nb * nb;
  for (int e = tid; e < nb * nb; e += kQrBlock) {
    const int i = e / nb, j = e - i * nb;
    double2 m;
    if (i < j) m = make_double2(-Gp[e].x, -Gp[e].y);
    else if (i > j) m = make_double2(-Gp[j * nb + i].x, Gp[j * nb + i].y);
    else m = make_double2(-Gp[e].x, 0.0);
    M[i * ld + j] = m;
    Ri[i * ld + j] = Rp[e];
  }
  __syncthreads();
  double2* Ap = A + p * nb * nb;
  double2* Bp = Bm + p * nb * nb;
  for (int e = tid; e < nb * nb; e += kQrBlock) {
    const int i = e / nb, j = e - i * nb;
    // (R^-H)[k][j] = conj(Ri[j][k]), nonzero for k >= j
    Ap[e] = cconj(Ri[j * ld + i]);
    double2 acc = czero();
    for (int k = j; k < nb; ++k) acc = cadd(acc, cmul(M[i * ld + k], cconj(Ri[j * ld + k])));
    Bp[e] = acc;
  }
}

int launch_qr_bwd_small(const double2* G, const double2* rinv, double2* A,
                        double2* Bm, int64_t P, int nb, hipStream_t s) {
  size_t shm = 2 * (size_t)nb * (nb + 1) * sizeof(double2);
  hipLaunchKernelGGL(qr_bwd_small_kernel, dim3((unsigned)P), dim3(kQrBlock),
                     shm, s, G, rinv, A, Bm, nb);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

__global__ __launch_bounds__(kQrBlock) void apply_bwd_kernel(
    const double2* __restrict__ qbar, const double2* __restrict__ q,
    const double2* __restrict__ A, const double2* __restrict__ Bm,
    double* __restrict__ dwr, double* __restrict__ dwi, int ng, int nb) {
  extern __shared__ double2 sm[];
  const int ld = kQrRows + 1;
  double2* tb = sm;            // [nb][kQrRows+1] Qbar tile
  double2* tq = sm + nb * ld;  // [nb][kQrRows+1] Q tile
  const int64_t p = blockIdx.y;
  const int gt = blockIdx.x * kQrRows;
  const int nr = min(kQrRows, ng - gt);
  const int tid = threadIdx.x;
  for (int e = tid; e < nb * nr; e += kQrBlock) {
    const int i = e / nr, r = e - i * nr;
    const int64_t src = (p * nb + i) * ng + gt + r;
    tb[i * ld + r] = qbar[src];
    tq[i * ld + r] = q[src];
  }
  __syncthreads();
  const double2* Ap = A + p * nb * nb;
  const double2* Bp = Bm + p * nb * nb;
  const int nrg = kQrBlock / nb;
  const int j = tid % nb, rg = tid / nb;
  if (rg >= nrg) return;
  for (int r = rg; r < nr; r += nrg) {
    double2 acc = czero();
    for (int i = j; i < nb; ++i) acc = cadd(acc, cmul(tb[i * ld + r], Ap[i * nb + j]));
    for (int i = 0; i < nb; ++i) acc = cadd(acc, cmul(tq[i * ld + r], Bp[i * nb + j]));
    const int64_t o = (p * ng + gt + r) * nb + j;
    dwr[o] = acc.x;
    dwi[o] = acc.y;
  }
}

int launch_apply_bwd(const double2* qbar, const double2* q, const double2* A,
                     const double2* Bm, double* dwr, double* dwi, int64_t P,
                     int ng, int nb, hipStream_t s) {
  size_t shm = 2 * (size_t)nb * (kQrRows + 1) * sizeof(double2);
  dim3 grid(ceil_div(ng, kQrRows), (unsigned)P), block(kQrBlock);
  hipLaunchKernelGGL(apply_bwd_kernel, grid, block, shm, s, qbar, q, A, Bm,
                     dwr, dwi, ng, nb);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

}  // namespace jrx
