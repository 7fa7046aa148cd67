// Rayleigh-Ritz rotation of a minimiser's gauge reset, and the eigenvalues
// of the subspace Hamiltonian, on the device.
//
// After an evaluation with a gradient, S = Q^H Qbar holds
// S_ij = 2 f_j q_i^H H q_j (Qbar = 2 H Q F).  Within each run of
// consecutive bands with equal occupation f > 0 of a (spin, k) problem, V
// holds the eigenvectors (ascending eigenvalues) of the subspace
// Hamiltonian h_ij = S_ij / (2 f_j), Hermitised; V is the identity
// elsewhere.  In the whole-block form every band of the problem is one run
// (all f > 0 required), which gives the eigenvalues of q_i^H H q_j over
// all bands.  The eigenvectors come from the cyclic Jacobi
// method for Hermitian matrices (Golub & Van Loan, Matrix Computations,
// 4th ed., 8.5, with a diagonal phase that makes each pivot real), run in
// parallel over disjoint pivot pairs in round-robin order.
#include "qr_kernels.h"

namespace jrx {

namespace {

constexpr int kRitzBlock = 256;
constexpr int kRitzLds = 64;      // runs up to this size live in LDS
constexpr int kRitzSweeps = 30;   // Jacobi sweeps at most

// runs[p][0] = number of runs of problem p; runs[p][1 + 2r], [2 + 2r] =
// first band and length of run r (length >= 2, equal f > 0); whole: the
// one run [0, nb).
__global__ void ritz_runs_kernel(const double* __restrict__ occ, int64_t P,
                                 int nb, int whole, int* __restrict__ runs) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const double* f = occ + p * nb;
  int* out = runs + p * (nb + 1);
  if (whole) {
    out[0] = 1;
    out[1] = 0;
    out[2] = nb;
    return;
  }
  int n = 0, b = 0;
  while (b < nb) {
    int e = b;
    while (e + 1 < nb && f[e + 1] == f[b]) ++e;
    if (f[b] > 0.0 && e > b) {
      out[1 + 2 * n] = b;
      out[2 + 2 * n] = e - b + 1;
      ++n;
    }
    b = e + 1;
  }
  out[0] = n;
}

// V[p] = I
__global__ void ritz_identity_kernel(double2* __restrict__ V, int64_t P,
                                     int nb) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= P * nb * nb) return;
  const int i = (int)((e / nb) % nb), j = (int)(e % nb);
  V[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
}

// Pair t of round r of the circle method over mp (even) indices.
__device__ inline void round_pair(int r, int t, int mp, int* a, int* b) {
  auto pos = [&](int k) { return k == 0 ? 0 : ((k - 1 + r) % (mp - 1)) + 1; };
  *a = pos(t);
  *b = pos(mp - 1 - t);
}

// One block per (run r, problem p).  A (the run's h) and the accumulated
// rotations W are m x m with row stride m + 1, in LDS up to kRitzLds,
// else in the problem's global scratch.  V (optional) receives the
// eigenvectors, evals (optional) the ascending eigenvalues at the run's
// bands; a run with an occupation <= 0 (whole form) gives NaN eigenvalues.
__global__ __launch_bounds__(kRitzBlock) void ritz_jacobi_kernel(
    const double2* __restrict__ S, const double* __restrict__ occ,
    const int* __restrict__ runs, int nb, double2* __restrict__ V,
    double* __restrict__ evals, double2* __restrict__ gscratch) {
  extern __shared__ double2 sm[];
  const int64_t p = blockIdx.y;
  const int* rp = runs + p * (nb + 1);
  const int r = blockIdx.x;
  if (r >= rp[0]) return;  // uniform over the block
  const int b0 = rp[1 + 2 * r], m = rp[2 + 2 * r];
  const int ld = m + 1, mp = m + (m & 1), np = mp / 2;
  const double* f = occ + p * nb + b0;
  double2* A;
  double2* W;
  double2* small = sm;
  if (m <= kRitzLds) {
    A = sm;
    W = sm + m * ld;
    small = W + m * ld;
  } else {
    A = gscratch + p * 2 * (int64_t)nb * (nb + 1);
    W = A + m * ld;
  }
  double2* rot = small;                   // [np]: (c, s)
  double2* ph = small + np;               // [np]: e^{-i phi}
  int* pa = reinterpret_cast<int*>(small + 2 * np);  // [np]
  int* pb = pa + np;                                 // [np]
  double* red = reinterpret_cast<double*>(pb + np);  // [kRitzBlock / 64]
  const int tid = threadIdx.x;
  int bad = 0;
  for (int i = tid; i < m; i += kRitzBlock) bad |= !(f[i] > 0.0);
  if (__syncthreads_or(bad)) {  // uniform
    if (evals)
      for (int i = tid; i < m; i += kRitzBlock)
        evals[p * nb + b0 + i] = __builtin_nan("");
    return;
  }
  const double2* Sp = S + p * nb * nb;
  for (int e = tid; e < m * m; e += kRitzBlock) {
    const int i = e / m, j = e - i * m;
    const double2 x = Sp[(b0 + i) * nb + b0 + j];
    const double2 y = Sp[(b0 + j) * nb + b0 + i];
    // h = (S F^-1 + (S F^-1)^H) / 4
    const double sj = 0.25 / f[j], si = 0.25 / f[i];
    A[i * ld + j] = make_double2(sj * x.x + si * y.x, sj * x.y - si * y.y);
    W[i * ld + j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  for (int sweep = 0; sweep < kRitzSweeps; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int e = tid; e < m * m; e += kRitzBlock) {
      const int i = e / m, j = e - i * m;
      const double2 a = A[i * ld + j];
      if (i < j) off += a.x * a.x + a.y * a.y;
      else if (i == j) dia += a.x * a.x;
    }
    off = block_sum(off, red);
    dia = block_sum(dia, red);
    // converged: off-diagonal norm at 1e-14 of the diagonal (quadratic
    // convergence takes it to rounding within the next sweep)
    if (off <= 1e-28 * dia || off < 1e-300) break;  // uniform
    for (int rr = 0; rr < mp - 1; ++rr) {
      if (tid < np) {
        int i, j;
        round_pair(rr, tid, mp, &i, &j);
        if (i > j) { const int t = i; i = j; j = t; }
        pa[tid] = i;
        pb[tid] = j;
        double c = 1.0, s = 0.0;
        double2 e = make_double2(1.0, 0.0);
        if (j < m) {
          const double app = A[i * ld + i].x, aqq = A[j * ld + j].x;
          const double2 apq = A[i * ld + j];
          const double mag = sqrt(apq.x * apq.x + apq.y * apq.y);
          if (mag > 1e-300 && mag > 1e-17 * (fabs(app) + fabs(aqq))) {
            e = make_double2(apq.x / mag, -apq.y / mag);  // e^{-i phi}
            const double tau = (aqq - app) / (2.0 * mag);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) /
                             (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
          }
        }
        rot[tid] = make_double2(c, s);
        ph[tid] = e;
      }
      __syncthreads();
      // columns: A <- A G, W <- W G with G = [[c, s], [-s e, c e]] on (i, j),
      // e = e^{-i phi}
      for (int e = tid; e < np * m; e += kRitzBlock) {
        const int t = e / m, row = e - t * m;
        const int i = pa[t], j = pb[t];
        if (j >= m) continue;
        const double c = rot[t].x, s = rot[t].y;
        const double2 ep = ph[t];
        for (int w = 0; w < 2; ++w) {
          double2* M = w ? W : A;
          const double2 x = M[row * ld + i], y = cmul(M[row * ld + j], ep);
          M[row * ld + i] = make_double2(c * x.x - s * y.x, c * x.y - s * y.y);
          M[row * ld + j] = make_double2(s * x.x + c * y.x, s * x.y + c * y.y);
        }
      }
      __syncthreads();
      // rows: A <- G^H A
      for (int e = tid; e < np * m; e += kRitzBlock) {
        const int t = e / m, col = e - t * m;
        const int i = pa[t], j = pb[t];
        if (j >= m) continue;
        const double c = rot[t].x, s = rot[t].y;
        const double2 x = A[i * ld + col];
        const double2 y = cmul(cconj(ph[t]), A[j * ld + col]);
        A[i * ld + col] = make_double2(c * x.x - s * y.x, c * x.y - s * y.y);
        A[j * ld + col] = make_double2(s * x.x + c * y.x, s * x.y + c * y.y);
      }
      __syncthreads();
    }
  }
  // ascending eigenvalues: column i of W goes to column rank(i)
  auto rank_of = [&](int i) {
    const double li = A[i * ld + i].x;
    int rank = 0;
    for (int j = 0; j < m; ++j) {
      const double lj = A[j * ld + j].x;
      rank += (lj < li || (lj == li && j < i)) ? 1 : 0;
    }
    return rank;
  };
  if (evals)
    for (int i = tid; i < m; i += kRitzBlock)
      evals[p * nb + b0 + rank_of(i)] = A[i * ld + i].x;
  if (!V) return;
  double2* Vp = V + p * nb * nb;
  for (int e = tid; e < m * m; e += kRitzBlock) {
    const int row = e / m, i = e - row * m;
    Vp[(b0 + row) * nb + b0 + rank_of(i)] = W[row * ld + i];
  }
}

// H[p]_ij = (S_ij / (2 f_j) + conj(S_ji) / (2 f_i)) / 2: q_i^H H q_j of the
// last evaluation with a gradient (NaN where an occupation is <= 0).
__global__ void subspace_h_kernel(const double2* __restrict__ S,
                                  const double* __restrict__ occ, int64_t P,
                                  int nb, double2* __restrict__ H) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= P * nb * nb) return;
  const int64_t p = e / ((int64_t)nb * nb);
  const int i = (int)((e / nb) % nb), j = (int)(e % nb);
  const double fi = occ[p * nb + i], fj = occ[p * nb + j];
  if (!(fi > 0.0) || !(fj > 0.0)) {
    H[e] = make_double2(__builtin_nan(""), __builtin_nan(""));
    return;
  }
  const double2 x = S[e];
  const double2 y = S[(p * nb + j) * nb + i];
  const double sj = 0.25 / fj, si = 0.25 / fi;
  H[e] = make_double2(sj * x.x + si * y.x, sj * x.y - si * y.y);
}

}  // namespace

size_t ritz_scratch_bytes(int64_t P, int nb) {
  size_t b = (size_t)P * (nb + 1) * sizeof(int);
  b = (b + 255) & ~(size_t)255;
  if (nb > kRitzLds) b += (size_t)P * 2 * nb * (nb + 1) * sizeof(double2);
  return b;
}

int launch_ritz(const double2* S, const double* occ, int64_t P, int nb,
                double2* V, double* evals, bool whole, void* scratch,
                hipStream_t s) {
  int* runs = static_cast<int*>(scratch);
  size_t off = ((size_t)P * (nb + 1) * sizeof(int) + 255) & ~(size_t)255;
  double2* g = reinterpret_cast<double2*>(static_cast<char*>(scratch) + off);
  if (V) {
    hipLaunchKernelGGL(ritz_identity_kernel,
                       dim3((unsigned)ceil_div(P * nb * nb, (int64_t)256)),
                       dim3(256), 0, s, V, P, nb);
    JRX_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(ritz_runs_kernel, dim3((unsigned)ceil_div(P, (int64_t)64)),
                     dim3(64), 0, s, occ, P, nb, whole ? 1 : 0, runs);
  JRX_LAUNCH_CHECK();
  const int mmax = nb < kRitzLds ? nb : kRitzLds;
  const int npmax = (nb + 1) / 2;
  const size_t shm = 2 * (size_t)mmax * (mmax + 1) * sizeof(double2) +
                     2 * (size_t)npmax * sizeof(double2) +
                     2 * (size_t)npmax * sizeof(int) +
                     (kRitzBlock / 64) * sizeof(double) + 64;
  hipLaunchKernelGGL(ritz_jacobi_kernel,
                     dim3(whole ? 1u : (unsigned)((nb + 1) / 2), (unsigned)P),
                     dim3(kRitzBlock), shm, s, S, occ, runs, nb, V, evals, g);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int launch_subspace_h(const double2* S, const double* occ, int64_t P, int nb,
                      double2* H, hipStream_t s) {
  hipLaunchKernelGGL(subspace_h_kernel,
                     dim3((unsigned)ceil_div(P * nb * nb, (int64_t)256)),
                     dim3(256), 0, s, S, occ, P, nb, H);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

}  // namespace jrx
