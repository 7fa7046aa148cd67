// Kernels behind the reference's public functions (full-box layout):
// coefficient scatter/gather, density from wave functions, box-layout
// kinetic energy, Hartree / external / LDA-x energies of a given density and
// their gradients, the Ewald constant, and the stand-alone QR.
//
// Formulas: utils.py:108-130,257-308 (absolute_square, expand/squeeze),
// pw.py:214-284 (density_grid), energy.py:29-243, potential.py:24-77,
// 121-166, braket.py:27-72, xc.py:54-64, ewald.py:21-86.
// Gradients use the torch convention for complex inputs
// (dL/dRe z + i dL/dIm z) and are scaled by the incoming gradient gout.
#include "energy_kernels.h"
#include "qr_kernels.h"

namespace jrx {
namespace {

constexpr int kB = 256;

__global__ void zero_kernel(double2* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i < n) p[i] = czero();
}

__global__ void scatter_kernel(const double2* __restrict__ q,
                               const int64_t* __restrict__ idx,
                               double2* __restrict__ box, int ng, int64_t N) {
  const int64_t st = blockIdx.y;
  const int g = blockIdx.x * kB + threadIdx.x;
  if (g < ng) box[st * N + idx[g]] = q[st * ng + g];
}

__global__ void gather_kernel(const double2* __restrict__ box,
                              const int64_t* __restrict__ idx,
                              double2* __restrict__ q, int ng, int64_t N) {
  const int64_t st = blockIdx.y;
  const int g = blockIdx.x * kB + threadIdx.x;
  if (g < ng) q[st * ng + g] = box[st * N + idx[g]];
}

// out[s][r] = sum_{states of spin s} occ |psi|^2 (occ given), otherwise
// out[st][r] = |psi_st(r)|^2.
__global__ void density_kernel(const double2* __restrict__ psi,
                               const double* __restrict__ occ,
                               double* __restrict__ out, int64_t sps,
                               int64_t N) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= N) return;
  if (occ) {
    const int64_t s = blockIdx.y;
    double a = 0.0;
    for (int64_t c = 0; c < sps; ++c) {
      const int64_t st = s * sps + c;
      const double2 v = psi[st * N + i];
      a += occ[st] * (v.x * v.x + v.y * v.y);
    }
    out[s * N + i] = a;
  } else {
    const int64_t st = blockIdx.y;
    const double2 v = psi[st * N + i];
    out[st * N + i] = v.x * v.x + v.y * v.y;
  }
}

__global__ void density_bwd_kernel(const double2* __restrict__ psi,
                                   const double* __restrict__ occ,
                                   const double* __restrict__ grho,
                                   double2* __restrict__ gpsi, int64_t sps,
                                   int64_t N) {
  const int64_t st = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= N) return;
  const double w = occ ? 2.0 * occ[st] * grho[(st / sps) * N + i]
                       : 2.0 * grho[st * N + i];
  gpsi[st * N + i] = cscale(psi[st * N + i], w);
}

__device__ __forceinline__ double gk2_at(const double* gvec,
                                         const double* kpts, int64_t i, int k) {
  const double gx = gvec[3 * i] + kpts[3 * k];
  const double gy = gvec[3 * i + 1] + kpts[3 * k + 1];
  const double gz = gvec[3 * i + 2] + kpts[3 * k + 2];
  return gx * gx + gy * gy + gz * gz;
}

__global__ void kinetic_box_kernel(const double2* __restrict__ c,
                                   const double* __restrict__ gvec,
                                   const double* __restrict__ kpts, int nk,
                                   int nb, int64_t N, double* __restrict__ part,
                                   int nblk) {
  __shared__ double red[kB];
  const int64_t st = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  double e = 0.0;
  if (i < N) {
    const double2 v = c[st * N + i];
    e = gk2_at(gvec, kpts, i, (int)((st / nb) % nk)) * (v.x * v.x + v.y * v.y);
  }
  const double t = block_sum(e, red);
  if (threadIdx.x == 0) part[st * nblk + blockIdx.x] = t;
}

// per[st] = 0.5 sum_b part[st][b]
__global__ void per_state_kernel(const double* __restrict__ part, int nblk,
                                 int64_t S, double* __restrict__ per) {
  const int64_t st = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (st >= S) return;
  double a = 0.0;
  for (int b = 0; b < nblk; ++b) a += part[st * nblk + b];
  per[st] = 0.5 * a;
}

// out[0] = scale * sum_i w[i] v[i] (w may be null: plain sum), one block.
__global__ void dot_kernel(const double* __restrict__ v,
                           const double* __restrict__ w, int64_t n,
                           double scale, double* __restrict__ out) {
  __shared__ double red[kB];
  double a = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kB) a += w ? w[i] * v[i] : v[i];
  const double t = block_sum(a, red);
  if (threadIdx.x == 0) out[0] = scale * t;
}

__global__ void kinetic_box_bwd_kernel(const double2* __restrict__ c,
                                       const double* __restrict__ gvec,
                                       const double* __restrict__ kpts,
                                       const double* __restrict__ occ,
                                       const double* __restrict__ gout,
                                       int nk, int nb, int64_t N,
                                       double2* __restrict__ gc) {
  const int64_t st = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= N) return;
  const double s = occ ? gout[0] * occ[st] : gout[st];
  gc[st * N + i] =
      cscale(c[st * N + i], s * gk2_at(gvec, kpts, i, (int)((st / nb) % nk)));
}

// E_H partials: (2 pi vol / N^2) sum_{G!=0} |rho_tot|^2 / |G|^2 (doubled for
// kohn_sham, whose potential is held fixed: gradient c * rho_tot).
__global__ void hartree_kernel(const double2* __restrict__ rr,
                               const double* __restrict__ gvec, int ns,
                               int64_t N, double vol, int ks,
                               double* __restrict__ part,
                               double2* __restrict__ grad,
                               const double* __restrict__ gout) {
  __shared__ double red[kB];
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  double e = 0.0;
  if (i < N) {
    double2 r = czero();
    for (int s = 0; s < ns; ++s) r = cadd(r, rr[s * N + i]);
    const double g2 = gvec[3 * i] * gvec[3 * i] +
                      gvec[3 * i + 1] * gvec[3 * i + 1] +
                      gvec[3 * i + 2] * gvec[3 * i + 2];
    const double dN = (double)N;
    const double c =
        (i == 0) ? 0.0 : (ks ? 2.0 : 1.0) * 2.0 * M_PI * vol / (dN * dN) / g2;
    e = c * (r.x * r.x + r.y * r.y);
    if (grad) {
      const double2 gv = cscale(r, gout[0] * (ks ? c : 2.0 * c));
      for (int s = 0; s < ns; ++s) grad[s * N + i] = gv;
    }
  }
  const double t = block_sum(e, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// E_ext partials: -(4 pi / N) Re sum_{G!=0} S(G) rho_tot(G) / (|G|^2+1e-10).
__global__ void external_kernel(const double2* __restrict__ rr,
                                const double* __restrict__ gvec,
                                const double* __restrict__ pos,
                                const double* __restrict__ chg, int natom,
                                int ns, int64_t N, double* __restrict__ part,
                                double2* __restrict__ grad,
                                const double* __restrict__ gout) {
  __shared__ double red[kB];
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  double e = 0.0;
  if (i < N) {
    const double gx = gvec[3 * i], gy = gvec[3 * i + 1], gz = gvec[3 * i + 2];
    const double g2 = gx * gx + gy * gy + gz * gz;
    double sr = 0.0, si = 0.0;
    if (i != 0) {
      for (int a = 0; a < natom; ++a) {
        double sn, cs;
        sincos(gx * pos[3 * a] + gy * pos[3 * a + 1] + gz * pos[3 * a + 2],
               &sn, &cs);
        sr += chg[a] * cs;
        si += chg[a] * sn;
      }
    }
    const double c = (i == 0) ? 0.0 : -4.0 * M_PI / (double)N / (g2 + 1e-10);
    double2 r = czero();
    for (int s = 0; s < ns; ++s) r = cadd(r, rr[s * N + i]);
    e = c * (sr * r.x - si * r.y);
    if (grad) {
      const double2 gv = make_double2(gout[0] * c * sr, -gout[0] * c * si);
      for (int s = 0; s < ns; ++s) grad[s * N + i] = gv;
    }
  }
  const double t = block_sum(e, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__device__ __forceinline__ double lda_eps(double n, double cx) {
  return n > 1e-15 ? -cx * cbrt(n) : 0.0;
}
__device__ __forceinline__ double lda_deps(double n, double cx) {
  if (!(n > 1e-15)) return 0.0;
  const double c = cbrt(n);
  return -cx / (3.0 * c * c);
}

__global__ void xc_lda_kernel(const double* __restrict__ rho, int ns,
                              int64_t N, double vol, double cx,
                              double* __restrict__ part,
                              double* __restrict__ grad,
                              const double* __restrict__ gout) {
  __shared__ double red[kB];
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  double e = 0.0;
  if (i < N) {
    const double w = vol / (double)N;
    if (ns == 1) {
      const double r = rho[i];
      const double ex = lda_eps(r, cx);
      e = w * ex * r;
      if (grad) grad[i] = gout[0] * w * (4.0 / 3.0) * ex;
    } else {
      const double ru = rho[i], rd = rho[N + i], rt = ru + rd;
      const double ex = 0.5 * (lda_eps(2 * ru, cx) + lda_eps(2 * rd, cx));
      e = w * ex * rt;
      if (grad) {
        grad[i] = gout[0] * w * (ex + rt * lda_deps(2 * ru, cx));
        grad[N + i] = gout[0] * w * (ex + rt * lda_deps(2 * rd, cx));
      }
    }
  }
  const double t = block_sum(e, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Ewald: real-space pair sums over the translation list (ewald.py:52-60)
// and the reciprocal sum over the box in structure-factor form,
//   sum_{ab} Z_a Z_b cos(G.(R_b - R_a)) = |sum_a Z_a e^{i G.R_a}|^2.
__global__ void ewald_real_kernel(const double* __restrict__ pos,
                                  const double* __restrict__ chg, int natom,
                                  const double* __restrict__ trans,
                                  int64_t nt, double eta,
                                  double* __restrict__ part) {
  __shared__ double red[kB];
  const int64_t t = (int64_t)blockIdx.x * kB + threadIdx.x;
  double e = 0.0;
  if (t < nt) {
    const double tx = trans[3 * t], ty = trans[3 * t + 1], tz = trans[3 * t + 2];
    for (int a = 0; a < natom; ++a)
      for (int b = 0; b < natom; ++b) {
        const double dx = pos[3 * b] - pos[3 * a] - tx;
        const double dy = pos[3 * b + 1] - pos[3 * a + 1] - ty;
        const double dz = pos[3 * b + 2] - pos[3 * a + 2] - tz;
        double r = sqrt(dx * dx + dy * dy + dz * dz + 1e-20);
        if (r <= 1e-9) r = 1e20;
        e += chg[a] * chg[b] * erfc(eta * r) / r;
      }
  }
  const double s = block_sum(e, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void ewald_rec_kernel(const double* __restrict__ pos,
                                 const double* __restrict__ chg, int natom,
                                 const double* __restrict__ gvec, int64_t N,
                                 double eta, double* __restrict__ part) {
  __shared__ double red[kB];
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  double e = 0.0;
  if (i > 0 && i < N) {
    const double gx = gvec[3 * i], gy = gvec[3 * i + 1], gz = gvec[3 * i + 2];
    const double g2 = gx * gx + gy * gy + gz * gz;
    double sr = 0.0, si = 0.0;
    for (int a = 0; a < natom; ++a) {
      double sn, cs;
      sincos(gx * pos[3 * a] + gy * pos[3 * a + 1] + gz * pos[3 * a + 2], &sn,
             &cs);
      sr += chg[a] * cs;
      si += chg[a] * sn;
    }
    e = exp(-g2 / 4.0 / (eta * eta)) / g2 * (sr * sr + si * si);
  }
  const double s = block_sum(e, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void ewald_finalize_kernel(const double* __restrict__ preal,
                                      int nreal,
                                      const double* __restrict__ prec,
                                      int nrec, const double* __restrict__ chg,
                                      int natom, double vol, double eta,
                                      double* __restrict__ out) {
  __shared__ double red[kB];
  double a = 0.0;
  for (int i = threadIdx.x; i < nreal; i += kB) a += preal[i];
  const double sreal = block_sum(a, red);
  a = 0.0;
  for (int i = threadIdx.x; i < nrec; i += kB) a += prec[i];
  const double srec = block_sum(a, red);
  if (threadIdx.x == 0) {
    double z2 = 0.0, z = 0.0;
    for (int i = 0; i < natom; ++i) {
      z2 += chg[i] * chg[i];
      z += chg[i];
    }
    double e = 0.5 * (sreal + srec * 4.0 * M_PI / vol);
    e -= z2 * 2.0 * eta / sqrt(M_PI) / 2.0;
    e -= z * z * M_PI / (eta * eta) / vol / 2.0;
    out[0] = e;
  }
}

__global__ void real_to_complex_kernel(const double* __restrict__ x,
                                       double2* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i < n) y[i] = make_double2(x[i], 0.0);
}

__global__ void complex_real_kernel(const double2* __restrict__ x,
                                    double* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i < n) y[i] = x[i].x;
}

__global__ void scale_kernel(const double* __restrict__ x,
                             const double* __restrict__ alpha,
                             double* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i < n) y[i] = alpha[0] * x[i];
}

// optax.adam (scale_by_adam + scale(-lr)) on a flat float64 parameter:
//   m = b1 m + (1-b1) g,  v = b2 v + (1-b2) g^2,
//   p -= lr * (m / (1-b1^t)) / (sqrt(v / (1-b2^t) + eps_root) + eps).
__global__ void adam_kernel(double* __restrict__ p, const double* __restrict__ g,
                            double* __restrict__ m, double* __restrict__ v,
                            int64_t n, double lr, double b1, double b2,
                            double eps, double eps_root, double c1,
                            double c2) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const double gi = g[i];
  const double mi = b1 * m[i] + (1.0 - b1) * gi;
  const double vi = b2 * v[i] + (1.0 - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrt(vi / c2 + eps_root) + eps);
}

// Deterministic dot product: block b sums the contiguous slice
// [b*chunk, (b+1)*chunk) (fixed thread order and tree), then one block sums
// the block partials.
__global__ void vec_dot_part_kernel(const double* __restrict__ a,
                                    const double* __restrict__ b, int64_t n,
                                    int64_t chunk, double* __restrict__ part) {
  __shared__ double red[kB];
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < n ? lo + chunk : n;
  double acc = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kB) acc += a[i] * b[i];
  const double t = block_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// y and out may be the same buffer (in-place update).
__global__ void axpby_kernel(double alpha, const double* __restrict__ x,
                             double beta, const double* y, double* out,
                             int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i < n) out[i] = alpha * x[i] + beta * y[i];
}

inline dim3 grid2(int64_t n, int64_t y) {
  return dim3((unsigned)ceil_div(n, kB), (unsigned)y);
}

inline size_t parts_bytes(int64_t n) {
  return ((size_t)ceil_div(n, kB) + 8) * sizeof(double);
}

}  // namespace
}  // namespace jrx

using namespace jrx;

extern "C" {

size_t jrx_reduce_workspace_bytes(int64_t n) { return 2 * parts_bytes(n); }

int jrx_adam_update(double* param, const double* grad, double* m, double* v,
                    int64_t n, double lr, double b1, double b2, double eps,
                    double eps_root, int64_t step, void* stream) {
  JRX_REQUIRE(param && grad && m && v && n >= 0 && step >= 1,
              "invalid adam arguments");
  if (n == 0) return JRX_OK;
  const double c1 = 1.0 - pow(b1, (double)step);
  const double c2 = 1.0 - pow(b2, (double)step);
  hipLaunchKernelGGL(adam_kernel, dim3(ceil_div(n, kB)), dim3(kB), 0,
                     (hipStream_t)stream, param, grad, m, v, n, lr, b1, b2,
                     eps, eps_root, c1, c2);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

constexpr int kDotBlocks = 1024;

size_t jrx_dot_workspace_bytes(void) { return (kDotBlocks + 8) * sizeof(double); }

int jrx_vec_dot(const double* a, const double* b, int64_t n, double* out,
                void* workspace, size_t workspace_bytes, void* stream) {
  JRX_REQUIRE(a && b && out && workspace && n >= 0, "invalid dot arguments");
  JRX_REQUIRE(workspace_bytes >= jrx_dot_workspace_bytes(),
              "dot workspace too small");
  int64_t nb = ceil_div(n, (int64_t)kB * 16);
  if (nb < 1) nb = 1;
  if (nb > kDotBlocks) nb = kDotBlocks;
  const int64_t chunk = (n + nb - 1) / nb;
  hipStream_t s = (hipStream_t)stream;
  double* part = (double*)workspace;
  hipLaunchKernelGGL(vec_dot_part_kernel, dim3((unsigned)nb), dim3(kB), 0, s,
                     a, b, n, chunk > 0 ? chunk : 1, part);
  JRX_LAUNCH_CHECK();
  hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(kB), 0, s, part, nullptr, nb,
                     1.0, out);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int jrx_vec_axpby(double alpha, const double* x, double beta, const double* y,
                  double* out, int64_t n, void* stream) {
  JRX_REQUIRE(x && y && out && n >= 0, "invalid axpby arguments");
  if (n == 0) return JRX_OK;
  hipLaunchKernelGGL(axpby_kernel, dim3(ceil_div(n, kB)), dim3(kB), 0,
                     (hipStream_t)stream, alpha, x, beta, y, out, n);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int jrx_real_to_complex(const double* x, double* y, int64_t n, void* stream) {
  JRX_REQUIRE(x && y && n >= 0, "invalid real_to_complex arguments");
  if (n == 0) return JRX_OK;
  hipLaunchKernelGGL(real_to_complex_kernel, dim3(ceil_div(n, kB)), dim3(kB),
                     0, (hipStream_t)stream, x, (double2*)y, n);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int jrx_complex_real(const double* x, double* y, int64_t n, void* stream) {
  JRX_REQUIRE(x && y && n >= 0, "invalid complex_real arguments");
  if (n == 0) return JRX_OK;
  hipLaunchKernelGGL(complex_real_kernel, dim3(ceil_div(n, kB)), dim3(kB), 0,
                     (hipStream_t)stream, (const double2*)x, y, n);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int jrx_scale(const double* x, const double* alpha, double* y, int64_t n,
              void* stream) {
  JRX_REQUIRE(x && alpha && y && n >= 0, "invalid scale arguments");
  if (n == 0) return JRX_OK;
  hipLaunchKernelGGL(scale_kernel, dim3(ceil_div(n, kB)), dim3(kB), 0,
                     (hipStream_t)stream, x, alpha, y, n);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int jrx_expand_coefficient(const double* q, const int64_t* mask_index,
                           double* box, int64_t nstates, int ng, int64_t N,
                           void* stream) {
  JRX_REQUIRE(q && mask_index && box && nstates >= 0 && ng > 0 && N >= ng,
              "invalid expand_coefficient arguments");
  JRX_REQUIRE(nstates <= 65535, "too many states for one call");
  if (nstates == 0) return JRX_OK;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(zero_kernel, dim3(ceil_div(nstates * N, kB)), dim3(kB), 0,
                     s, (double2*)box, nstates * N);
  JRX_LAUNCH_CHECK();
  hipLaunchKernelGGL(scatter_kernel, grid2(ng, nstates), dim3(kB), 0, s,
                     (const double2*)q, mask_index, (double2*)box, ng, N);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int jrx_squeeze_coefficient(const double* box, const int64_t* mask_index,
                            double* q, int64_t nstates, int ng, int64_t N,
                            void* stream) {
  JRX_REQUIRE(q && mask_index && box && nstates >= 0 && ng > 0 && N >= ng,
              "invalid squeeze_coefficient arguments");
  JRX_REQUIRE(nstates <= 65535, "too many states for one call");
  if (nstates == 0) return JRX_OK;
  hipLaunchKernelGGL(gather_kernel, grid2(ng, nstates), dim3(kB), 0,
                     (hipStream_t)stream, (const double2*)box, mask_index,
                     (double2*)q, ng, N);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int jrx_density(const double* psi, const double* occ, double* out, int ns,
                int64_t states_per_spin, int64_t N, void* stream) {
  JRX_REQUIRE(psi && out && ns >= 1 && states_per_spin >= 1 && N >= 1,
              "invalid density arguments");
  const int64_t ny = occ ? ns : ns * states_per_spin;
  JRX_REQUIRE(ny <= 65535, "too many states for one call");
  hipLaunchKernelGGL(density_kernel, grid2(N, ny), dim3(kB), 0,
                     (hipStream_t)stream, (const double2*)psi, occ, out,
                     states_per_spin, N);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int jrx_density_backward(const double* psi, const double* occ,
                         const double* grad_rho, double* grad_psi, int ns,
                         int64_t states_per_spin, int64_t N, void* stream) {
  JRX_REQUIRE(psi && grad_rho && grad_psi && ns >= 1 && N >= 1 &&
                  states_per_spin >= 1,
              "invalid density_backward arguments");
  const int64_t S = ns * states_per_spin;
  JRX_REQUIRE(S <= 65535, "too many states for one call");
  hipLaunchKernelGGL(density_bwd_kernel, grid2(N, S), dim3(kB), 0,
                     (hipStream_t)stream, (const double2*)psi, occ, grad_rho,
                     (double2*)grad_psi, states_per_spin, N);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

size_t jrx_kinetic_workspace_bytes(int64_t nstates, int64_t N) {
  return ((size_t)nstates * (ceil_div(N, kB) + 1) + 8) * sizeof(double);
}

int jrx_kinetic_box(const double* coeff, const double* gvec,
                    const double* kpts, const double* occ, int ns, int nk,
                    int nb, int64_t N, double* out, void* workspace,
                    size_t workspace_bytes, void* stream) {
  JRX_REQUIRE(coeff && gvec && kpts && out && workspace,
              "invalid kinetic arguments");
  const int64_t S = (int64_t)ns * nk * nb;
  JRX_REQUIRE(S >= 1 && S <= 65535, "invalid number of states");
  JRX_REQUIRE(workspace_bytes >= jrx_kinetic_workspace_bytes(S, N),
              "kinetic workspace too small");
  const int nblk = ceil_div(N, kB);
  hipStream_t s = (hipStream_t)stream;
  double* part = (double*)workspace;
  double* per = part + S * nblk;
  hipLaunchKernelGGL(kinetic_box_kernel, grid2(N, S), dim3(kB), 0, s,
                     (const double2*)coeff, gvec, kpts, nk, nb, N, part, nblk);
  JRX_LAUNCH_CHECK();
  hipLaunchKernelGGL(per_state_kernel, dim3(ceil_div(S, kB)), dim3(kB), 0, s,
                     part, nblk, S, occ ? per : out);
  JRX_LAUNCH_CHECK();
  if (occ) {
    hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(kB), 0, s, per, occ, S, 1.0,
                       out);
    JRX_LAUNCH_CHECK();
  }
  return JRX_OK;
}

int jrx_kinetic_box_backward(const double* coeff, const double* gvec,
                             const double* kpts, const double* occ,
                             const double* grad_out, int ns, int nk, int nb,
                             int64_t N, double* grad_coeff, void* stream) {
  JRX_REQUIRE(coeff && gvec && kpts && grad_out && grad_coeff,
              "invalid kinetic_backward arguments");
  const int64_t S = (int64_t)ns * nk * nb;
  JRX_REQUIRE(S >= 1 && S <= 65535, "invalid number of states");
  hipLaunchKernelGGL(kinetic_box_bwd_kernel, grid2(N, S), dim3(kB), 0,
                     (hipStream_t)stream, (const double2*)coeff, gvec, kpts,
                     occ, grad_out, nk, nb, N, (double2*)grad_coeff);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int jrx_hartree(const double* rho_rec, const double* gvec, int ns, int64_t N,
                double vol, int kohn_sham, double* out, double* grad,
                const double* grad_out, void* workspace,
                size_t workspace_bytes, void* stream) {
  JRX_REQUIRE(rho_rec && gvec && out && workspace && ns >= 1 && N >= 1,
              "invalid hartree arguments");
  JRX_REQUIRE(!grad || grad_out, "grad requires grad_out");
  JRX_REQUIRE(workspace_bytes >= jrx_reduce_workspace_bytes(N),
              "reduction workspace too small");
  const int nb = ceil_div(N, kB);
  hipStream_t s = (hipStream_t)stream;
  double* part = (double*)workspace;
  hipLaunchKernelGGL(hartree_kernel, dim3(nb), dim3(kB), 0, s,
                     (const double2*)rho_rec, gvec, ns, N, vol, kohn_sham,
                     part, (double2*)grad, grad_out);
  JRX_LAUNCH_CHECK();
  hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(kB), 0, s, part, nullptr,
                     (int64_t)nb, 1.0, out);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int jrx_external(const double* rho_rec, const double* gvec,
                 const double* positions, const double* charges, int natom,
                 int ns, int64_t N, double* out, double* grad,
                 const double* grad_out, void* workspace,
                 size_t workspace_bytes, void* stream) {
  JRX_REQUIRE(rho_rec && gvec && out && workspace && ns >= 1 && N >= 1 &&
                  natom >= 0 && (natom == 0 || (positions && charges)),
              "invalid external arguments");
  JRX_REQUIRE(!grad || grad_out, "grad requires grad_out");
  JRX_REQUIRE(workspace_bytes >= jrx_reduce_workspace_bytes(N),
              "reduction workspace too small");
  const int nb = ceil_div(N, kB);
  hipStream_t s = (hipStream_t)stream;
  double* part = (double*)workspace;
  hipLaunchKernelGGL(external_kernel, dim3(nb), dim3(kB), 0, s,
                     (const double2*)rho_rec, gvec, positions, charges, natom,
                     ns, N, part, (double2*)grad, grad_out);
  JRX_LAUNCH_CHECK();
  hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(kB), 0, s, part, nullptr,
                     (int64_t)nb, 1.0, out);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int jrx_xc_lda(const double* rho, int ns, int64_t N, double vol, double* out,
               double* grad, const double* grad_out, void* workspace,
               size_t workspace_bytes, void* stream) {
  JRX_REQUIRE(rho && out && workspace && (ns == 1 || ns == 2) && N >= 1,
              "invalid xc arguments");
  JRX_REQUIRE(!grad || grad_out, "grad requires grad_out");
  JRX_REQUIRE(workspace_bytes >= jrx_reduce_workspace_bytes(N),
              "reduction workspace too small");
  const int nb = ceil_div(N, kB);
  hipStream_t s = (hipStream_t)stream;
  double* part = (double*)workspace;
  hipLaunchKernelGGL(xc_lda_kernel, dim3(nb), dim3(kB), 0, s, rho, ns, N, vol,
                     0.75 * cbrt(3.0 / M_PI), part, grad, grad_out);
  JRX_LAUNCH_CHECK();
  hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(kB), 0, s, part, nullptr,
                     (int64_t)nb, 1.0, out);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

size_t jrx_ewald_workspace_bytes(int64_t n_translations, int64_t N) {
  return parts_bytes(n_translations) + parts_bytes(N);
}

int jrx_ewald(const double* positions, const double* charges, int natom,
              const double* gvec, int64_t N, double vol, double eta,
              const double* translations, int64_t nt, double* out,
              void* workspace, size_t workspace_bytes, void* stream) {
  JRX_REQUIRE(positions && charges && gvec && translations && out &&
                  workspace && natom >= 1 && N >= 1 && nt >= 1,
              "invalid ewald arguments");
  JRX_REQUIRE(workspace_bytes >= jrx_ewald_workspace_bytes(nt, N),
              "ewald workspace too small");
  hipStream_t s = (hipStream_t)stream;
  double* preal = (double*)workspace;
  double* prec = (double*)((char*)workspace + parts_bytes(nt));
  const int nbr = ceil_div(nt, kB), nbk = ceil_div(N, kB);
  hipLaunchKernelGGL(ewald_real_kernel, dim3(nbr), dim3(kB), 0, s, positions,
                     charges, natom, translations, nt, eta, preal);
  JRX_LAUNCH_CHECK();
  hipLaunchKernelGGL(ewald_rec_kernel, dim3(nbk), dim3(kB), 0, s, positions,
                     charges, natom, gvec, N, eta, prec);
  JRX_LAUNCH_CHECK();
  hipLaunchKernelGGL(ewald_finalize_kernel, dim3(1), dim3(kB), 0, s, preal,
                     nbr, prec, nbk, charges, natom, vol, eta, out);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

// ------------------------------------------------------ stand-alone QR --
static size_t qr_ws_layout(int64_t P, int ng, int nb, size_t* o_part,
                           size_t* o_s, size_t* o_a, size_t* o_b) {
  const GramSplit sp = gram_split(P, ng, nb);
  const size_t mat = (size_t)P * nb * nb * sizeof(double2);
  size_t off = 0;
  auto take = [&](size_t b) {
    size_t o = off;
    off = (off + b + 255) & ~(size_t)255;
    return o;
  };
  *o_part = take((size_t)P * sp.nchunk * nb * nb * sizeof(double2));
  *o_s = take(mat);
  *o_a = take(mat);
  *o_b = take(mat);
  return off;
}

size_t jrx_qr_workspace_bytes(int64_t P, int ng, int nb) {
  size_t a, b, c, d;
  return qr_ws_layout(P, ng, nb, &a, &b, &c, &d);
}

int jrx_qr(const double* w_re, const double* w_im, int64_t P, int ng, int nb,
           double* q, double* rinv, void* workspace, size_t workspace_bytes,
           void* stream) {
  JRX_REQUIRE(w_re && w_im && q && rinv && workspace && P >= 1 && ng >= nb &&
                  nb >= 1,
              "invalid qr arguments (need ng >= nb)");
  JRX_REQUIRE(P <= 65535, "too many QR problems");
  if (nb > kMaxBands) {
    set_error("nb above 128 is not supported by this build");
    return JRX_ERR_UNSUPPORTED;
  }
  size_t op, os, oa, ob;
  const size_t need = qr_ws_layout(P, ng, nb, &op, &os, &oa, &ob);
  JRX_REQUIRE(workspace_bytes >= need && ((uintptr_t)workspace & 255) == 0,
              "qr workspace too small or misaligned");
  char* ws = (char*)workspace;
  double2* part = (double2*)(ws + op);
  double2* S = (double2*)(ws + os);
  double2* R1 = (double2*)(ws + oa);
  double2* R2 = (double2*)(ws + ob);
  const GramSplit sp = gram_split(P, ng, nb);
  hipStream_t s = (hipStream_t)stream;
  int rc;
  if ((rc = launch_gram_gmajor(w_re, w_im, nullptr, part, P, ng, nb, sp, s)) ||
      (rc = launch_reduce_partials(part, S, P, sp.nchunk, nb, s)) ||
      (rc = launch_chol_inv(S, R1, P, nb, s)) ||
      (rc = launch_gram_gmajor(w_re, w_im, R1, part, P, ng, nb, sp, s, nullptr,
                               (double2*)q)) ||
      (rc = launch_reduce_partials(part, S, P, sp.nchunk, nb, s)) ||
      (rc = launch_chol_inv(S, R2, P, nb, s)) ||
      (rc = launch_trmm_uu(R1, R2, (double2*)rinv, P, nb, s)) ||
      (rc = launch_apply_fwd(w_re, w_im, (const double2*)rinv, (double2*)q, P,
                             ng, nb, s)))
    return rc;
  return JRX_OK;
}

int jrx_qr_backward(const double* q, const double* qbar, const double* rinv,
                    int64_t P, int ng, int nb, double* dw_re, double* dw_im,
                    void* workspace, size_t workspace_bytes, void* stream) {
  JRX_REQUIRE(q && qbar && rinv && dw_re && dw_im && workspace && P >= 1 &&
                  ng >= nb && nb >= 1,
              "invalid qr_backward arguments");
  JRX_REQUIRE(P <= 65535, "too many QR problems");
  if (nb > kMaxBands) {
    set_error("nb above 128 is not supported by this build");
    return JRX_ERR_UNSUPPORTED;
  }
  size_t op, os, oa, ob;
  const size_t need = qr_ws_layout(P, ng, nb, &op, &os, &oa, &ob);
  JRX_REQUIRE(workspace_bytes >= need && ((uintptr_t)workspace & 255) == 0,
              "qr workspace too small or misaligned");
  char* ws = (char*)workspace;
  double2* part = (double2*)(ws + op);
  double2* G = (double2*)(ws + os);
  double2* A = (double2*)(ws + oa);
  double2* Bm = (double2*)(ws + ob);
  const GramSplit sp = gram_split(P, ng, nb);
  hipStream_t s = (hipStream_t)stream;
  int rc;
  if ((rc = launch_gram_bmajor((const double2*)q, (const double2*)qbar, part,
                               P, ng, nb, sp, s)) ||
      (rc = launch_reduce_partials(part, G, P, sp.nchunk, nb, s)) ||
      (rc = launch_qr_bwd_small(G, (const double2*)rinv, A, Bm, P, nb, s)) ||
      (rc = launch_apply_bwd((const double2*)qbar, (const double2*)q, A, Bm,
                             dw_re, dw_im, P, ng, nb, s)))
    return rc;
  return JRX_OK;
}

}  // extern "C"
