// Exchange-correlation beyond LDA exchange: PW92 correlation, PBE exchange
// and PBE correlation (unpolarised), for the potential stage of the fused
// evaluation and the stand-alone jrx_xc entry point.
//
// Energy (jrystal/_src/energy.py:185-211, xc.py:67-106,222-253):
//   E_xc = (vol/N) sum_r eps(n(r), sigma(r)) n(r),
//   sigma = sum_d |grad_d n|^2,  grad_d n = ifftn(i G_d fftn(n))  (complex,
//   as xc.py:88-106 forms it).
// Its derivative, the adjoint of that discretisation:
//   dE/dn = (vol/N) d(n eps)/dn + 2 Re sum_d ifftn(-i G_d fftn(w grad_d n)),
//   w = (vol/N) d(n eps)/dsigma.
// Forms: PW92 (Perdew & Wang, PRB 45, 13244 (1992), eq. 10, unpolarised
// Table I row; PBE uses libxc's more precise A = 0.0310907), PBE (Perdew,
// Burke & Ernzerhof, PRL 77, 3865 (1996)).  The derivatives below are
// written out by hand from these forms.
#include "energy_kernels.h"
#include "fft_kernels.h"

namespace jrx {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kDensThreshold = 1e-15;

struct Pw92 {
  double A, a1, b1, b2, b3, b4;
};
constexpr Pw92 kPw92 = {0.031091, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294};
constexpr Pw92 kPw92Mod = {0.0310907, 0.21370, 7.5957, 3.5876, 1.6382,
                           0.49294};
constexpr double kKappa = 0.804;
constexpr double kBeta = 0.06672455060314922;

// eps_c(n) and d eps_c / dn of PW92 eq. 10 (p = 1).
__device__ inline void pw92(double n, const Pw92& p, double* ec,
                            double* dec) {
  const double rs = cbrt(3.0 / (4.0 * kPi * n));
  const double srs = sqrt(rs);
  const double q = 2.0 * p.A *
                   (p.b1 * srs + p.b2 * rs + p.b3 * rs * srs + p.b4 * rs * rs);
  const double dq = 2.0 * p.A * (0.5 * p.b1 / srs + p.b2 + 1.5 * p.b3 * srs +
                                 2.0 * p.b4 * rs);
  const double l = log1p(1.0 / q);
  const double g = -2.0 * p.A * (1.0 + p.a1 * rs) * l;
  const double dg = -2.0 * p.A * p.a1 * l +
                    2.0 * p.A * (1.0 + p.a1 * rs) * dq / (q * (q + 1.0));
  *ec = g;
  *dec = dg * (-rs / (3.0 * n));  // drs/dn = -rs / (3 n)
}

// Accumulates e = n eps, dn = d(n eps)/dn, ds = d(n eps)/dsigma over the
// terms selected by `flags` (JRX_XC_* bits).
__device__ inline void xc_point(int flags, double n, double sigma, double* e,
                                double* dn, double* ds) {
  *e = 0.0;
  *dn = 0.0;
  *ds = 0.0;
  if (!(n > kDensThreshold)) return;
  const double n13 = cbrt(n);
  const double cx = 0.75 * cbrt(3.0 / kPi);
  if (flags & JRX_XC_LDA_X) {
    *e += -cx * n13 * n;
    *dn += -(4.0 / 3.0) * cx * n13;
  }
  if (flags & JRX_XC_LDA_C_PW) {
    double ec, dec;
    pw92(n, kPw92, &ec, &dec);
    *e += n * ec;
    *dn += ec + n * dec;
  }
  const double kf = cbrt(3.0 * kPi * kPi * n);
  if (flags & JRX_XC_GGA_X_PBE) {
    const double mu = kBeta * kPi * kPi / 3.0;
    const double s2 = sigma / (4.0 * kf * kf * n * n);
    const double den = 1.0 + mu * s2 / kKappa;
    const double f = 1.0 + kKappa - kKappa / den;
    const double df = mu / (den * den);  // dF / d s^2
    *e += -cx * n13 * n * f;
    *dn += -cx * n13 * ((4.0 / 3.0) * f - (8.0 / 3.0) * s2 * df);
    *ds += -cx * n13 * n * df / (4.0 * kf * kf * n * n);  // ds2/dsigma
  }
  if (flags & JRX_XC_GGA_C_PBE) {
    const double gam = (1.0 - log(2.0)) / (kPi * kPi);
    const double bg = kBeta / gam;
    double ec, dec;
    pw92(n, kPw92Mod, &ec, &dec);
    const double t2 = sigma * kPi / (16.0 * kf * n * n);
    const double ex = exp(-ec / gam);
    const double a = bg / (ex - 1.0);
    const double y = a * t2;
    const double nn = 1.0 + y, dd = 1.0 + y + y * y;
    const double r = bg * t2 * nn / dd;
    const double h = gam * log1p(r);
    const double dh_dr = gam / (1.0 + r);
    const double dr_dt2 = bg * (nn / dd - y * y * (2.0 + y) / (dd * dd));
    const double dr_da = -bg * t2 * t2 * y * (2.0 + y) / (dd * dd);
    const double da_dec = a * a * ex / kBeta;
    const double dh_dec = dh_dr * dr_da * da_dec;
    const double dh_dt2 = dh_dr * dr_dt2;
    *e += n * (ec + h);
    // dt2/dn = -(7/3) t2 / n
    *dn += ec + h + n * (dec * (1.0 + dh_dec)) - (7.0 / 3.0) * t2 * dh_dt2;
    *ds += n * dh_dt2 * kPi / (16.0 * kf * n * n);
  }
}

// Cartesian G of flat box index i (box axes nx, ny, nz of cp).
__device__ inline void box_g(const CellParams& cp, int64_t i, double* g) {
  const int iz = (int)(i % cp.nz);
  const int iy = (int)((i / cp.nz) % cp.ny);
  const int ix = (int)(i / ((int64_t)cp.ny * cp.nz));
  const int mx = ix < (cp.nx + 1) / 2 ? ix : ix - cp.nx;
  const int my = iy < (cp.ny + 1) / 2 ? iy : iy - cp.ny;
  const int mz = iz < (cp.nz + 1) / 2 ? iz : iz - cp.nz;
  for (int d = 0; d < 3; ++d)
    g[d] = mx * cp.B[d] + my * cp.B[3 + d] + mz * cp.B[6 + d];
}

// G of box point i: from the caller's G grid gvec [N][3] when given, else
// from cp's reciprocal basis.
__device__ inline void point_g(const CellParams& cp, const double* gvec,
                               int64_t i, double* g) {
  if (gvec) {
    for (int d = 0; d < 3; ++d) g[d] = gvec[3 * i + d];
  } else {
    box_g(cp, i, g);
  }
}

// gb[d][i] = i G_d rho_rec[i]  (d = 0, 1, 2)
__global__ void xc_grad_kernel(const double2* __restrict__ rho_rec,
                               double2* __restrict__ gb, CellParams cp,
                               const double* __restrict__ gvec, int64_t N) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double g[3];
  point_g(cp, gvec, i, g);
  const double2 r = rho_rec[i];
  for (int d = 0; d < 3; ++d) gb[d * N + i] = make_double2(-g[d] * r.y, g[d] * r.x);
}

// Per point: v[i] = vbase[i] (real part of vbox, or 0) + (vol/N) dn, block
// partials of (vol/N) n eps, and for GGA gb[d][i] <- (vol/N) ds gb[d][i].
__global__ __launch_bounds__(kRedBlock) void xc_point_kernel(
    const double2* __restrict__ vbox, const double* __restrict__ rho,
    double* __restrict__ v, double2* __restrict__ gb,
    double* __restrict__ exc_part, int64_t N, double w, int flags) {
  __shared__ double red[kRedBlock];
  const int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x;
  double exc = 0.0;
  if (i < N) {
    double sigma = 0.0;
    if (gb) {
      for (int d = 0; d < 3; ++d) {
        const double2 a = gb[d * N + i];
        sigma += a.x * a.x + a.y * a.y;
      }
    }
    double e, dn, ds;
    xc_point(flags, rho[i], sigma, &e, &dn, &ds);
    exc = w * e;
    v[i] = (vbox ? vbox[i].x : 0.0) + w * dn;
    if (gb)
      for (int d = 0; d < 3; ++d) gb[d * N + i] = cscale(gb[d * N + i], w * ds);
  }
  const double t = block_sum(exc, red);
  if (threadIdx.x == 0) exc_part[blockIdx.x] = t;
}

// out[i] = sum_d (-i G_d) gb[d][i]  (out may alias gb[0])
__global__ void xc_div_kernel(const double2* gb, double2* out, CellParams cp,
                              const double* __restrict__ gvec, int64_t N) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double g[3];
  point_g(cp, gvec, i, g);
  double2 acc = czero();
  for (int d = 0; d < 3; ++d) {
    const double2 a = gb[d * N + i];
    acc.x += g[d] * a.y;   // -i g (a.x + i a.y) = g a.y - i g a.x
    acc.y -= g[d] * a.x;
  }
  out[i] = acc;
}

__global__ void xc_add_kernel(double* __restrict__ v,
                              const double2* __restrict__ y, int64_t N) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) v[i] += 2.0 * y[i].x;
}

// out[0] = sum_j part[j] (one block, fixed order).
__global__ __launch_bounds__(kRedBlock) void xc_sum_kernel(
    const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double red[kRedBlock];
  double a = 0.0;
  for (int j = threadIdx.x; j < n; j += kRedBlock) a += part[j];
  const double t = block_sum(a, red);
  if (threadIdx.x == 0) out[0] = t;
}

// v[i] *= s[0]
__global__ void xc_scale_kernel(double* __restrict__ v,
                                const double* __restrict__ s, int64_t N) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) v[i] *= s[0];
}

}  // namespace

bool xc_is_gga(int flags) {
  return (flags & (JRX_XC_GGA_X_PBE | JRX_XC_GGA_C_PBE)) != 0;
}

int xc_stage_gradients(const double2* rho_rec, double2* gb,
                       const CellParams& cp, const FftLen& L0,
                       const FftLen& L1, const FftLen& L2, hipStream_t s,
                       const double* gvec) {
  const int64_t N = (int64_t)cp.nx * cp.ny * cp.nz;
  hipLaunchKernelGGL(xc_grad_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, s,
                     rho_rec, gb, cp, gvec, N);
  JRX_LAUNCH_CHECK();
  return fft3d(gb, gb, 3, cp.nx, cp.ny, cp.nz, true, 1.0 / (double)N, L0, L1,
               L2, s);
}

int xc_stage_potential(const double2* vbox, const double* rho, double* v,
                       double2* gb, double* exc_part, int* nparts,
                       const CellParams& cp, int flags, const FftLen& L0,
                       const FftLen& L1, const FftLen& L2, hipStream_t s,
                       const double* gvec) {
  const int64_t N = (int64_t)cp.nx * cp.ny * cp.nz;
  const int nb = ceil_div(N, kRedBlock);
  const bool gga = xc_is_gga(flags);
  hipLaunchKernelGGL(xc_point_kernel, dim3(nb), dim3(kRedBlock), 0, s, vbox,
                     rho, v, gga ? gb : nullptr, exc_part, N,
                     cp.vol / (double)N, flags);
  JRX_LAUNCH_CHECK();
  *nparts = nb;
  if (!gga) return JRX_OK;
  int rc = fft3d(gb, gb, 3, cp.nx, cp.ny, cp.nz, false, 1.0, L0, L1, L2, s);
  if (rc) return rc;
  hipLaunchKernelGGL(xc_div_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, s,
                     gb, gb, cp, gvec, N);
  JRX_LAUNCH_CHECK();
  rc = fft3d(gb, gb, 1, cp.nx, cp.ny, cp.nz, true, 1.0 / (double)N, L0, L1,
             L2, s);
  if (rc) return rc;
  hipLaunchKernelGGL(xc_add_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, s, v,
                     gb, N);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

}  // namespace jrx

using namespace jrx;

namespace {
// workspace of jrx_xc: fftn(rho) box, three gradient boxes (GGA), the
// potential when the caller wants no gradient, and the energy partials
struct XcWs {
  size_t box, gb, v, part, total;
};
XcWs xc_ws_layout(int64_t N, int flags) {
  XcWs w;
  size_t off = 0;
  auto take = [&](size_t b) {
    const size_t o = off;
    off = (off + b + 255) & ~(size_t)255;
    return o;
  };
  w.box = take((size_t)N * sizeof(double2));
  w.gb = take(xc_is_gga(flags) ? 3 * (size_t)N * sizeof(double2) : 0);
  w.v = take((size_t)N * sizeof(double));
  w.part = take((size_t)ceil_div(N, kRedBlock) * sizeof(double));
  w.total = off;
  return w;
}
}  // namespace

extern "C" {

size_t jrx_xc_workspace_bytes(int nx, int ny, int nz, int xc_flags) {
  return xc_ws_layout((int64_t)nx * ny * nz, xc_flags).total;
}

int jrx_xc(const double* rho, const double* gvec, int nx, int ny, int nz,
           double vol, int xc_flags, double* out, double* grad,
           const double* grad_out, void* workspace, size_t workspace_bytes,
           void* stream) {
  JRX_REQUIRE(rho && gvec && out && workspace && nx >= 1 && ny >= 1 &&
                  nz >= 1 && vol > 0.0,
              "invalid xc arguments");
  JRX_REQUIRE(!grad || grad_out, "grad requires grad_out");
  JRX_REQUIRE(xc_flags >= 1 && xc_flags <= 15,
              "xc_flags must be a nonzero OR of JRX_XC_* terms");
  const int64_t N = (int64_t)nx * ny * nz;
  const XcWs w = xc_ws_layout(N, xc_flags);
  JRX_REQUIRE(workspace_bytes >= w.total && ((uintptr_t)workspace & 255) == 0,
              "xc workspace too small or misaligned");
  FftLen Lx, Ly, Lz;
  int rc;
  if ((rc = cached_fftlen(nx, &Lx)) || (rc = cached_fftlen(ny, &Ly)) ||
      (rc = cached_fftlen(nz, &Lz)))
    return rc;
  char* ws = static_cast<char*>(workspace);
  double2* box = reinterpret_cast<double2*>(ws + w.box);
  double2* gb = xc_is_gga(xc_flags) ? reinterpret_cast<double2*>(ws + w.gb)
                                    : nullptr;
  double* v = grad ? grad : reinterpret_cast<double*>(ws + w.v);
  double* part = reinterpret_cast<double*>(ws + w.part);
  hipStream_t s = (hipStream_t)stream;
  CellParams cp{};
  cp.nx = nx;
  cp.ny = ny;
  cp.nz = nz;
  cp.vol = vol;
  if (gb) {
    if ((rc = launch_rho_to_complex(rho, box, 1, N, s)) ||
        (rc = fft3d(box, box, 1, nx, ny, nz, false, 1.0, Lx, Ly, Lz, s)) ||
        (rc = xc_stage_gradients(box, gb, cp, Lx, Ly, Lz, s, gvec)))
      return rc;
  }
  int nparts = 0;
  if ((rc = xc_stage_potential(nullptr, rho, v, gb, part, &nparts, cp,
                               xc_flags, Lx, Ly, Lz, s, gvec)))
    return rc;
  hipLaunchKernelGGL(xc_sum_kernel, dim3(1), dim3(kRedBlock), 0, s, part,
                     nparts, out);
  JRX_LAUNCH_CHECK();
  if (grad) {
    hipLaunchKernelGGL(xc_scale_kernel, dim3(ceil_div(N, 256)), dim3(256), 0,
                       s, grad, grad_out, N);
    JRX_LAUNCH_CHECK();
  }
  return JRX_OK;
}

}  // extern "C"
