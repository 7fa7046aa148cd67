// Reciprocal-space Hartree / external energies, LDA exchange, kinetic
// energy and the potential v(r) = dE/drho(r) for the adjoint.
//
// Formulas follow jrystal/_src/potential.py:24-77,121-166 (note the literal
// +1e-10 regulariser of the external term), braket.py:27-72,
// energy.py:29-211 and xc.py:54-64.
#include "energy_kernels.h"

namespace jrx {

__device__ __forceinline__ int freq(int j, int n) {
  return j < (n + 1) / 2 ? j : j - n;
}

__global__ __launch_bounds__(kRedBlock) void rho_to_complex_kernel(
    const double* __restrict__ rho, double2* __restrict__ out, int ns,
    int64_t N) {
  const int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x;
  if (i >= N) return;
  double r = 0.0;
  for (int s = 0; s < ns; ++s) r += rho[s * N + i];
  out[i] = make_double2(r, 0.0);
}

int launch_rho_to_complex(const double* rho, double2* out, int ns, int64_t N,
                          hipStream_t s) {
  hipLaunchKernelGGL(rho_to_complex_kernel, dim3(ceil_div(N, kRedBlock)),
                     dim3(kRedBlock), 0, s, rho, out, ns, N);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

// Cartesian G and |G|^2 of flat box index i (box axes nx, ny, nz of cp).
__device__ inline double box_g2(const CellParams& cp, int64_t i, double* g) {
  const int iz = (int)(i % cp.nz);
  const int iy = (int)((i / cp.nz) % cp.ny);
  const int ix = (int)(i / ((int64_t)cp.ny * cp.nz));
  const int mx = freq(ix, cp.nx), my = freq(iy, cp.ny), mz = freq(iz, cp.nz);
  for (int d = 0; d < 3; ++d)
    g[d] = mx * cp.B[d] + my * cp.B[3 + d] + mz * cp.B[6 + d];
  return g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
}

// S(G) = sum_a Z_a e^{+i G.R_a}
__device__ inline double2 structure_factor(const CellParams& cp,
                                           const double* g) {
  double sr = 0.0, si = 0.0;
  for (int a = 0; a < cp.natom; ++a) {
    const double ph =
        g[0] * cp.pos[3 * a] + g[1] * cp.pos[3 * a + 1] + g[2] * cp.pos[3 * a + 2];
    double sn, cs;
    sincos(ph, &sn, &cs);
    sr += cp.charge[a] * cs;
    si += cp.charge[a] * sn;
  }
  return make_double2(sr, si);
}

// out[i] = S(G_i) / (|G_i|^2 + 1e-10), 0 at G = 0.
__global__ __launch_bounds__(kRedBlock) void structure_term_kernel(
    CellParams cp, double2* __restrict__ out) {
  const int64_t N = (int64_t)cp.nx * cp.ny * cp.nz;
  const int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x;
  if (i >= N) return;
  double g[3];
  const double g2 = box_g2(cp, i, g);
  out[i] = i == 0 ? czero() : cscale(structure_factor(cp, g), 1.0 / (g2 + 1e-10));
}

int launch_structure_term(const CellParams& cp, double2* out, hipStream_t s) {
  const int64_t N = (int64_t)cp.nx * cp.ny * cp.nz;
  hipLaunchKernelGGL(structure_term_kernel, dim3(ceil_div(N, kRedBlock)),
                     dim3(kRedBlock), 0, s, cp, out);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

__global__ __launch_bounds__(kRedBlock) void potential_rec_kernel(
    const double2* __restrict__ rho_rec, double2* __restrict__ vrec,
    double* __restrict__ eh_part, double* __restrict__ ee_part,
    CellParams cp, const double2* __restrict__ sterm) {
  __shared__ double red[kRedBlock];
  const int64_t N = (int64_t)cp.nx * cp.ny * cp.nz;
  const int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x;
  double eh = 0.0, ee = 0.0;
  if (i < N) {
    double g[3];
    const double g2 = box_g2(cp, i, g);
    double2 out = czero();
    if (i != 0) {
      // t = S(G) / (|G|^2 + 1e-10), precomputed per plan or formed here
      const double2 t =
          sterm ? sterm[i]
                : cscale(structure_factor(cp, g), 1.0 / (g2 + 1e-10));
      const double2 r = rho_rec[i];
      const double inv_g2 = 1.0 / g2;
      const double dN = (double)N;
      eh = (2.0 * M_PI * cp.vol / (dN * dN)) * (r.x * r.x + r.y * r.y) * inv_g2;
      ee = -(4.0 * M_PI / dN) * (t.x * r.x - t.y * r.y);
      const double ch = 4.0 * M_PI * cp.vol / dN * inv_g2;
      out = make_double2(ch * r.x - 4.0 * M_PI * t.x,
                         ch * r.y + 4.0 * M_PI * t.y);
    }
    vrec[i] = out;
  }
  const double teh = block_sum(eh, red);
  const double tee = block_sum(ee, red);
  if (threadIdx.x == 0) {
    eh_part[blockIdx.x] = teh;
    ee_part[blockIdx.x] = tee;
  }
}

int launch_potential_rec(const double2* rho_rec, double2* vrec,
                         double* eh_part, double* ee_part,
                         const CellParams& cp, int* nparts, hipStream_t s,
                         const double2* sterm) {
  const int64_t N = (int64_t)cp.nx * cp.ny * cp.nz;
  const int nb = ceil_div(N, kRedBlock);
  hipLaunchKernelGGL(potential_rec_kernel, dim3(nb), dim3(kRedBlock), 0, s,
                     rho_rec, vrec, eh_part, ee_part, cp, sterm);
  JRX_LAUNCH_CHECK();
  *nparts = nb;
  return JRX_OK;
}

// eps_x(n) = -cx n^(1/3) above the density threshold, else 0.
__device__ __forceinline__ double lda_x_eps(double n, double cx) {
  return n > 1e-15 ? -cx * cbrt(n) : 0.0;
}
__device__ __forceinline__ double lda_x_deps(double n, double cx) {
  if (!(n > 1e-15)) return 0.0;
  const double c = cbrt(n);
  return -cx / (3.0 * c * c);
}

__global__ __launch_bounds__(kRedBlock) void potential_real_kernel(
    const double2* __restrict__ vbox, const double* __restrict__ rho,
    double* __restrict__ v, double* __restrict__ exc_part, int ns, int64_t N,
    double vol, double cx, int ks) {
  __shared__ double red[kRedBlock];
  const int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x;
  double exc = 0.0;
  if (i < N) {
    const double w = vol / (double)N;
    const double vhe = vbox[i].x;
    if (ns == 1) {
      const double r = rho[i];
      const double e = lda_x_eps(r, cx);
      exc = w * e * r;
      v[i] = vhe + w * (4.0 / 3.0) * e;
    } else {
      // Spin-polarised exchange as in xc.py:56-59: eps = (e(2ru)+e(2rd))/2
      // multiplies the total density.  v_s = d E / d rho_s = eps + rt
      // d eps / d rho_s; the Kohn-Sham form (ks) of the band driver is
      // vxc_lda (xc.py:109-124), eps + rho_s d eps / d rho_s.
      const double ru = rho[i], rd = rho[N + i];
      const double e = 0.5 * (lda_x_eps(2.0 * ru, cx) + lda_x_eps(2.0 * rd, cx));
      const double rt = ru + rd;
      exc = w * e * rt;
      v[i] = vhe + w * (e + (ks ? ru : rt) * lda_x_deps(2.0 * ru, cx));
      v[N + i] = vhe + w * (e + (ks ? rd : rt) * lda_x_deps(2.0 * rd, cx));
    }
  }
  const double t = block_sum(exc, red);
  if (threadIdx.x == 0) exc_part[blockIdx.x] = t;
}

int launch_potential_real(const double2* vbox, const double* rho, double* v,
                          double* exc_part, int ns, int64_t N, double vol,
                          double cx, int* nparts, hipStream_t s, bool ks) {
  const int nb = ceil_div(N, kRedBlock);
  hipLaunchKernelGGL(potential_real_kernel, dim3(nb), dim3(kRedBlock), 0, s,
                     vbox, rho, v, exc_part, ns, N, vol, cx, ks ? 1 : 0);
  JRX_LAUNCH_CHECK();
  *nparts = nb;
  return JRX_OK;
}

// g_m[3g..3g+2] are the integer frequencies (mx, my, mz) of mask point g.
__global__ __launch_bounds__(kRedBlock) void kinetic_kernel(
    const double2* __restrict__ q, const double* __restrict__ occ,
    const double* __restrict__ kpts, const short* __restrict__ g_m,
    CellParams cp, int ng, int nk, int nb, int gblocks,
    double* __restrict__ kin_part) {
  __shared__ double red[kRedBlock];
  const int64_t st = blockIdx.x / gblocks;
  const int gbk = (int)(blockIdx.x % gblocks);
  const int g = gbk * kRedBlock + threadIdx.x;
  double kin = 0.0;
  if (g < ng) {
    const int k = (int)((st / nb) % nk);
    const int mx = g_m[3 * g], my = g_m[3 * g + 1], mz = g_m[3 * g + 2];
    const double gx = mx * cp.B[0] + my * cp.B[3] + mz * cp.B[6] + kpts[3 * k];
    const double gy = mx * cp.B[1] + my * cp.B[4] + mz * cp.B[7] + kpts[3 * k + 1];
    const double gz = mx * cp.B[2] + my * cp.B[5] + mz * cp.B[8] + kpts[3 * k + 2];
    const double2 c = q[st * ng + g];
    kin = (gx * gx + gy * gy + gz * gz) * (c.x * c.x + c.y * c.y);
    (void)occ;
  }
  const double t = block_sum(kin, red);
  if (threadIdx.x == 0) kin_part[st * gblocks + gbk] = t;
}

int launch_kinetic(const double2* q, const double* occ, const double* kpts,
                   const short* g_m, const CellParams& cp, int64_t nstates,
                   int ng, int nk, int nb, double* kin_part, int* nparts,
                   hipStream_t s) {
  const int gb = ceil_div(ng, kRedBlock);
  if ((int64_t)gb * nstates > 0x7fffffff) {
    set_error("too many states for the kinetic reduction");
    return JRX_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL(kinetic_kernel, dim3((unsigned)(gb * nstates)),
                     dim3(kRedBlock), 0, s, q, occ, kpts, g_m, cp, ng, nk, nb,
                     gb, kin_part);
  JRX_LAUNCH_CHECK();
  *nparts = (int)(gb * nstates);
  return JRX_OK;
}

__device__ double strided_sum(const double* p, int n, double* red) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += kRedBlock) acc += p[i];
  return block_sum(acc, red);
}

// tb[st] = sum_j part[st * nper + j] (fixed order): |G+k|^2-weighted norm
// of each state, i.e. twice its kinetic energy.
__global__ __launch_bounds__(kRedBlock) void state_sums_kernel(
    const double* __restrict__ part, int nper, int64_t S,
    double* __restrict__ tb) {
  const int64_t st = (int64_t)blockIdx.x * kRedBlock + threadIdx.x;
  if (st >= S) return;
  double a = 0.0;
  for (int j = 0; j < nper; ++j) a += part[st * nper + j];
  tb[st] = a;
}

int launch_state_sums(const double* part, int nper, int64_t S, double* tb,
                      hipStream_t s) {
  hipLaunchKernelGGL(state_sums_kernel, dim3(ceil_div(S, kRedBlock)),
                     dim3(kRedBlock), 0, s, part, nper, S, tb);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

__global__ __launch_bounds__(kRedBlock) void finalize_energies_kernel(
    const double* tb, const double* occ, int64_t S, const double* eh,
    const double* ee, int nrec, const double* exc, int nexc,
    double* energies) {
  __shared__ double red[kRedBlock];
  double a = 0.0;
  for (int64_t i = threadIdx.x; i < S; i += kRedBlock) a += occ[i] * tb[i];
  const double k = 0.5 * block_sum(a, red);
  const double h = strided_sum(eh, nrec, red);
  const double e = strided_sum(ee, nrec, red);
  const double x = strided_sum(exc, nexc, red);
  if (threadIdx.x == 0) {
    energies[0] = k;
    energies[1] = h;
    energies[2] = e;
    energies[3] = x;
    energies[4] = ((k + h) + e) + x;
  }
}

int launch_finalize_energies(const double* tb, const double* occ, int64_t S,
                             const double* eh, const double* ee, int nrec,
                             const double* exc, int nexc, double* energies,
                             hipStream_t s) {
  hipLaunchKernelGGL(finalize_energies_kernel, dim3(1), dim3(kRedBlock), 0, s,
                     tb, occ, S, eh, ee, nrec, exc, nexc, energies);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

__global__ void fill_kernel(double* p, double v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

int launch_fill(double* p, double v, int64_t n, hipStream_t s) {
  if (n <= 0) return JRX_OK;
  hipLaunchKernelGGL(fill_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, p,
                     v, n);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

}  // namespace jrx
