// Non-local (Kleinman-Bylander) part of a norm-conserving pseudopotential
// on the device, for the layout of the fused evaluation: Q [P][nb][ng]
// (band-major, problem p = (spin, k)), projectors proj [nk][np][ng].
//
//   F[p][j][b]  = sum_g proj[k(p)][j][g] Q[p][b][g]        (no conjugate)
//   e[p][b]     = sum_j |F[p][j][b]|^2 / vol               (per band)
//   E           = sum_{p,b} f[p][b] e[p][b]
//   Qbar[p][b][g] += (2 f[p][b] / vol) sum_j F[p][j][b] conj(proj[k][j][g])
//
// (pseudopotential/nloc.py:153-227 of the reference: hamiltonian_nonlocal
// and energy_nonlocal; the last line is the gradient in the convention
// dE/dRe Q + i dE/dIm Q.)  F is formed in two fixed-order stages: per-chunk
// partial sums over g, then a sum over the chunks.
#include "nonlocal_kernels.h"

namespace jrx {

namespace {

constexpr int kNlBlock = 256;
constexpr int kNlTile = 32;      // g per LDS tile
constexpr int kNlPairs = 2048;   // (j, b) pairs of one block (8 per thread)

// Partial F over the g-range of chunk c: part[p][c][pair], pair = j nb + b,
// for the npb pairs of this block starting at pr0 = blockIdx.z kNlPairs.
// Thread t serves the pairs t / S + (kNlBlock / S) r and, of every tile of
// 32 g, the g with g % S == t % S.  S > 1 only when the block has at most
// kNlBlock / S pairs (a single pass over them), so that few pairs still
// occupy every thread; the S slice sums of a pair are then added in slice
// order.
// Dynamic LDS: the projector tile [32][33] and the Q tile [nb][33].
__global__ __launch_bounds__(kNlBlock) void nl_project_kernel(
    const double2* __restrict__ proj, const double2* __restrict__ q, int nk,
    int np_, int ng, int nb, int rows, int nchunk, int S,
    double2* __restrict__ part) {
  extern __shared__ double2 nl_lds[];
  double2* sp = nl_lds;                                // [32][33]
  double2* sq = nl_lds + kNlTile * (kNlTile + 1);      // [nb][33]
  const int c = blockIdx.x;
  const int64_t p = blockIdx.y;
  const int pr0 = blockIdx.z * kNlPairs;
  const int k = (int)(p % nk);
  const int npairs = np_ * nb;
  const int npb = min(kNlPairs, npairs - pr0);
  const int tid = threadIdx.x;
  const int slots = kNlBlock / S, ps = tid / S, sl = tid - ps * S;
  const int g0 = c * rows, g1 = min(ng, g0 + rows);
  const double2* P = proj + (int64_t)k * np_ * ng;
  const double2* Qp = q + p * nb * ng;
  double2 acc[kNlPairs / kNlBlock];
#pragma unroll
  for (int i = 0; i < kNlPairs / kNlBlock; ++i) acc[i] = czero();
  for (int gt = g0; gt < g1; gt += kNlTile) {
    const int w = min(kNlTile, g1 - gt);
    for (int jt = 0; jt < np_; jt += kNlTile) {  // projector tiles of 32
      const int jn = min(kNlTile, np_ - jt);
      __syncthreads();
      for (int e = tid; e < kNlTile * kNlTile; e += kNlBlock) {
        const int j = e / kNlTile, g = e % kNlTile;
        sp[j * (kNlTile + 1) + g] =
            (j < jn && g < w) ? P[(int64_t)(jt + j) * ng + gt + g] : czero();
      }
      if (jt == 0)
        for (int e = tid; e < nb * kNlTile; e += kNlBlock) {
          const int b = e / kNlTile, g = e % kNlTile;
          sq[b * (kNlTile + 1) + g] =
              g < w ? Qp[(int64_t)b * ng + gt + g] : czero();
        }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < kNlPairs / kNlBlock; ++i) {
        const int lp = ps + i * slots;
        if (lp >= npb) continue;
        const int pair = pr0 + lp;
        const int j = pair / nb, b = pair - j * nb;
        if (j < jt || j >= jt + jn) continue;
        const double2* a_row = sp + (j - jt) * (kNlTile + 1);
        const double2* b_row = sq + b * (kNlTile + 1);
        double2 a = acc[i];
        for (int g = sl; g < w; g += S) a = cadd(a, cmul(a_row[g], b_row[g]));
        acc[i] = a;
      }
    }
  }
  double2* out = part + (p * nchunk + c) * npairs + pr0;
  if (S == 1) {
#pragma unroll
    for (int i = 0; i < kNlPairs / kNlBlock; ++i) {
      const int lp = tid + i * kNlBlock;
      if (lp < npb) out[lp] = acc[i];
    }
    return;
  }
  __syncthreads();  // the tiles are no longer read: reuse them
  nl_lds[tid] = acc[0];
  __syncthreads();
  if (sl == 0 && ps < npb) {
    double2 a = nl_lds[tid];
    for (int t = 1; t < S; ++t) a = cadd(a, nl_lds[tid + t]);
    out[ps] = a;
  }
}

// F[p][pair] = sum_c part[p][c][pair] (fixed order): one thread per output
// with four sums in flight, for few chunks.
__global__ void nl_reduce_kernel(const double2* __restrict__ part,
                                 int64_t P, int nchunk, int npairs,
                                 double2* __restrict__ F) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= P * npairs) return;
  const int64_t p = e / npairs;
  const int pair = (int)(e - p * npairs);
  const double2* src = part + p * nchunk * npairs + pair;
  double2 a = czero(), b = czero(), c2 = czero(), d = czero();
  int c = 0;
  for (; c + 3 < nchunk; c += 4) {
    a = cadd(a, src[(int64_t)c * npairs]);
    b = cadd(b, src[(int64_t)(c + 1) * npairs]);
    c2 = cadd(c2, src[(int64_t)(c + 2) * npairs]);
    d = cadd(d, src[(int64_t)(c + 3) * npairs]);
  }
  for (; c < nchunk; ++c) a = cadd(a, src[(int64_t)c * npairs]);
  F[e] = cadd(cadd(a, b), cadd(c2, d));
}

// The same sum with one wave per output, for many chunks: lane l adds the
// chunks l, l + 64, ... in order, then the lanes are combined by a fixed
// butterfly.  Four waves per block; a wave past the end returns whole.
__global__ __launch_bounds__(256) void nl_reduce_wave_kernel(
    const double2* __restrict__ part, int64_t P, int nchunk, int npairs,
    double2* __restrict__ F) {
  const int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (e >= P * npairs) return;
  const int64_t p = e / npairs;
  const int pair = (int)(e - p * npairs);
  const double2* src = part + p * nchunk * npairs + pair;
  double2 a = czero();
  for (int c = lane; c < nchunk; c += 64) a = cadd(a, src[(int64_t)c * npairs]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a.x += __shfl_xor(a.x, off);
    a.y += __shfl_xor(a.y, off);
  }
  if (lane == 0) F[e] = a;
}

// e[p][b] = sum_j |F[p][j][b]|^2 / vol
__global__ void nl_energy_kernel(const double2* __restrict__ F, int64_t P,
                                 int np_, int nb, double inv_vol,
                                 double* __restrict__ e) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P * nb) return;
  const int64_t p = s / nb;
  const int b = (int)(s - p * nb);
  const double2* Fp = F + p * np_ * nb;
  double a = 0.0;
  for (int j = 0; j < np_; ++j) {
    const double2 v = Fp[j * nb + b];
    a += v.x * v.x + v.y * v.y;
  }
  e[s] = a * inv_vol;
}

// Qbar[p][b][g] += (2 f / vol) sum_j F[p][j][b] conj(proj[k][j][g]); block
// (64 g, problem p), four groups of 64 threads over the bands.
__global__ __launch_bounds__(kNlBlock) void nl_apply_kernel(
    const double2* __restrict__ proj, const double2* __restrict__ F,
    const double* __restrict__ occ, int nk, int np_, int ng, int nb,
    double two_over_vol, double2* __restrict__ qbar) {
  extern __shared__ double2 sf[];  // F[p] [np][nb], scaled by 2 f / vol
  const int64_t p = blockIdx.y;
  const int k = (int)(p % nk);
  const int tid = threadIdx.x;
  for (int e = tid; e < np_ * nb; e += kNlBlock) {
    const int b = e % nb;
    sf[e] = cscale(F[p * np_ * nb + e], two_over_vol * occ[p * nb + b]);
  }
  __syncthreads();
  const int g = blockIdx.x * 64 + (tid & 63);
  if (g >= ng) return;
  const double2* P = proj + (int64_t)k * np_ * ng + g;
  double2* Qb = qbar + p * nb * ng + g;
  for (int b = tid >> 6; b < nb; b += kNlBlock / 64) {
    double2 a = czero();
    for (int j = 0; j < np_; ++j) {
      const double2 pj = P[(int64_t)j * ng];
      const double2 fj = sf[j * nb + b];
      // F conj(P)
      a.x += fj.x * pj.x + fj.y * pj.y;
      a.y += fj.y * pj.x - fj.x * pj.y;
    }
    Qb[(int64_t)b * ng] = cadd(Qb[(int64_t)b * ng], a);
  }
}

// energies[2] += E, energies[4] += E with E = sum f e (one block, fixed
// order).
__global__ __launch_bounds__(256) void nl_finalize_kernel(
    const double* __restrict__ e, const double* __restrict__ occ, int64_t S,
    double* __restrict__ energies) {
  __shared__ double red[4];
  double a = 0.0;
  for (int64_t i = threadIdx.x; i < S; i += 256) a += occ[i] * e[i];
  a = block_sum(a, red);
  if (threadIdx.x == 0) {
    energies[2] += a;
    energies[4] += a;
  }
}

}  // namespace

int nl_rows(int ng, int64_t P) {
  // about 256 blocks or more: enough to spread the projections, few enough
  // partials for the chunk sum
  int rows = 1024;
  while (rows > kNlTile && (int64_t)ceil_div(ng, rows) * P < 256) rows >>= 1;
  return rows;
}

size_t nl_scratch_bytes(int64_t P, int np_, int nb, int ng) {
  const int nchunk = ceil_div(ng, nl_rows(ng, P));
  return (size_t)P * np_ * nb * sizeof(double2) * (size_t)(nchunk + 1) +
         (size_t)P * nb * sizeof(double);
}

int launch_nonlocal(const double2* proj, const double2* q, const double* occ,
                    int64_t P, int nk, int np_, int ng, int nb, double vol,
                    double2* qbar, void* scratch, double* e_out,
                    hipStream_t s) {
  if (np_ <= 0) return JRX_OK;
  if (nb > 128) {
    set_error("non-local projectors: more than 128 bands");
    return JRX_ERR_UNSUPPORTED;
  }
  const int rows = nl_rows(ng, P), nchunk = ceil_div(ng, rows);
  const int npairs = np_ * nb;
  double2* part = static_cast<double2*>(scratch);
  double2* F = part + (size_t)P * nchunk * npairs;
  // g slices per pair: the threads of a block not needed for its pairs
  const int npb = npairs < kNlPairs ? npairs : kNlPairs;
  int S = 1;
  while (S < kNlTile && npb * S * 2 <= kNlBlock) S *= 2;
  const size_t shm =
      (size_t)(kNlTile + nb) * (kNlTile + 1) * sizeof(double2);
  hipLaunchKernelGGL(nl_project_kernel,
                     dim3((unsigned)nchunk, (unsigned)P,
                          (unsigned)ceil_div(npairs, kNlPairs)),
                     dim3(kNlBlock), shm, s, proj, q, nk, np_, ng, nb, rows,
                     nchunk, S, part);
  JRX_LAUNCH_CHECK();
  if (nchunk > 16)
    hipLaunchKernelGGL(nl_reduce_wave_kernel,
                       dim3((unsigned)ceil_div(P * npairs, (int64_t)4)),
                       dim3(256), 0, s, part, P, nchunk, npairs, F);
  else
    hipLaunchKernelGGL(nl_reduce_kernel,
                       dim3((unsigned)ceil_div(P * npairs, (int64_t)256)),
                       dim3(256), 0, s, part, P, nchunk, npairs, F);
  JRX_LAUNCH_CHECK();
  hipLaunchKernelGGL(nl_energy_kernel,
                     dim3((unsigned)ceil_div(P * nb, (int64_t)256)),
                     dim3(256), 0, s, F, P, np_, nb, 1.0 / vol, e_out);
  JRX_LAUNCH_CHECK();
  if (qbar) {
    const size_t shm = (size_t)npairs * sizeof(double2);
    if (shm > 64 * 1024) {
      set_error("non-local projectors: projectors x bands above 4096");
      return JRX_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL(nl_apply_kernel,
                       dim3((unsigned)ceil_div(ng, 64), (unsigned)P),
                       dim3(kNlBlock), shm, s, proj, F, occ, nk, np_, ng, nb,
                       2.0 / vol, qbar);
    JRX_LAUNCH_CHECK();
  }
  return JRX_OK;
}

int launch_nonlocal_finalize(const double* e, const double* occ, int64_t S,
                             double* energies, hipStream_t s) {
  hipLaunchKernelGGL(nl_finalize_kernel, dim3(1), dim3(256), 0, s, e, occ, S,
                     energies);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

}  // namespace jrx
