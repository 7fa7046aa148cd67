// FFT kernels: generic batched 3D transforms and the sphere-pruned fused
// passes of the energy/gradient evaluation.
#include "fft_kernels.h"

namespace jrx {

// Lengths with compile-time schedules (RadixOf): the x-line passes use
// all of them; the plane passes use those whose n x n plane fits in LDS
// (n <= 80), for cubic grids.
#define JRX_CT_PLANE_SIZES(X)                                               \
  X(32) X(35) X(36) X(40) X(42) X(45) X(48) X(49) X(50) X(54) X(56) X(60)   \
  X(63) X(64) X(70) X(72) X(75) X(80)
#define JRX_CT_SIZES(X) JRX_CT_PLANE_SIZES(X) X(128)


// ------------------------------------------------------------ generic 3D --
constexpr int kLineBlock = 256;
constexpr int kLineCap = 2048;

// Lines (o, q): base = o*os + q*qs, element stride es.  A block transforms
// LB consecutive q of one o.  ELEM_MAJOR keeps LDS as [j][LB] (lines
// adjacent, for strided axes); otherwise [l][n+1] (for contiguous lines).
// With REAL the input is instead the real box sum_s rin[s][.] (ns spin
// channels of rn points each, added in spin order).
template <bool INV, bool ELEM_MAJOR, bool REAL = false>
__global__ __launch_bounds__(kLineBlock) void line_fft_kernel(
    const double2* __restrict__ in, double2* __restrict__ out, FftLen L,
    int64_t os, int64_t Q, int64_t qs, int64_t es, int LB, double scale,
    const double* __restrict__ rin = nullptr, int ns = 0, int64_t rn = 0) {
  extern __shared__ double2 sm[];
  const int n = L.n;
  const int64_t qblocks = (Q + LB - 1) / LB;
  const int64_t o = blockIdx.x / qblocks;
  const int64_t q0 = (blockIdx.x % qblocks) * LB;
  const int nl = (int)((Q - q0) < LB ? (Q - q0) : LB);
  const int64_t off = o * os + q0 * qs;
  const double2* src = in + off;
  double2* dst = out + off;
  const int tot = nl * n;
  for (int e = threadIdx.x; e < tot; e += kLineBlock) {
    int l, j;
    if (ELEM_MAJOR) { l = e % nl; j = e / nl; } else { j = e % n; l = e / n; }
    double2 x;
    if constexpr (REAL) {
      double r = 0.0;
      for (int c = 0; c < ns; ++c) r += rin[c * rn + off + l * qs + j * es];
      x = make_double2(r, 0.0);
    } else {
      x = src[l * qs + j * es];
    }
    sm[ELEM_MAJOR ? j * LB + l : l * (n + 1) + j] = x;
  }
  __syncthreads();
  if (ELEM_MAJOR)
    lds_fft<INV, kLineBlock, kLineCap, true>(sm, nl, 1, LB, nullptr, L);
  else
    lds_fft<INV, kLineBlock, kLineCap, false>(sm, nl, n + 1, 1, nullptr, L);
  for (int e = threadIdx.x; e < tot; e += kLineBlock) {
    int l, j;
    if (ELEM_MAJOR) { l = e % nl; j = e / nl; } else { j = e % n; l = e / n; }
    dst[l * qs + j * es] =
        cscale(sm[ELEM_MAJOR ? j * LB + l : l * (n + 1) + j], scale);
  }
}

int launch_line_fft(const double2* in, double2* out, int64_t B, int n0, int n1,
                    int n2, int axis, bool inverse, double scale,
                    const FftLen& L, hipStream_t s, const double* rin,
                    int ns) {
  if (rin && (axis != 2 || inverse || B != 1)) {
    set_error("launch_line_fft: a real input is for one forward z pass");
    return JRX_ERR_INVALID;
  }
  const int n = L.n;
  if (n > kLineCap) {
    set_error("FFT length above 2048 is not supported");
    return JRX_ERR_UNSUPPORTED;
  }
  int LB = kLineCap / n;
  int64_t nouter, os, Q, qs, es;
  bool elem_major;
  if (axis == 2) {
    nouter = 1; os = 0; Q = B * n0 * n1; qs = n2; es = 1; elem_major = false;
  } else if (axis == 1) {
    nouter = B * n0; os = (int64_t)n1 * n2; Q = n2; qs = 1; es = n2;
    elem_major = true;
  } else {
    nouter = B; os = (int64_t)n0 * n1 * n2; Q = (int64_t)n1 * n2; qs = 1;
    es = (int64_t)n1 * n2; elem_major = true;
  }
  // Lines per block: few enough that one box spreads over about 1024
  // blocks, but at least 8 (on the strided axes 128 bytes of each element
  // row; one line per block measured slower on the contiguous axis too);
  // large batches keep the LDS-filling kLineCap / n.
  const int64_t lines = nouter * Q;
  int want = lines / 1024 > 1 ? (int)(lines / 1024 < LB ? lines / 1024 : LB)
                              : 1;
  if (want < 8) want = 8;
  if (want < LB) LB = want;
  if (Q < LB) LB = (int)Q;
  const int64_t qblocks = (Q + LB - 1) / LB;
  const int64_t nblocks = nouter * qblocks;
  if (nblocks <= 0) return JRX_OK;
  if (nblocks > 0x7fffffff) {
    set_error("FFT batch too large");
    return JRX_ERR_UNSUPPORTED;
  }
  size_t shm = elem_major ? (size_t)n * LB * sizeof(double2)
                          : (size_t)LB * (n + 1) * sizeof(double2);
  dim3 grid((unsigned)nblocks), block(kLineBlock);
  if (rin) {
    hipLaunchKernelGGL((line_fft_kernel<false, false, true>), grid, block,
                       shm, s, nullptr, out, L, os, Q, qs, es, LB, scale, rin,
                       ns, (int64_t)n0 * n1 * n2);
    JRX_LAUNCH_CHECK();
    return JRX_OK;
  }
  if (inverse) {
    if (elem_major)
      hipLaunchKernelGGL((line_fft_kernel<true, true>), grid, block, shm, s,
                         in, out, L, os, Q, qs, es, LB, scale);
    else
      hipLaunchKernelGGL((line_fft_kernel<true, false>), grid, block, shm, s,
                         in, out, L, os, Q, qs, es, LB, scale);
  } else {
    if (elem_major)
      hipLaunchKernelGGL((line_fft_kernel<false, true>), grid, block, shm, s,
                         in, out, L, os, Q, qs, es, LB, scale);
    else
      hipLaunchKernelGGL((line_fft_kernel<false, false>), grid, block, shm, s,
                         in, out, L, os, Q, qs, es, LB, scale);
  }
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int fft3d(const double2* in, double2* out, int64_t B, int n0, int n1, int n2,
          bool inverse, double scale, const FftLen& Lx, const FftLen& Ly,
          const FftLen& Lz, hipStream_t s) {
  int rc = launch_line_fft(in, out, B, n0, n1, n2, 2, inverse, 1.0, Lz, s);
  if (rc) return rc;
  rc = launch_line_fft(out, out, B, n0, n1, n2, 1, inverse, 1.0, Ly, s);
  if (rc) return rc;
  return launch_line_fft(out, out, B, n0, n1, n2, 0, inverse, scale, Lx, s);
}

int fft3d_real_forward(const double* rin, int ns, double2* out, int n0,
                       int n1, int n2, const FftLen& Lx, const FftLen& Ly,
                       const FftLen& Lz, hipStream_t s) {
  int rc = launch_line_fft(nullptr, out, 1, n0, n1, n2, 2, false, 1.0, Lz, s,
                           rin, ns);
  if (rc) return rc;
  rc = launch_line_fft(out, out, 1, n0, n1, n2, 1, false, 1.0, Ly, s);
  if (rc) return rc;
  return launch_line_fft(out, out, 1, n0, n1, n2, 0, false, 1.0, Lx, s);
}

// ------------------------------------------------------- fused, pruned --
constexpr int kPlaneBlock = 512;

__device__ __forceinline__ int fftfreq_index(int j, int n) {
  return j < (n + 1) / 2 ? j : j - n;
}

// Scatter the mask points of one x-plane of one state into LDS, inverse
// transform along z (only rows holding mask points) and along y, store the
// plane.  NY, NZ > 0 select a compile-time schedule.
template <int NY, int NZ, int CAP>
__global__ __launch_bounds__(kPlaneBlock) void plane_ifft_kernel(
    const double2* __restrict__ q, double2* __restrict__ T, int64_t state0,
    BoxGeom geo, FftLen Ly, FftLen Lz) {
  extern __shared__ double2 sm[];
  const int ny = NY > 0 ? NY : geo.ny, nz = NZ > 0 ? NZ : geo.nz;
  const int P = ny * nz;
  const int xi = blockIdx.x;
  const int c = blockIdx.y;
  const int64_t st = state0 + c;
  for (int i = threadIdx.x; i < P; i += kPlaneBlock) sm[i] = czero();
  const int r0 = geo.row_off[xi];
  const int nr = geo.row_off[xi + 1] - r0;
  const short* rows = geo.rows + r0;  // y of the rows holding mask points
  __syncthreads();
  const int g0 = geo.plane_g0[xi], g1 = geo.plane_g0[xi + 1];
  const double2* qs = q + st * geo.ng;
  for (int g = g0 + threadIdx.x; g < g1; g += kPlaneBlock)
    sm[geo.g_yz[g]] = qs[g];
  __syncthreads();
  fft_lines<NZ, true, kPlaneBlock, CAP, false>(sm, nr, nz, 1, rows, Lz);
  fft_lines<NY, true, kPlaneBlock, CAP, true>(sm, nz, 1, nz, nullptr, Ly);
  double2* dst = T + ((int64_t)c * geo.nxa + xi) * geo.ps;
  for (int i = threadIdx.x; i < P; i += kPlaneBlock) dst[i] = sm[i];
}

int launch_plane_ifft(const double2* q, double2* T, int64_t state0,
                      int nstates, const BoxGeom& geo, const FftLen& Ly,
                      const FftLen& Lz, hipStream_t s) {
  const int P = geo.ny * geo.nz;
  size_t shm = (size_t)P * sizeof(double2);
  dim3 grid(geo.nxa, nstates), block(kPlaneBlock);
  const int cube = geo.ny == geo.nz ? geo.ny : 0;
#define JRX_PI(NN)                                                          \
  case NN:                                                                  \
    hipLaunchKernelGGL((plane_ifft_kernel<NN, NN, NN * NN>), grid, block,   \
                       shm, s, q, T, state0, geo, Ly, Lz);                  \
    break;
  switch (cube) {
    JRX_CT_PLANE_SIZES(JRX_PI)
    default:
      hipLaunchKernelGGL((plane_ifft_kernel<0, 0, kPlaneCap>), grid, block,
                         shm, s, q, T, state0, geo, Ly, Lz);
  }
#undef JRX_PI
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

constexpr int kXBlock = 256;
constexpr int kXCap = 2048;
constexpr int kXEpt = kXCap / kXBlock;


// The x-line passes stage LZ lines of nx points in LDS, slot e = x * LZ + l
// (line l = z - z0).  Thread t owns slots t + k * kXBlock; goff[k] is the
// slot's offset in one state's plane buffer, or -1 when the slot is outside
// the block's lines or its x-plane holds no mask point (those slots stay 0).
__device__ __forceinline__ void xline_slots(int (&goff)[kXEpt],
                                            const BoxGeom& geo, int nxl,
                                            int LZ, int nl) {
#pragma unroll
  for (int k = 0; k < kXEpt; ++k) {
    const int e = threadIdx.x + k * kXBlock;
    const int x = e / LZ, l = e - x * LZ;
    const int xi = (e < nxl && l < nl) ? geo.xmap[x] : -1;
    goff[k] = xi >= 0 ? xi * geo.ps + l : -1;
  }
}

// Registers of one state's slots (prefetched while the previous state's
// lines are transformed).
__device__ __forceinline__ void xline_load(double2 (&pre)[kXEpt],
                                           const int (&goff)[kXEpt],
                                           const double2* src) {
#pragma unroll
  for (int k = 0; k < kXEpt; ++k)
    pre[k] = goff[k] >= 0 ? nt_load(src + goff[k]) : czero();
}

int xlines_per_block(int nx, int nz) {
  int LZ = kXCap / nx;
  if (LZ > kXLines) LZ = kXLines;
  if (LZ < 1) return 1;
  // as few z-tiles as this allows, of equal size (48 lines: 2 x 24, not
  // 32 + 16)
  const int tiles = (nz + LZ - 1) / LZ;
  return (nz + tiles - 1) / tiles;
}

int xline_groups(const BoxGeom& geo, int nstates, int target_blocks) {
  const int LZ = xlines_per_block(geo.nx, geo.nz);
  const int tiles = geo.ny * ((geo.nz + LZ - 1) / LZ);
  const int target = target_blocks > 0 ? target_blocks : 4096;
  int G = (target + tiles - 1) / tiles;
  if (G > kMaxXGroups) G = kMaxXGroups;
  if (G > nstates) G = nstates;
  return G < 1 ? 1 : G;
}

// Block (tile, group): fixed y, z in [z0, z0+nl), states of the chunk in
// group blockIdx.y.  Inverse transform along x; accumulates occ |psi|^2 in
// registers and adds them to the group's partial density.
template <int NX>
__global__ __launch_bounds__(kXBlock) void xline_density_kernel(
    const double2* __restrict__ T, double* __restrict__ rho_part,
    const double* __restrict__ occ, int64_t state0, int nstates,
    double inv_vol, BoxGeom geo, FftLen Lx, int LZ, int store) {
  extern __shared__ double2 sm[];
  const int nx = NX > 0 ? NX : geo.nx;
  const int ny = geo.ny, nz = geo.nz, P = ny * nz;
  const int nzt = (nz + LZ - 1) / LZ;
  const int y = blockIdx.x / nzt;
  const int z0 = (blockIdx.x % nzt) * LZ;
  const int nl = min(LZ, nz - z0);
  const int G = gridDim.y, gi = blockIdx.y;
  const int per = (nstates + G - 1) / G;
  const int cb = gi * per, ce = min(nstates, cb + per);
  const int nxl = nx * LZ;
  double* part = rho_part + (int64_t)gi * geo.nx * P;
  if (cb >= ce) {  // uniform over the block: an empty group
    if (store)
      for (int e = threadIdx.x; e < nxl; e += kXBlock) {
        const int x = e / LZ, l = e % LZ;
        if (l < nl) part[((int64_t)x * ny + y) * nz + z0 + l] = 0.0;
      }
    return;
  }
  int goff[kXEpt];
  xline_slots(goff, geo, nxl, LZ, nl);
  const double2* src0 = T + (int64_t)y * nz + z0;
  const int64_t sstride = (int64_t)geo.nxa * geo.ps;
  double2 pre[kXEpt];
  xline_load(pre, goff, src0 + (int64_t)cb * sstride);
  double acc[kXEpt];
#pragma unroll
  for (int k = 0; k < kXEpt; ++k) acc[k] = 0.0;
  for (int c = cb; c < ce; ++c) {
#pragma unroll
    for (int k = 0; k < kXEpt; ++k) {
      const int e = threadIdx.x + k * kXBlock;
      if (e < nxl) sm[e] = pre[k];
    }
    if (c + 1 < ce) xline_load(pre, goff, src0 + (int64_t)(c + 1) * sstride);
    __syncthreads();
    fft_lines<NX, true, kXBlock, kXCap, true>(sm, nl, 1, LZ, nullptr, Lx);
    const double f = occ[state0 + c];
#pragma unroll
    for (int k = 0; k < kXEpt; ++k) {
      const int e = threadIdx.x + k * kXBlock;
      if (e < nxl && (e % LZ) < nl) {
        const double2 v = sm[e];
        acc[k] += f * (v.x * v.x + v.y * v.y);
      }
    }
    __syncthreads();
  }
  // the group's partial: stored by the first sweep chunk, added to after
#pragma unroll
  for (int k = 0; k < kXEpt; ++k) {
    const int e = threadIdx.x + k * kXBlock;
    const int x = e / LZ, l = e % LZ;
    if (e < nxl && l < nl) {
      const int64_t idx = ((int64_t)x * ny + y) * nz + z0 + l;
      part[idx] = store ? acc[k] * inv_vol : part[idx] + acc[k] * inv_vol;
    }
  }
}

int launch_xline_density(const double2* T, double* rho_part, int groups,
                         const double* occ, int64_t state0, int nstates,
                         double inv_vol, const BoxGeom& geo, const FftLen& Lx,
                         bool store, hipStream_t s) {
  if (geo.nx > kXCap) {
    set_error("x grid size above 2048 is not supported");
    return JRX_ERR_UNSUPPORTED;
  }
  const int LZ = xlines_per_block(geo.nx, geo.nz);
  const int nzt = (geo.nz + LZ - 1) / LZ;
  size_t shm = (size_t)geo.nx * LZ * sizeof(double2) + (size_t)geo.nxa * 4 + 16;
  dim3 grid(geo.ny * nzt, groups), block(kXBlock);
#define JRX_XD(NN)                                                          \
  case NN:                                                                  \
    hipLaunchKernelGGL((xline_density_kernel<NN>), grid, block, shm, s, T,  \
                       rho_part, occ, state0, nstates, inv_vol, geo, Lx,    \
                       LZ, store ? 1 : 0);                                  \
    break;
  switch (geo.nx) {
    JRX_CT_SIZES(JRX_XD)
    default:
      hipLaunchKernelGGL((xline_density_kernel<0>), grid, block, shm, s, T,
                         rho_part, occ, state0, nstates, inv_vol, geo, Lx,
                         LZ, store ? 1 : 0);
  }
#undef JRX_XD
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

// T[c] <- FFT_x(v * IFFT_x(T[c])) on the active planes, for every state.
template <int NX>
__global__ __launch_bounds__(kXBlock) void xline_vpsi_kernel(
    double2* __restrict__ T, const double* __restrict__ v, int nstates,
    BoxGeom geo, FftLen Lx, int LZ) {
  extern __shared__ double2 sm[];
  const int nx = NX > 0 ? NX : geo.nx;
  const int ny = geo.ny, nz = geo.nz;
  const int nzt = (nz + LZ - 1) / LZ;
  const int y = blockIdx.x / nzt;
  const int z0 = (blockIdx.x % nzt) * LZ;
  const int nl = min(LZ, nz - z0);
  const int G = gridDim.y, gi = blockIdx.y;
  const int per = (nstates + G - 1) / G;
  const int cb = gi * per, ce = min(nstates, cb + per);
  if (cb >= ce) return;  // uniform over the block
  const int nxl = nx * LZ;
  int goff[kXEpt];
  xline_slots(goff, geo, nxl, LZ, nl);
  double vr[kXEpt];
#pragma unroll
  for (int k = 0; k < kXEpt; ++k) {
    const int e = threadIdx.x + k * kXBlock;
    const int x = e / LZ, l = e % LZ;
    vr[k] = (e < nxl && l < nl) ? v[((int64_t)x * ny + y) * nz + z0 + l] : 0.0;
  }
  double2* buf0 = T + (int64_t)y * nz + z0;
  const int64_t sstride = (int64_t)geo.nxa * geo.ps;
  double2 pre[kXEpt];
  xline_load(pre, goff, buf0 + (int64_t)cb * sstride);
  for (int c = cb; c < ce; ++c) {
#pragma unroll
    for (int k = 0; k < kXEpt; ++k) {
      const int e = threadIdx.x + k * kXBlock;
      if (e < nxl) sm[e] = pre[k];
    }
    if (c + 1 < ce) xline_load(pre, goff, buf0 + (int64_t)(c + 1) * sstride);
    __syncthreads();
    fft_lines<NX, true, kXBlock, kXCap, true>(sm, nl, 1, LZ, nullptr, Lx);
#pragma unroll
    for (int k = 0; k < kXEpt; ++k) {
      const int e = threadIdx.x + k * kXBlock;
      if (e < nxl) sm[e] = cscale(sm[e], vr[k]);
    }
    __syncthreads();
    fft_lines<NX, false, kXBlock, kXCap, true>(sm, nl, 1, LZ, nullptr, Lx);
    double2* buf = buf0 + (int64_t)c * sstride;
#pragma unroll
    for (int k = 0; k < kXEpt; ++k)
      if (goff[k] >= 0) nt_store(buf + goff[k], sm[threadIdx.x + k * kXBlock]);
    __syncthreads();
  }
}

int launch_xline_vpsi(double2* T, const double* v, int groups, int nstates,
                      const BoxGeom& geo, const FftLen& Lx, hipStream_t s) {
  if (geo.nx > kXCap) {
    set_error("x grid size above 2048 is not supported");
    return JRX_ERR_UNSUPPORTED;
  }
  const int LZ = xlines_per_block(geo.nx, geo.nz);
  const int nzt = (geo.nz + LZ - 1) / LZ;
  size_t shm = (size_t)geo.nx * LZ * sizeof(double2) + (size_t)geo.nxa * 4 + 16;
  dim3 grid(geo.ny * nzt, groups), block(kXBlock);
#define JRX_XV(NN)                                                          \
  case NN:                                                                  \
    hipLaunchKernelGGL((xline_vpsi_kernel<NN>), grid, block, shm, s, T, v,  \
                       nstates, geo, Lx, LZ);                               \
    break;
  switch (geo.nx) {
    JRX_CT_SIZES(JRX_XV)
    default:
      hipLaunchKernelGGL((xline_vpsi_kernel<0>), grid, block, shm, s, T, v,
                         nstates, geo, Lx, LZ);
  }
#undef JRX_XV
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

// out[i] = sum_g part[g][i], four sums in flight (groups g = 4m + u go to
// sum u), added in a fixed order.
__global__ void reduce_groups_kernel(const double* __restrict__ part,
                                     double* __restrict__ out, int G,
                                     int64_t N) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int g = 0;
  for (; g + 3 < G; g += 4) {
    a0 += part[(int64_t)g * N + i];
    a1 += part[(int64_t)(g + 1) * N + i];
    a2 += part[(int64_t)(g + 2) * N + i];
    a3 += part[(int64_t)(g + 3) * N + i];
  }
  for (; g < G; ++g) a0 += part[(int64_t)g * N + i];
  out[i] = (a0 + a1) + (a2 + a3);
}

int launch_reduce_groups(const double* part, double* out, int G, int64_t N,
                         hipStream_t s) {
  hipLaunchKernelGGL(reduce_groups_kernel, dim3(ceil_div(N, 256)), dim3(256),
                     0, s, part, out, G, N);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

// Forward transform of one plane (y, then z on the mask rows), gather at the
// mask points and combine with the kinetic term.
template <int NY, int NZ, int CAP>
__global__ __launch_bounds__(kPlaneBlock) void plane_fft_gather_kernel(
    const double2* __restrict__ T, const double2* __restrict__ q,
    double2* __restrict__ qbar, double* __restrict__ kin_partial,
    int64_t state0, BoxGeom geo, FftLen Ly, FftLen Lz, GatherParams gp) {
  extern __shared__ double2 sm[];
  const int nx = geo.nx;
  const int ny = NY > 0 ? NY : geo.ny, nz = NZ > 0 ? NZ : geo.nz;
  const int P = ny * nz;
  const int xi = blockIdx.x;
  const int c = blockIdx.y;
  const int64_t st = state0 + c;
  const double2* src = T + ((int64_t)c * geo.nxa + xi) * geo.ps;
  for (int i = threadIdx.x; i < P; i += kPlaneBlock) sm[i] = src[i];
  const int r0 = geo.row_off[xi];
  const int nr = geo.row_off[xi + 1] - r0;
  const short* rows = geo.rows + r0;
  __syncthreads();
  fft_lines<NY, false, kPlaneBlock, CAP, true>(sm, nz, 1, nz, nullptr, Ly);
  fft_lines<NZ, false, kPlaneBlock, CAP, false>(sm, nr, nz, 1, rows, Lz);
  const int g0 = geo.plane_g0[xi], g1 = geo.plane_g0[xi + 1];
  const int k = (int)((st / gp.nb) % gp.nk);
  const double f = gp.occ[st];
  const double kx = gp.kpts[3 * k], ky = gp.kpts[3 * k + 1],
               kz = gp.kpts[3 * k + 2];
  const int mx = fftfreq_index(geo.plane_x[xi], nx);
  // G + k = mx b1 + my b2 + mz b3 + k
  const double ax = mx * gp.B[0] + kx, ay = mx * gp.B[1] + ky,
               az = mx * gp.B[2] + kz;
  const double2* qs = q + st * geo.ng;
  double2* qb = qbar + st * geo.ng;
  const bool be = gp.pot_partial != nullptr;
  double kin = 0.0, pe = 0.0;
  for (int g = g0 + threadIdx.x; g < g1; g += kPlaneBlock) {
    const int yz = geo.g_yz[g];
    const int iy = yz / nz, iz = yz - iy * nz;
    const int my = fftfreq_index(iy, ny), mz = fftfreq_index(iz, nz);
    const double gx = ax + my * gp.B[3] + mz * gp.B[6];
    const double gy = ay + my * gp.B[4] + mz * gp.B[7];
    const double gz = az + my * gp.B[5] + mz * gp.B[8];
    const double gk2 = gx * gx + gy * gy + gz * gz;
    const double2 cv = qs[g];
    const double2 hv = sm[yz];
    qb[g] = make_double2(f * (gp.two_over_vol * hv.x + gk2 * cv.x),
                         f * (gp.two_over_vol * hv.y + gk2 * cv.y));
    kin += gk2 * (cv.x * cv.x + cv.y * cv.y);
    pe += cv.x * hv.x + cv.y * hv.y;
  }
  __syncthreads();  // the plane is dead: reuse its LDS for the reduction
  double* red = reinterpret_cast<double*>(sm);  // [2][kPlaneBlock]
  red[threadIdx.x] = kin;
  if (be) red[kPlaneBlock + threadIdx.x] = pe;
  __syncthreads();
  if (threadIdx.x < 64) {
    const double tot = wave0_sum(red);
    const double tpe = be ? wave0_sum(red + kPlaneBlock) : 0.0;
    if (threadIdx.x == 0) {
      kin_partial[st * geo.nxa + xi] = tot;
      if (be) gp.pot_partial[st * geo.nxa + xi] = tpe;
    }
  }
}

int launch_plane_fft_gather(const double2* T, const double2* q, double2* qbar,
                            double* kin_partial, int64_t state0, int nstates,
                            const BoxGeom& geo, const FftLen& Ly,
                            const FftLen& Lz, const GatherParams& gp,
                            hipStream_t s) {
  const int P = geo.ny * geo.nz;
  size_t shm = (size_t)P * sizeof(double2);
  if (shm < 2 * kPlaneBlock * sizeof(double))  // the final reductions
    shm = 2 * kPlaneBlock * sizeof(double);
  dim3 grid(geo.nxa, nstates), block(kPlaneBlock);
  const int cube = geo.ny == geo.nz ? geo.ny : 0;
#define JRX_PG(NN)                                                          \
  case NN:                                                                  \
    hipLaunchKernelGGL((plane_fft_gather_kernel<NN, NN, NN * NN>), grid,    \
                       block, shm, s, T, q, qbar, kin_partial, state0, geo, \
                       Ly, Lz, gp);                                         \
    break;
  switch (cube) {
    JRX_CT_PLANE_SIZES(JRX_PG)
    default:
      hipLaunchKernelGGL((plane_fft_gather_kernel<0, 0, kPlaneCap>), grid,
                         block, shm, s, T, q, qbar, kin_partial, state0, geo,
                         Ly, Lz, gp);
  }
#undef JRX_PG
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}


// ----------------------------------------- register kernels for 64-point --
// Specialisations for nx = ny = nz = 64 built on fft64_8x8: no zero fills,
// no block barriers inside the transforms, loads and stores straight from
// registers.  Plane rows are padded to 66 entries so that the column
// accesses of 16-lane groups hit distinct LDS banks.
// Streaming access (nt_load / nt_store) to the plane buffer (written once and read once per
// sweep: gigabytes, far beyond the caches) and to the coefficients the
// inverse plane pass scatters into LDS carries the nontemporal hint
// (measured at 64^3: x-line passes 5-13% and the inverse plane pass 20%
// faster).  Loads of data that another wave of the same kernel re-reads
// (the QR Gram inputs) keep the default policy, which measured faster.

constexpr int kPL = 66;
constexpr int kExLine = 72;  // padded exchange slots per x-line

__global__ __launch_bounds__(kPlaneBlock) void plane_ifft64_kernel(
    const double2* __restrict__ q, double2* __restrict__ T, int64_t state0,
    BoxGeom geo, const double2* __restrict__ tw64) {
  extern __shared__ double2 sm[];
  double2* pl = sm;              // [64][kPL]
  double2* stw = sm + 64 * kPL;  // [64]
  const int tid = threadIdx.x, j = tid & 7, line = tid >> 3;
  const int xi = blockIdx.x, c = blockIdx.y;
  const int64_t st = state0 + c;
  for (int i = tid; i < 64 * kPL; i += kPlaneBlock) pl[i] = czero();
  if (tid < 64) stw[tid] = tw64[tid];
  __syncthreads();
  const int g0 = geo.plane_g0[xi], g1 = geo.plane_g0[xi + 1];
  const double2* qs = q + st * geo.ng;
  for (int g = g0 + tid; g < g1; g += kPlaneBlock) {
    const int yz = geo.g_yz[g];
    pl[(yz >> 6) * kPL + (yz & 63)] = nt_load(qs + g);
  }
  __syncthreads();
  const int r0 = geo.row_off[xi], nr = geo.row_off[xi + 1] - r0;
  if (line < nr) {  // z transform of the mask rows
    double2* base = pl + geo.rows[r0 + line] * kPL;
    double2 v[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = base[j + 8 * m];
    fft64_8x8<true>(v, base, 1, j, stw);
#pragma unroll
    for (int t = 0; t < 8; ++t) base[j + 8 * t] = v[t];
  }
  __syncthreads();
  {  // y transform of column z = line, straight to the plane buffer T
    double2* base = pl + line;
    double2 v[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = base[(j + 8 * m) * kPL];
    fft64_8x8<true>(v, base, kPL, j, stw);
    double2* dst = T + ((int64_t)c * geo.nxa + xi) * geo.ps + line;
#pragma unroll
    for (int t = 0; t < 8; ++t) nt_store(dst + (j + 8 * t) * 64, v[t]);
  }
}

// Persistent: block b handles items b, b + gridDim.x, ... (item = state *
// nxa + active plane) and loads the next item's columns from the plane
// buffer while it transforms and gathers the current one (only two blocks
// fit a CU's LDS, so in-block prefetch is what hides the load latency;
// measured 7% faster than one block per item).
// Two blocks fit a CU's LDS (72.7 KB each); both are resident only at 4
// waves per SIMD, i.e. at most 128 VGPRs, hence the second bound.
template <bool BE>
__global__ __launch_bounds__(kPlaneBlock, 4) void plane_fft_gather64_kernel(
    const double2* __restrict__ T, const double2* __restrict__ q,
    double2* __restrict__ qbar, double* __restrict__ kin_partial,
    int64_t state0, int nstates, BoxGeom geo,
    const double2* __restrict__ tw64, GatherParams gp) {
  extern __shared__ double2 sm[];
  double2* pl = sm;
  double2* stw = sm + 64 * kPL;
  // [512] kinetic, then (BE) [512] band-energy partials
  double* red = reinterpret_cast<double*>(sm + 64 * kPL + 64);
  const int tid = threadIdx.x, j = tid & 7, line = tid >> 3;
  const int nitems = geo.nxa * nstates;
  if (tid < 64) stw[tid] = tw64[tid];
  __syncthreads();
  int item = blockIdx.x;
  double2 nv[8];
  if (item < nitems) {
    const int xi0 = item % geo.nxa, c0 = item / geo.nxa;
    const double2* src = T + ((int64_t)c0 * geo.nxa + xi0) * geo.ps + line;
#pragma unroll
    for (int m = 0; m < 8; ++m) nv[m] = nt_load(src + (j + 8 * m) * 64);
  }
  for (; item < nitems; item += gridDim.x) {
    const int xi = item % geo.nxa, c = item / geo.nxa;
    const int64_t st = state0 + c;
    double2 v[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = nv[m];
    const int nxt = item + gridDim.x;
    if (nxt < nitems) {  // prefetch the next item's columns
      const int xn = nxt % geo.nxa, cn = nxt / geo.nxa;
      const double2* src = T + ((int64_t)cn * geo.nxa + xn) * geo.ps + line;
#pragma unroll
      for (int m = 0; m < 8; ++m) nv[m] = nt_load(src + (j + 8 * m) * 64);
    }
    {  // y transform of column z = line
      double2* base = pl + line;
      fft64_8x8<false>(v, base, kPL, j, stw);
#pragma unroll
      for (int t = 0; t < 8; ++t) base[(j + 8 * t) * kPL] = v[t];
    }
    __syncthreads();
    const int r0 = geo.row_off[xi], nr = geo.row_off[xi + 1] - r0;
    if (line < nr) {  // z transform of the mask rows
      double2* base = pl + geo.rows[r0 + line] * kPL;
      double2 w[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) w[m] = base[j + 8 * m];
      fft64_8x8<false>(w, base, 1, j, stw);
#pragma unroll
      for (int t = 0; t < 8; ++t) base[j + 8 * t] = w[t];
    }
    __syncthreads();
    const int g0 = geo.plane_g0[xi], g1 = geo.plane_g0[xi + 1];
    const int k = (int)((st / gp.nb) % gp.nk);
    const double f = gp.occ[st];
    const int mx = fftfreq_index(geo.plane_x[xi], 64);
    const double ax = mx * gp.B[0] + gp.kpts[3 * k];
    const double ay = mx * gp.B[1] + gp.kpts[3 * k + 1];
    const double az = mx * gp.B[2] + gp.kpts[3 * k + 2];
    const double2* qs = q + st * geo.ng;
    double2* qb = qbar + st * geo.ng;
    double kin = 0.0, pe = 0.0;
    for (int g = g0 + tid; g < g1; g += kPlaneBlock) {
      const int yz = geo.g_yz[g];
      const int iy = yz >> 6, iz = yz & 63;
      const int my = fftfreq_index(iy, 64), mz = fftfreq_index(iz, 64);
      const double gx = ax + my * gp.B[3] + mz * gp.B[6];
      const double gy = ay + my * gp.B[4] + mz * gp.B[7];
      const double gz = az + my * gp.B[5] + mz * gp.B[8];
      const double gk2 = gx * gx + gy * gy + gz * gz;
      const double2 cv = qs[g];
      const double2 hv = pl[iy * kPL + iz];
      qb[g] = make_double2(f * (gp.two_over_vol * hv.x + gk2 * cv.x),
                           f * (gp.two_over_vol * hv.y + gk2 * cv.y));
      kin += gk2 * (cv.x * cv.x + cv.y * cv.y);
      if constexpr (BE) pe += cv.x * hv.x + cv.y * hv.y;
    }
    // red lies outside pl; wave 0 reads it before its next barrier, and no
    // thread stores to it again before passing that barrier
    red[tid] = kin;
    if constexpr (BE) red[kPlaneBlock + tid] = pe;
    __syncthreads();  // pl is free for the next item, red is complete
    if (tid < 64) {
      const double tot = wave0_sum(red);
      if (tid == 0) kin_partial[st * geo.nxa + xi] = tot;
      if constexpr (BE) {
        const double tpe = wave0_sum(red + kPlaneBlock);
        if (tid == 0) gp.pot_partial[st * geo.nxa + xi] = tpe;
      }
    }
  }
}

// x-lines: block = 32 lines (fixed y, z0..z0+31), 8 lanes per line.
__global__ __launch_bounds__(kXBlock) void xline_density64_kernel(
    const double2* __restrict__ T, double* __restrict__ rho_part,
    const double* __restrict__ occ, int64_t state0, int nstates,
    double inv_vol, BoxGeom geo, const double2* __restrict__ tw64,
    int store) {
  __shared__ double2 ex[(kXBlock / 8) * kExLine];
  __shared__ double2 stw[64];
  const int tid = threadIdx.x, j = tid & 7, line = tid >> 3;
  const int nz = geo.nz, ny = geo.ny, P = ny * nz;
  const int nzt = (nz + 31) / 32;
  const int y = blockIdx.x / nzt;
  const int z = (blockIdx.x % nzt) * 32 + line;
  const bool zok = z < nz;
  const int G = gridDim.y, gi = blockIdx.y;
  const int per = (nstates + G - 1) / G;
  const int cb = gi * per, ce = min(nstates, cb + per);
  double* part = rho_part + (int64_t)gi * 64 * P + (int64_t)y * nz + z;
  if (tid < 64) stw[tid] = tw64[tid];
  __syncthreads();
  if (cb >= ce) {  // uniform over the block: an empty group
    if (store && zok)
#pragma unroll
      for (int t = 0; t < 8; ++t) part[(int64_t)(j + 8 * t) * P] = 0.0;
    return;
  }
  int xo[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int xi = geo.xmap[j + 8 * m];
    xo[m] = (xi >= 0 && zok) ? xi * geo.ps : -1;
  }
  double2* base = ex + line * kExLine;
  double acc[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) acc[t] = 0.0;
  const double2* src0 = T + (int64_t)y * nz + z;
  double2 nxt[8];
#pragma unroll
  for (int m = 0; m < 8; ++m)
    nxt[m] = xo[m] >= 0 ? nt_load(src0 + (int64_t)cb * geo.nxa * geo.ps + xo[m])
                        : czero();
  for (int c = cb; c < ce; ++c) {
    double2 v[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = nxt[m];
    if (c + 1 < ce) {
      const double2* src = src0 + (int64_t)(c + 1) * geo.nxa * geo.ps;
#pragma unroll
      for (int m = 0; m < 8; ++m)
        nxt[m] = xo[m] >= 0 ? nt_load(src + xo[m]) : czero();
    }
    fft64_8x8<true>(v, base, 1, j, stw);
    const double f = occ[state0 + c];
#pragma unroll
    for (int t = 0; t < 8; ++t) acc[t] += f * (v[t].x * v[t].x + v[t].y * v[t].y);
  }
  if (!zok) return;
  // the group's partial: stored by the first sweep chunk, added to after
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    double* dst = part + (int64_t)(j + 8 * t) * P;
    *dst = store ? acc[t] * inv_vol : *dst + acc[t] * inv_vol;
  }
}

__global__ __launch_bounds__(kXBlock) void xline_vpsi64_kernel(
    double2* __restrict__ T, const double* __restrict__ v, int nstates,
    BoxGeom geo, const double2* __restrict__ tw64) {
  __shared__ double2 ex[(kXBlock / 8) * kExLine];
  __shared__ double2 stw[64];
  const int tid = threadIdx.x, j = tid & 7, line = tid >> 3;
  const int nz = geo.nz, ny = geo.ny, P = ny * nz;
  const int nzt = (nz + 31) / 32;
  const int y = blockIdx.x / nzt;
  const int z = (blockIdx.x % nzt) * 32 + line;
  const bool zok = z < nz;
  const int G = gridDim.y, gi = blockIdx.y;
  const int per = (nstates + G - 1) / G;
  const int cb = gi * per, ce = min(nstates, cb + per);
  if (tid < 64) stw[tid] = tw64[tid];
  __syncthreads();
  if (cb >= ce) return;
  int xo[8];
  double vr[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int xi = geo.xmap[j + 8 * m];
    xo[m] = (xi >= 0 && zok) ? xi * geo.ps : -1;
    vr[m] = zok ? v[(int64_t)(j + 8 * m) * P + (int64_t)y * nz + z] : 0.0;
  }
  double2* base = ex + line * kExLine;
  double2* buf0 = T + (int64_t)y * nz + z;
  double2 nxt[8];
#pragma unroll
  for (int m = 0; m < 8; ++m)
    nxt[m] = xo[m] >= 0 ? nt_load(buf0 + (int64_t)cb * geo.nxa * geo.ps + xo[m])
                        : czero();
  for (int c = cb; c < ce; ++c) {
    double2 w[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) w[m] = nxt[m];
    double2* buf = buf0 + (int64_t)c * geo.nxa * geo.ps;
    if (c + 1 < ce) {
      const double2* src = buf + (int64_t)geo.nxa * geo.ps;
#pragma unroll
      for (int m = 0; m < 8; ++m)
        nxt[m] = xo[m] >= 0 ? nt_load(src + xo[m]) : czero();
    }
    fft64_8x8<true>(w, base, 1, j, stw);
#pragma unroll
    for (int t = 0; t < 8; ++t) w[t] = cscale(w[t], vr[t]);
    fft64_8x8<false>(w, base, 1, j, stw);
#pragma unroll
    for (int m = 0; m < 8; ++m)
      if (xo[m] >= 0) nt_store(buf + xo[m], w[m]);
  }
}

bool use_reg64(const BoxGeom& geo) {
  return geo.nx == 64 && geo.ny == 64 && geo.nz == 64;
}

int launch_plane_ifft64(const double2* q, double2* T, int64_t state0,
                        int nstates, const BoxGeom& geo, const double2* tw64,
                        hipStream_t s) {
  size_t shm = (size_t)(64 * kPL + 64) * sizeof(double2);
  hipLaunchKernelGGL(plane_ifft64_kernel, dim3(geo.nxa, nstates),
                     dim3(kPlaneBlock), shm, s, q, T, state0, geo, tw64);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int launch_plane_fft_gather64(const double2* T, const double2* q,
                              double2* qbar, double* kin_partial,
                              int64_t state0, int nstates, const BoxGeom& geo,
                              const double2* tw64, const GatherParams& gp,
                              hipStream_t s) {
  const bool be = gp.pot_partial != nullptr;
  size_t shm = (size_t)(64 * kPL + 64) * sizeof(double2) +
               (be ? 2 : 1) * kPlaneBlock * sizeof(double);
  const int64_t nitems = (int64_t)geo.nxa * nstates;
  const int64_t nblk = nitems < 1024 ? nitems : 1024;
  if (be)
    hipLaunchKernelGGL(plane_fft_gather64_kernel<true>, dim3((unsigned)nblk),
                       dim3(kPlaneBlock), shm, s, T, q, qbar, kin_partial,
                       state0, nstates, geo, tw64, gp);
  else
    hipLaunchKernelGGL(plane_fft_gather64_kernel<false>, dim3((unsigned)nblk),
                       dim3(kPlaneBlock), shm, s, T, q, qbar, kin_partial,
                       state0, nstates, geo, tw64, gp);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int launch_xline_density64(const double2* T, double* rho_part, int groups,
                           const double* occ, int64_t state0, int nstates,
                           double inv_vol, const BoxGeom& geo,
                           const double2* tw64, bool store, hipStream_t s) {
  const int nzt = (geo.nz + 31) / 32;
  hipLaunchKernelGGL(xline_density64_kernel, dim3(geo.ny * nzt, groups),
                     dim3(kXBlock), 0, s, T, rho_part, occ, state0, nstates,
                     inv_vol, geo, tw64, store ? 1 : 0);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int launch_xline_vpsi64(double2* T, const double* v, int groups, int nstates,
                        const BoxGeom& geo, const double2* tw64,
                        hipStream_t s) {
  const int nzt = (geo.nz + 31) / 32;
  hipLaunchKernelGGL(xline_vpsi64_kernel, dim3(geo.ny * nzt, groups),
                     dim3(kXBlock), 0, s, T, v, nstates, geo, tw64);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

}  // namespace jrx
