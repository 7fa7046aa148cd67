// Row-buffer path of the fused evaluation, for cubic boxes of 64 or 128
// points per side whose mask spans at most 64 x-planes.
//
// A state's 3D transform runs through a row buffer R[state][z][row], where
// the rows are the (x, y) lines along z that hold mask points (numbered in
// C order, so the mask points of a row are a contiguous range of g):
//   row_ifft        sphere -> inverse z transform of every row -> R
//   zplane_density  per z-plane: inverse y transform of the active x
//                   columns, inverse x transform of every y line,
//                   occupation-weighted |psi|^2 accumulated per state group
//   zplane_vpsi     the same inverse transforms, times v(r), forward x and
//                   y transforms back to the rows of R (in place)
//   row_fft_gather  forward z transform of every row, gather at the mask,
//                   qbar and the kinetic / band-energy partials
// R holds N * rows points per state (0.81 MB at 64^3 with 793 rows),
// against N * N * active planes (2.0 MB) for a buffer of y,z-transformed
// planes, and it is crossed four times per evaluation.  Densities and
// potentials on this path are z-major: [z][y][x].
#include "fft_kernels.h"

namespace jrx {

// LDS layouts, chosen by measuring the bank-conflict counter on gfx950 for
// the exact exchange patterns of these kernels (conflict cycles per 16-byte
// access of one wave): line exchanges in rows of stride 72 cost 0 (stride
// 66: 9 on reads); column exchanges through a stride-72 plane cost 56 on
// writes unless the column position is XOR-swizzled by the row
// (ZPlane::at: 0 on writes, 3 on reads); transposed tiles of stride
// RPB + 1 cost 0 on writes and 4 on reads.
// LDS stride (elements) of one row of a z-plane: 64 exchange slots + pad.
constexpr int kZRow = 72;
// row-pass staging rows (N + kRowPad) and transposed tiles (RPB + 1)
constexpr int kRowPad = 8;

template <int N>
struct RowCfg;  // rows per block of the row passes (8 lanes per row)
template <>
struct RowCfg<64> {
  static constexpr int RPB = 64;
};
template <>
struct RowCfg<128> {
  static constexpr int RPB = 32;
};

// Launch shape of the 64-point z-plane passes (compile-time tuning knobs):
// threads per block, the waves per SIMD the register budget is set for (0:
// the compiler's choice), branch-free row gathers / scatters, and register
// prefetch of the next state's column inputs.  Measured at 64^3 (64 k, 32
// bands): 256 threads with prefetch and branches 2.07 ms (v psi pass);
// 512 threads at 4 waves per SIMD 2.61 ms; branch-free forms 3.05-3.23 ms.
#ifndef JRX_ZBLOCK64
#define JRX_ZBLOCK64 256
#endif
#ifndef JRX_ZWPE64
#define JRX_ZWPE64 0
#endif
#ifndef JRX_ZBRANCHFREE
#define JRX_ZBRANCHFREE 0
#endif
#ifndef JRX_ZPREFETCH64
#define JRX_ZPREFETCH64 1
#endif

template <int N>
struct ZCfg;  // threads of the z-plane passes (8 lanes per line) and the
              // waves per SIMD the register budget is set for
template <>
struct ZCfg<64> {
  static constexpr int BLOCK = JRX_ZBLOCK64, WPE = JRX_ZWPE64;
  static constexpr bool PF = JRX_ZPREFETCH64 != 0;
};
template <>
struct ZCfg<128> {
  static constexpr int BLOCK = 512, WPE = 2;  // LDS: one block per CU
  static constexpr bool PF = false;
};

template <int N>
constexpr int row_lds_elems() {
  constexpr int RPB = RowCfg<N>::RPB;
  constexpr int a = RPB * (N + kRowPad), b = N * (RPB + 1);
  return (a > b ? a : b) + N;  // + twiddles
}

// ------------------------------------------------------------ row_ifft --
// Block = RPB rows of one state.  The mask points of the rows are scattered
// into zeroed LDS rows, each row is transformed along z by its eight lanes,
// and the result is stored transposed, R[c][z][r0 .. r0 + RPB).
template <int N>
__global__ __launch_bounds__(RowCfg<N>::RPB * 8) void row_ifft_kernel(
    const double2* __restrict__ q, double2* __restrict__ R, int64_t state0,
    BoxGeom geo, const double2* __restrict__ tw) {
  constexpr int RPB = RowCfg<N>::RPB, NT = RPB * 8, V = N / 8;
  constexpr int RS = N + kRowPad, TS = RPB + 1;
  extern __shared__ double2 sm[];
  double2* stw = sm + (row_lds_elems<N>() - N);
  const int tid = threadIdx.x, j = tid & 7, rl = tid >> 3;
  const int r0 = blockIdx.x * RPB;
  const int nr = min(RPB, geo.nrows - r0);
  const int c = blockIdx.y;
  const int64_t st = state0 + c;
  for (int i = tid; i < RPB * RS; i += NT) sm[i] = czero();
  for (int i = tid; i < N; i += NT) stw[i] = tw[i];
  __syncthreads();
  const int g0 = geo.row_g0[r0], g1 = geo.row_g0[r0 + nr];
  const double2* qs = q + st * geo.ng;
  for (int g = g0 + tid; g < g1; g += NT)
    sm[(geo.g_row[g] - r0) * RS + (geo.g_yz[g] & (N - 1))] = nt_load(qs + g);
  __syncthreads();
  double2 v[V];
  double2* row = sm + rl * RS;
#pragma unroll
  for (int m = 0; m < V; ++m) v[m] = row[j + 8 * m];
  RegFft<N>::template run<true>(v, row, 1, j, stw);
  __syncthreads();  // every row is in registers: reuse LDS as [z][row]
#pragma unroll
  for (int t = 0; t < V; ++t) sm[(j + 8 * t) * TS + rl] = v[t];
  __syncthreads();
  double2* dst = R + (int64_t)c * N * geo.rstride + r0;
  const int r = tid % RPB;
  if (r < nr)
    for (int z = tid / RPB; z < N; z += NT / RPB)
      nt_store(dst + (int64_t)z * geo.rstride + r, sm[z * TS + r]);
}

// ------------------------------------------------------- row_fft_gather --
// Block = RPB rows of one state: R[c][z][rows] -> LDS [z][row], forward z
// transform of each row, then qbar at the block's mask points and the
// block's partials of sum |G+k|^2 |q|^2 (and, with BE, of Re conj(q) h).
template <int N, bool BE>
__global__ __launch_bounds__(RowCfg<N>::RPB * 8) void row_fft_gather_kernel(
    const double2* __restrict__ R, const double2* __restrict__ q,
    double2* __restrict__ qbar, double* __restrict__ kin_partial,
    int64_t state0, BoxGeom geo, const double2* __restrict__ tw,
    GatherParams gp) {
  constexpr int RPB = RowCfg<N>::RPB, NT = RPB * 8, V = N / 8;
  constexpr int TS = RPB + 1;
  extern __shared__ double2 sm[];
  double2* stw = sm + N * TS;
  double* red = reinterpret_cast<double*>(stw + N);  // [2][NT]
  const int tid = threadIdx.x, j = tid & 7, rl = tid >> 3;
  const int r0 = blockIdx.x * RPB;
  const int nr = min(RPB, geo.nrows - r0);
  const int c = blockIdx.y;
  const int64_t st = state0 + c;
  const double2* src = R + (int64_t)c * N * geo.rstride + r0;
  {
    const int r = tid % RPB;
    for (int z = tid / RPB; z < N; z += NT / RPB)
      sm[z * TS + r] =
          r < nr ? nt_load(src + (int64_t)z * geo.rstride + r) : czero();
  }
  for (int i = tid; i < N; i += NT) stw[i] = tw[i];
  __syncthreads();
  {
    double2 v[V];
    double2* col = sm + rl;
#pragma unroll
    for (int m = 0; m < V; ++m) v[m] = col[(j + 8 * m) * TS];
    RegFft<N>::template run<false>(v, col, TS, j, stw);
#pragma unroll
    for (int t = 0; t < V; ++t) col[(j + 8 * t) * TS] = v[t];
  }
  __syncthreads();
  const int g0 = geo.row_g0[r0], g1 = geo.row_g0[r0 + nr];
  const int k = (int)((st / gp.nb) % gp.nk);
  const double f = gp.occ[st];
  const double kx = gp.kpts[3 * k], ky = gp.kpts[3 * k + 1],
               kz = gp.kpts[3 * k + 2];
  const double2* qs = q + st * geo.ng;
  double2* qb = qbar + st * geo.ng;
  double kin = 0.0, pe = 0.0;
  for (int g = g0 + tid; g < g1; g += NT) {
    const int mx = gp.g_m[3 * g], my = gp.g_m[3 * g + 1],
              mz = gp.g_m[3 * g + 2];
    const double gx = mx * gp.B[0] + my * gp.B[3] + mz * gp.B[6] + kx;
    const double gy = mx * gp.B[1] + my * gp.B[4] + mz * gp.B[7] + ky;
    const double gz = mx * gp.B[2] + my * gp.B[5] + mz * gp.B[8] + kz;
    const double gk2 = gx * gx + gy * gy + gz * gz;
    const double2 cv = nt_load(qs + g);
    const double2 hv =
        sm[(geo.g_yz[g] & (N - 1)) * TS + (geo.g_row[g] - r0)];
    nt_store(qb + g, make_double2(f * (gp.two_over_vol * hv.x + gk2 * cv.x),
                                  f * (gp.two_over_vol * hv.y + gk2 * cv.y)));
    kin += gk2 * (cv.x * cv.x + cv.y * cv.y);
    if constexpr (BE) pe += cv.x * hv.x + cv.y * hv.y;
  }
  red[tid] = kin;
  if constexpr (BE) red[NT + tid] = pe;
  __syncthreads();
  if (tid < 64) {
    const double tot = wave0_sum(red);
    if (tid == 0) kin_partial[st * gridDim.x + blockIdx.x] = tot;
    if constexpr (BE) {
      const double tpe = wave0_sum(red + NT);
      if (tid == 0) gp.pot_partial[st * gridDim.x + blockIdx.x] = tpe;
    }
  }
}

// ------------------------------------------------------ z-plane passes --
// Block (z, state group).  LDS holds one z-plane P[y][xa] of the state in
// hand: the active x columns (compact, xa < nxa <= 64), rows of kZRow
// elements.  Lanes (column xa, j) own column xa in the y transforms (its
// first 64 rows are the column's exchange slots); lanes (line y, j) own
// row y in the x transforms (its 64 slots are the line's exchange slots,
// used once the line's inputs are in registers).  CP column passes cover
// nxa <= CP * BLK / 8.
template <int N, int CP, int BLK = ZCfg<N>::BLOCK>
struct ZPlane {
  static constexpr int BLOCK = BLK, V = N / 8, LPB = BLOCK / 8;
  static constexpr int YP = N / LPB;  // line passes
  // Prefetch of the next state's column inputs into registers while the
  // current state is transformed.
  static constexpr bool PF = ZCfg<N>::PF;
  static constexpr int PV = PF ? V : 1;
  // The density pass at 128 points runs the x transforms of each z-plane in
  // two blocks (y halves; both redo the column transforms), which halves
  // the accumulators each lane holds; the v(r) psi pass re-reads v(r) per
  // line pass instead of holding it in registers.
  static constexpr int YS = N == 64 ? 1 : 2;
  static constexpr int YPD = N / YS / LPB;  // line passes of a density block
  static constexpr bool VREG = N == 64;

  // Per-lane index tables held as pairs of 16-bit signed values per
  // register (row indices < 4096 and x-plane indices < 64 fit), -1 = off
  // the mask.
  struct Packed {
    unsigned w[V / 2];
    __device__ int operator[](int m) const {
      return (int)(short)(w[m >> 1] >> (16 * (m & 1)));
    }
    __device__ void set(int m, int v) {
      const unsigned h = (unsigned)(v & 0xffff) << (16 * (m & 1));
      w[m >> 1] = (m & 1) ? ((w[m >> 1] & 0xffffu) | h) : h;
    }
  };
  // Row index of (column ln + p*LPB, y = j + 8m), or -1.
  __device__ static void row_table(Packed (&ri)[CP], const BoxGeom& geo,
                                   int ln, int j) {
#pragma unroll
    for (int p = 0; p < CP; ++p) {
      const int xa = ln + p * LPB;
#pragma unroll
      for (int m = 0; m < V; ++m)
        ri[p].set(m, xa < geo.nxa ? geo.rowidx[xa * N + j + 8 * m] : -1);
    }
  }
  __device__ static void x_table(Packed& xo, const BoxGeom& geo, int j) {
#pragma unroll
    for (int m = 0; m < V; ++m) xo.set(m, geo.xmap[j + 8 * m]);
  }
  // Loads without branches: off-mask lanes read row 0 and drop the value.
  __device__ static void load(double2 (&pre)[CP][V], const Packed (&ri)[CP],
                              const double2* src) {
#pragma unroll
    for (int p = 0; p < CP; ++p)
#pragma unroll
      for (int m = 0; m < V; ++m) {
        const int r = ri[p][m];
        if (JRX_ZBRANCHFREE) {
          const double2 a = nt_load(src + (unsigned)(r < 0 ? 0 : r));
          pre[p][m] = r < 0 ? czero() : a;
        } else {
          pre[p][m] = r >= 0 ? nt_load(src + r) : czero();
        }
      }
  }
  // Element (y, xa) of P sits at y * kZRow + (xa ^ (y & 7)) (see kZRow).
  __device__ static int at(int y, int xa) {
    return y * kZRow + (xa ^ (y & 7));
  }
  // Exchange slots of the transform of column xa, and of line y.
  struct ColAddr {
    double2* sm;
    int xa;
    __device__ double2* operator()(int s) const { return sm + at(s, xa); }
  };
  struct LineAddr {
    double2* row;
    __device__ double2* operator()(int s) const { return row + s; }
  };
  // Lane twiddles of the 64-point transform kept in registers (the
  // 128-point one reads its table: its registers are taken).
  struct Tw {
    double2 w[8];
    __device__ void init(const double2* stw, int j) {
      if constexpr (N == 64) {
#pragma unroll
        for (int a = 0; a < 8; ++a) w[a] = stw[(j * a) & 63];
      }
    }
  };
  template <bool INV, class Addr>
  __device__ static void fft(double2 (&v)[V], Addr addr, int j,
                             const double2* stw, const Tw& tw) {
    if constexpr (N == 64) fft64_8x8_rw<INV>(v, addr, j, tw.w);
    else fft128_16x8_a<INV>(v, addr, j, stw);
  }
  // Inverse y transform of the active columns (inputs in pre) into P.
  __device__ static void y_inverse(double2* sm, const double2 (&pre)[CP][V],
                                   const BoxGeom& geo, int ln, int j,
                                   const double2* stw, const Tw& tw) {
#pragma unroll
    for (int p = 0; p < CP; ++p) {
      const int xa = ln + p * LPB;
      if (xa < geo.nxa) {
        double2 v[V];
#pragma unroll
        for (int m = 0; m < V; ++m) v[m] = pre[p][m];
        fft<true>(v, ColAddr{sm, xa}, j, stw, tw);
#pragma unroll
        for (int t = 0; t < V; ++t) sm[at(j + 8 * t, xa)] = v[t];
      }
    }
  }
  // Line y of P into registers: v[m] = P[y][xmap[j + 8m]] (0 off the mask).
  __device__ static void line_load(double2 (&v)[V], const double2* sm, int y,
                                   const Packed& xo) {
#pragma unroll
    for (int m = 0; m < V; ++m) {
      const int x = xo[m];
      if (JRX_ZBRANCHFREE) {
        const double2 a = sm[at(y, x < 0 ? 0 : x)];
        v[m] = x < 0 ? czero() : a;
      } else {
        v[m] = x >= 0 ? sm[at(y, x)] : czero();
      }
    }
  }
};

template <int N, int CP, int BLK>
__global__ __launch_bounds__(BLK, ZCfg<N>::WPE > 0 ? ZCfg<N>::WPE : 1) void
zplane_density_kernel(
    const double2* __restrict__ R, double* __restrict__ rho_part,
    const double* __restrict__ occ, int64_t state0, int nstates,
    double inv_vol, BoxGeom geo, const double2* __restrict__ tw, int store) {
  using Z = ZPlane<N, CP, BLK>;
  constexpr int BLOCK = Z::BLOCK, V = Z::V, LPB = Z::LPB, YP = Z::YPD;
  extern __shared__ double2 sm[];  // P [N][kZRow], twiddles [N]
  double2* stw = sm + N * kZRow;
  const int tid = threadIdx.x, j = tid & 7, ln = tid >> 3;
  const int z = blockIdx.x;
  const int G = gridDim.y, gi = blockIdx.y;
  const int per = (nstates + G - 1) / G;
  const int cb = gi * per, ce = min(nstates, cb + per);
  const int y0 = blockIdx.z * (N / Z::YS);  // first line of this block
  double* part = rho_part + ((int64_t)gi * N + z) * N * N;
  for (int i = tid; i < N; i += BLOCK) stw[i] = tw[i];
  __syncthreads();
  if (cb >= ce) {  // uniform over the block: an empty group
    if (store)
#pragma unroll
      for (int p = 0; p < YP; ++p)
#pragma unroll
        for (int t = 0; t < V; ++t)
          part[(y0 + ln + p * LPB) * N + j + 8 * t] = 0.0;
    return;
  }
  typename Z::Packed ri[CP], xo;
  Z::row_table(ri, geo, ln, j);
  Z::x_table(xo, geo, j);
  typename Z::Tw rtw;
  rtw.init(stw, j);
  double acc[YP][V];
#pragma unroll
  for (int p = 0; p < YP; ++p)
#pragma unroll
    for (int t = 0; t < V; ++t) acc[p][t] = 0.0;
  const int64_t sstride = (int64_t)N * geo.rstride;
  const double2* src0 = R + (int64_t)z * geo.rstride;
  double2 pre[CP][Z::PV];
  if constexpr (Z::PF) Z::load(pre, ri, src0 + (int64_t)cb * sstride);
  for (int c = cb; c < ce; ++c) {
    {
      double2 cur[CP][V];
      if constexpr (Z::PF) {
#pragma unroll
        for (int p = 0; p < CP; ++p)
#pragma unroll
          for (int m = 0; m < V; ++m) cur[p][m] = pre[p][m];
        if (c + 1 < ce) Z::load(pre, ri, src0 + (int64_t)(c + 1) * sstride);
      } else {
        Z::load(cur, ri, src0 + (int64_t)c * sstride);
      }
      Z::y_inverse(sm, cur, geo, ln, j, stw, rtw);
    }
    __syncthreads();
    const double f = occ[state0 + c];
#pragma unroll
    for (int p = 0; p < YP; ++p) {
      const int y = y0 + ln + p * LPB;
      double2* row = sm + y * kZRow;
      double2 v[V];
      Z::line_load(v, sm, y, xo);
      Z::template fft<true>(v, typename Z::LineAddr{row}, j, stw, rtw);
#pragma unroll
      for (int t = 0; t < V; ++t)
        acc[p][t] += f * (v[t].x * v[t].x + v[t].y * v[t].y);
    }
    __syncthreads();
  }
  // the group's plane: stored by the first sweep chunk, added to after
#pragma unroll
  for (int p = 0; p < YP; ++p) {
    const int y = y0 + ln + p * LPB;
#pragma unroll
    for (int t = 0; t < V; ++t) {
      double* dst = part + y * N + j + 8 * t;
      *dst = store ? acc[p][t] * inv_vol : *dst + acc[p][t] * inv_vol;
    }
  }
}

template <int N, int CP, int BLK>
__global__ __launch_bounds__(BLK, ZCfg<N>::WPE > 0 ? ZCfg<N>::WPE : 1) void
zplane_vpsi_kernel(
    double2* __restrict__ R, const double* __restrict__ vpot, int nstates,
    BoxGeom geo, const double2* __restrict__ tw) {
  using Z = ZPlane<N, CP, BLK>;
  constexpr int BLOCK = Z::BLOCK, V = Z::V, LPB = Z::LPB, YP = Z::YP;
  extern __shared__ double2 sm[];
  double2* stw = sm + N * kZRow;
  const int tid = threadIdx.x, j = tid & 7, ln = tid >> 3;
  const int z = blockIdx.x;
  const int G = gridDim.y, gi = blockIdx.y;
  const int per = (nstates + G - 1) / G;
  const int cb = gi * per, ce = min(nstates, cb + per);
  for (int i = tid; i < N; i += BLOCK) stw[i] = tw[i];
  __syncthreads();
  if (cb >= ce) return;
  typename Z::Packed ri[CP], xo;
  Z::row_table(ri, geo, ln, j);
  Z::x_table(xo, geo, j);
  typename Z::Tw rtw;
  rtw.init(stw, j);
  const double* vz = vpot + (int64_t)z * N * N;
  double vr[Z::VREG ? YP : 1][V];
  if constexpr (Z::VREG) {
#pragma unroll
    for (int p = 0; p < YP; ++p)
#pragma unroll
      for (int t = 0; t < V; ++t)
        vr[p][t] = vz[(ln + p * LPB) * N + j + 8 * t];
  }
  const int64_t sstride = (int64_t)N * geo.rstride;
  double2* buf0 = R + (int64_t)z * geo.rstride;
  double2 pre[CP][Z::PV];
  if constexpr (Z::PF) Z::load(pre, ri, buf0 + (int64_t)cb * sstride);
  for (int c = cb; c < ce; ++c) {
    double2* buf = buf0 + (int64_t)c * sstride;
    {
      double2 cur[CP][V];
      if constexpr (Z::PF) {
#pragma unroll
        for (int p = 0; p < CP; ++p)
#pragma unroll
          for (int m = 0; m < V; ++m) cur[p][m] = pre[p][m];
        if (c + 1 < ce) Z::load(pre, ri, buf + sstride);
      } else {
        Z::load(cur, ri, buf);
      }
      Z::y_inverse(sm, cur, geo, ln, j, stw, rtw);
    }
    __syncthreads();
    auto line_pass = [&](int p) {
      const int y = ln + p * LPB;
      double2* row = sm + y * kZRow;
      double2 v[V];
      Z::line_load(v, sm, y, xo);
      Z::template fft<true>(v, typename Z::LineAddr{row}, j, stw, rtw);
      if constexpr (Z::VREG) {
#pragma unroll
        for (int t = 0; t < V; ++t) v[t] = cscale(v[t], vr[p][t]);
      } else {
        const double* vl = vz + (ln + p * LPB) * N + j;
#pragma unroll
        for (int t = 0; t < V; ++t) v[t] = cscale(v[t], vl[8 * t]);
      }
      Z::template fft<false>(v, typename Z::LineAddr{row}, j, stw, rtw);
      // (branch-free form: off-mask x go to the row's pad slot kZRow - 1,
      // which nothing reads)
#pragma unroll
      for (int t = 0; t < V; ++t) {
        const int x = xo[t];
        if (JRX_ZBRANCHFREE)
          sm[x < 0 ? y * kZRow + kZRow - 1 : Z::at(y, x)] = v[t];
        else if (x >= 0)
          sm[Z::at(y, x)] = v[t];
      }
    };
    // v(r) in registers needs compile-time pass indices; the 128-point
    // form keeps the pass loop rolled (register pressure)
    if constexpr (Z::VREG) {
#pragma unroll
      for (int p = 0; p < YP; ++p) line_pass(p);
    } else {
#pragma unroll 1
      for (int p = 0; p < YP; ++p) line_pass(p);
    }
    __syncthreads();
    // forward y transform of the active columns, back to the rows of R
    // (the lanes of a column are those that own it in y_inverse, so the
    // next state's y_inverse needs no barrier after this)
#pragma unroll
    for (int p = 0; p < CP; ++p) {
      const int xa = ln + p * LPB;
      if (xa < geo.nxa) {
        double2 v[V];
#pragma unroll
        for (int m = 0; m < V; ++m) v[m] = sm[Z::at(j + 8 * m, xa)];
        Z::template fft<false>(v, typename Z::ColAddr{sm, xa}, j, stw, rtw);
        // off-mask rows go to the slice's pad element (index nrows).  Plain
        // stores: a column's rows start anywhere in a cache line, and the
        // L2 merges the partial lines that neighbouring columns complete
        // (nontemporal stores wrote them back separately: +50% writes).
#pragma unroll
        for (int t = 0; t < V; ++t) {
          const int r = ri[p][t];
          if (JRX_ZBRANCHFREE) buf[(unsigned)(r < 0 ? geo.nrows : r)] = v[t];
          else if (r >= 0) buf[r] = v[t];
        }
      }
    }
  }
}

// ------------------------------------------------------------ launchers --
bool use_rowpath(const BoxGeom& geo) {
  const int n = geo.nx;
  return (n == 64 || n == 128) && geo.ny == n && geo.nz == n &&
         geo.nxa <= 64 && geo.nrows > 0;
}

int row_chunks(const BoxGeom& geo) {
  const int rpb = geo.nx == 64 ? RowCfg<64>::RPB : RowCfg<128>::RPB;
  return (geo.nrows + rpb - 1) / rpb;
}

int zplane_groups(const BoxGeom& geo, int nstates) {
  // about 2048 blocks (z-planes x groups) when every group gets 16 states
  // or more; with fewer states, fewer groups down to about 512 blocks, each
  // group then taking several states (one state per block repeats the
  // block's set-up for little work)
  const int hi = (2048 + geo.nz - 1) / geo.nz, lo = (512 + geo.nz - 1) / geo.nz;
  int G = nstates / 16;
  if (G > hi) G = hi;
  if (G < lo) G = lo;
  if (G > nstates) G = nstates;
  return G < 1 ? 1 : G;
}

template <int N>
static int row_ifft_n(const double2* q, double2* R, int64_t state0,
                      int nstates, const BoxGeom& geo, const double2* tw,
                      hipStream_t s) {
  const size_t shm = (size_t)row_lds_elems<N>() * sizeof(double2);
  hipLaunchKernelGGL(row_ifft_kernel<N>, dim3(row_chunks(geo), nstates),
                     dim3(RowCfg<N>::RPB * 8), shm, s, q, R, state0, geo,
                     tw);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int launch_row_ifft(const double2* q, double2* R, int64_t state0, int nstates,
                    const BoxGeom& geo, const double2* tw, hipStream_t s) {
  if (!use_rowpath(geo)) {
    set_error("row path: unsupported box");
    return JRX_ERR_UNSUPPORTED;
  }
  return geo.nx == 64 ? row_ifft_n<64>(q, R, state0, nstates, geo, tw, s)
                      : row_ifft_n<128>(q, R, state0, nstates, geo, tw, s);
}

template <int N>
static int row_gather_n(const double2* R, const double2* q, double2* qbar,
                        double* kin_partial, int64_t state0, int nstates,
                        const BoxGeom& geo, const double2* tw,
                        const GatherParams& gp, hipStream_t s) {
  constexpr int RPB = RowCfg<N>::RPB;
  const size_t shm = (size_t)(N * (RPB + 1) + N) * sizeof(double2) +
                     2 * RPB * 8 * sizeof(double);
  const dim3 grid(row_chunks(geo), nstates), block(RPB * 8);
  if (gp.pot_partial)
    hipLaunchKernelGGL((row_fft_gather_kernel<N, true>), grid, block, shm, s,
                       R, q, qbar, kin_partial, state0, geo, tw, gp);
  else
    hipLaunchKernelGGL((row_fft_gather_kernel<N, false>), grid, block, shm, s,
                       R, q, qbar, kin_partial, state0, geo, tw, gp);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int launch_row_fft_gather(const double2* R, const double2* q, double2* qbar,
                          double* kin_partial, int64_t state0, int nstates,
                          const BoxGeom& geo, const double2* tw,
                          const GatherParams& gp, hipStream_t s) {
  if (!use_rowpath(geo)) {
    set_error("row path: unsupported box");
    return JRX_ERR_UNSUPPORTED;
  }
  return geo.nx == 64
             ? row_gather_n<64>(R, q, qbar, kin_partial, state0, nstates, geo,
                                tw, gp, s)
             : row_gather_n<128>(R, q, qbar, kin_partial, state0, nstates,
                                 geo, tw, gp, s);
}

template <int N, int CP, int BLK = ZCfg<N>::BLOCK>
static int zdens_n(const double2* R, double* rho_part, int groups,
                   const double* occ, int64_t state0, int nstates,
                   double inv_vol, const BoxGeom& geo, const double2* tw,
                   bool store, hipStream_t s) {
  const size_t shm = (size_t)(N * kZRow + N) * sizeof(double2);
  hipLaunchKernelGGL((zplane_density_kernel<N, CP, BLK>),
                     dim3(N, groups, ZPlane<N, CP, BLK>::YS),
                     dim3(BLK), shm, s, R, rho_part, occ, state0,
                     nstates, inv_vol, geo, tw, store ? 1 : 0);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int launch_zplane_density(const double2* R, double* rho_part, int groups,
                          const double* occ, int64_t state0, int nstates,
                          double inv_vol, const BoxGeom& geo,
                          const double2* tw, bool store, hipStream_t s) {
  if (!use_rowpath(geo)) {
    set_error("row path: unsupported box");
    return JRX_ERR_UNSUPPORTED;
  }
  if (geo.nx == 128)
    return zdens_n<128, 1>(R, rho_part, groups, occ, state0, nstates, inv_vol,
                           geo, tw, store, s);
  if (geo.nxa <= ZCfg<64>::BLOCK / 8)
    return zdens_n<64, 1>(R, rho_part, groups, occ, state0, nstates, inv_vol,
                          geo, tw, store, s);
  // more active columns than a 256-thread block has lane groups: a
  // 512-thread block (one column pass; a second pass of 256-thread blocks
  // doubled the registers and halved the occupancy)
  return zdens_n<64, 1, 512>(R, rho_part, groups, occ, state0, nstates,
                             inv_vol, geo, tw, store, s);
}

template <int N, int CP, int BLK = ZCfg<N>::BLOCK>
static int zvpsi_n(double2* R, const double* v, int groups, int nstates,
                   const BoxGeom& geo, const double2* tw, hipStream_t s) {
  const size_t shm = (size_t)(N * kZRow + N) * sizeof(double2);
  hipLaunchKernelGGL((zplane_vpsi_kernel<N, CP, BLK>), dim3(N, groups),
                     dim3(BLK), shm, s, R, v, nstates, geo, tw);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int launch_zplane_vpsi(double2* R, const double* v, int groups, int nstates,
                       const BoxGeom& geo, const double2* tw, hipStream_t s) {
  if (!use_rowpath(geo)) {
    set_error("row path: unsupported box");
    return JRX_ERR_UNSUPPORTED;
  }
  if (geo.nx == 128) return zvpsi_n<128, 1>(R, v, groups, nstates, geo, tw, s);
  if (geo.nxa <= ZCfg<64>::BLOCK / 8)
    return zvpsi_n<64, 1>(R, v, groups, nstates, geo, tw, s);
  return zvpsi_n<64, 1, 512>(R, v, groups, nstates, geo, tw, s);
}

// out[x][y][z] = in[z][y][x] for one (s) box of n^3 (32 x 32 tiles).
__global__ __launch_bounds__(256) void zmajor_to_xmajor_kernel(
    const double* __restrict__ in, double* __restrict__ out, int n) {
  __shared__ double tile[32][33];
  const int y = blockIdx.z;
  const int x0 = blockIdx.x * 32, z0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int z = z0 + k, x = x0 + tx;
    tile[k][tx] = (z < n && x < n) ? in[((int64_t)z * n + y) * n + x] : 0.0;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int x = x0 + k, z = z0 + tx;
    if (x < n && z < n) out[((int64_t)x * n + y) * n + z] = tile[tx][k];
  }
}

int launch_zmajor_to_xmajor(const double* in, double* out, int ns, int n,
                            hipStream_t s) {
  const int64_t N = (int64_t)n * n * n;
  const int t = (n + 31) / 32;
  for (int sp = 0; sp < ns; ++sp) {
    hipLaunchKernelGGL(zmajor_to_xmajor_kernel, dim3(t, t, n), dim3(256), 0,
                       s, in + sp * N, out + sp * N, n);
    JRX_LAUNCH_CHECK();
  }
  return JRX_OK;
}

}  // namespace jrx
