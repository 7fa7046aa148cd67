// Workgroup-cooperative mixed-radix FFT on lines held in LDS (FP64).
//
// Lines of length n live in LDS; element j of line l is at
//   buf[base(l) + j*es],  base(l) = (lmap ? lmap[l] : l) * ls.
// Each stage is a radix-R Stockham autosort pass (decimation in time):
//   thread-butterfly i in [0, n/R), k = i mod p (p = product of the radices
//   already applied), u_j = x[i + j n/R] * w^{j k n/(pR)}, y = DFT_R(u),
//   x'[(i-k) R + k + t p] = y_t.
// Reads of a stage complete (barrier) before its writes, so one buffer
// suffices.  All transforms are unnormalised; INV selects e^{+i}.
#pragma once
#include <type_traits>

#include "common.h"

namespace jrx {

constexpr int kMaxStages = 12;

// Streaming access to the FFT intermediates (written once and read once per
// sweep: gigabytes, far beyond the caches) carries the nontemporal hint.
__device__ __forceinline__ void nt_store(double2* p, double2 v) {
  __builtin_nontemporal_store(v.x, &p->x);
  __builtin_nontemporal_store(v.y, &p->y);
}
__device__ __forceinline__ double2 nt_load(const double2* p) {
  return make_double2(__builtin_nontemporal_load(&p->x),
                      __builtin_nontemporal_load(&p->y));
}

struct FftLen {
  int n;
  int nstages;
  int radix[kMaxStages];
  const double2* tw;  // device, tw[t] = exp(-2 pi i t / n), t in [0, n)
};

// Factor n into radices 8, 4, 2, 3, 5, 7 (n must be 7-smooth).  Returns
// false if n has another prime factor.
inline bool fft_factorize(int n, FftLen* L) {
  L->n = n;
  L->nstages = 0;
  int m = n;
  auto push = [&](int r) {
    if (L->nstages >= kMaxStages) return false;
    L->radix[L->nstages++] = r;
    return true;
  };
  while (m % 8 == 0) { if (!push(8)) return false; m /= 8; }
  while (m % 4 == 0) { if (!push(4)) return false; m /= 4; }
  while (m % 2 == 0) { if (!push(2)) return false; m /= 2; }
  for (int r : {3, 5, 7})
    while (m % r == 0) { if (!push(r)) return false; m /= r; }
  return m == 1;
}

template <bool INV>
__device__ __forceinline__ double2 twid(const double2* __restrict__ tw,
                                        int idx) {
  double2 w = tw[idx];
  if (INV) w.y = -w.y;
  return w;
}

// Multiply by -i (forward) or +i (inverse).
template <bool INV>
__device__ __forceinline__ double2 mul_mi(double2 a) {
  return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

template <int R, bool INV>
struct Dft;

template <bool INV>
struct Dft<2, INV> {
  __device__ __forceinline__ static void run(double2* u, const double2*, int) {
    double2 a = u[0], b = u[1];
    u[0] = cadd(a, b);
    u[1] = csub(a, b);
  }
};

template <bool INV>
struct Dft<4, INV> {
  __device__ __forceinline__ static void run(double2* u, const double2*, int) {
    double2 a0 = cadd(u[0], u[2]), a1 = csub(u[0], u[2]);
    double2 b0 = cadd(u[1], u[3]), b1 = mul_mi<INV>(csub(u[1], u[3]));
    u[0] = cadd(a0, b0);
    u[2] = csub(a0, b0);
    u[1] = cadd(a1, b1);
    u[3] = csub(a1, b1);
  }
};

template <bool INV>
struct Dft<8, INV> {
  __device__ __forceinline__ static void run(double2* u, const double2*, int) {
    constexpr double c = 0.70710678118654752440;  // 1/sqrt(2)
    double2 e[4] = {u[0], u[2], u[4], u[6]};
    double2 o[4] = {u[1], u[3], u[5], u[7]};
    Dft<4, INV>::run(e, nullptr, 0);
    Dft<4, INV>::run(o, nullptr, 0);
    // o1 *= w8, o2 *= w8^2, o3 *= w8^3 with w8 = exp(-+ i pi/4)
    double2 t1 = o[1], t3 = o[3];
    if (!INV) {
      o[1] = make_double2(c * (t1.x + t1.y), c * (t1.y - t1.x));
      o[3] = make_double2(c * (t3.y - t3.x), -c * (t3.x + t3.y));
    } else {
      o[1] = make_double2(c * (t1.x - t1.y), c * (t1.x + t1.y));
      o[3] = make_double2(-c * (t3.x + t3.y), c * (t3.x - t3.y));
    }
    o[2] = mul_mi<INV>(o[2]);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      u[t] = cadd(e[t], o[t]);
      u[t + 4] = csub(e[t], o[t]);
    }
  }
};

// Odd radices: direct R-point DFT with roots taken from the length-n table
// (w_R^m = tw[m n/R]), so no separately rounded constants are involved.
template <int R, bool INV>
struct DftOdd {
  __device__ __forceinline__ static void run(double2* u,
                                             const double2* __restrict__ tw,
                                             int n) {
    const int step = n / R;
    double2 y[R];
#pragma unroll
    for (int t = 0; t < R; ++t) {
      double2 acc = u[0];
#pragma unroll
      for (int j = 1; j < R; ++j) {
        acc = cadd(acc, cmul(u[j], twid<INV>(tw, ((j * t) % R) * step)));
      }
      y[t] = acc;
    }
#pragma unroll
    for (int t = 0; t < R; ++t) u[t] = y[t];
  }
};
template <bool INV>
struct Dft<3, INV> : DftOdd<3, INV> {};
template <bool INV>
struct Dft<5, INV> : DftOdd<5, INV> {};
template <bool INV>
struct Dft<7, INV> : DftOdd<7, INV> {};

// One radix-R Stockham stage over `nlines` lines.  Requires
// n * nlines <= CAP and blockDim.x == BLOCK.  LINE_FAST maps consecutive
// threads to consecutive lines (use when lines are adjacent in LDS, es > 1);
// otherwise consecutive threads walk along a line (use when es == 1).
template <int R, bool INV, int BLOCK, int CAP, bool LINE_FAST>
__device__ __forceinline__ void fft_stage(double2* buf, int n, int nlines,
                                          int ls, int es,
                                          const short* lmap, int p,
                                          const double2* __restrict__ tw) {
  constexpr int MAXB = (CAP + BLOCK * R - 1) / (BLOCK * R);
  const int T = n / R;
  const int total = T * nlines;
  const int st = n / (p * R);
  double2 v[MAXB][R];
#pragma unroll
  for (int c = 0; c < MAXB; ++c) {
    const int b = threadIdx.x + c * BLOCK;
    if (b < total) {
      const int l = LINE_FAST ? b % nlines : b / T;
      const int i = LINE_FAST ? b / nlines : b - l * T;
      const int base = (lmap ? (int)lmap[l] : l) * ls;
      const int k = i % p;
#pragma unroll
      for (int j = 0; j < R; ++j) v[c][j] = buf[base + (i + j * T) * es];
#pragma unroll
      for (int j = 1; j < R; ++j)
        v[c][j] = cmul(v[c][j], twid<INV>(tw, j * k * st));
      Dft<R, INV>::run(v[c], tw, n);
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < MAXB; ++c) {
    const int b = threadIdx.x + c * BLOCK;
    if (b < total) {
      const int l = LINE_FAST ? b % nlines : b / T;
      const int i = LINE_FAST ? b / nlines : b - l * T;
      const int base = (lmap ? (int)lmap[l] : l) * ls;
      const int k = i % p;
      const int o = (i - k) * R + k;
#pragma unroll
      for (int t = 0; t < R; ++t) buf[base + (o + t * p) * es] = v[c][t];
    }
  }
  __syncthreads();
}

// Full transform of `nlines` lines.  Caller guarantees L.n * nlines <= CAP
// and that the buffer is visible to all threads (barrier before the call).
template <bool INV, int BLOCK, int CAP, bool LINE_FAST>
__device__ void lds_fft(double2* buf, int nlines, int ls, int es,
                        const short* lmap, const FftLen& L) {
  int p = 1;
  for (int s = 0; s < L.nstages; ++s) {
    const int R = L.radix[s];
    switch (R) {
      case 8: fft_stage<8, INV, BLOCK, CAP, LINE_FAST>(buf, L.n, nlines, ls, es, lmap, p, L.tw); break;
      case 4: fft_stage<4, INV, BLOCK, CAP, LINE_FAST>(buf, L.n, nlines, ls, es, lmap, p, L.tw); break;
      case 2: fft_stage<2, INV, BLOCK, CAP, LINE_FAST>(buf, L.n, nlines, ls, es, lmap, p, L.tw); break;
      case 3: fft_stage<3, INV, BLOCK, CAP, LINE_FAST>(buf, L.n, nlines, ls, es, lmap, p, L.tw); break;
      case 5: fft_stage<5, INV, BLOCK, CAP, LINE_FAST>(buf, L.n, nlines, ls, es, lmap, p, L.tw); break;
      default: fft_stage<7, INV, BLOCK, CAP, LINE_FAST>(buf, L.n, nlines, ls, es, lmap, p, L.tw); break;
    }
    p *= R;
  }
}

// ------------------------------------------------- compile-time schedules --
// Radix schedules of the lengths with specialised kernels; the stage loop is
// unrolled and every index constant folds.
template <int... Rs>
struct RadixList {};
template <int N>
struct RadixOf {
  using type = void;  // no compile-time schedule: use lds_fft
};
// Same radix order as fft_factorize (8s, 4s, 2s, then 3, 5, 7), so the
// compile-time and runtime schedules give identical results.
template <> struct RadixOf<16> { using type = RadixList<8, 2>; };
template <> struct RadixOf<32> { using type = RadixList<8, 4>; };
template <> struct RadixOf<35> { using type = RadixList<5, 7>; };
template <> struct RadixOf<36> { using type = RadixList<4, 3, 3>; };
template <> struct RadixOf<40> { using type = RadixList<8, 5>; };
template <> struct RadixOf<42> { using type = RadixList<2, 3, 7>; };
template <> struct RadixOf<45> { using type = RadixList<3, 3, 5>; };
template <> struct RadixOf<48> { using type = RadixList<8, 2, 3>; };
template <> struct RadixOf<49> { using type = RadixList<7, 7>; };
template <> struct RadixOf<50> { using type = RadixList<2, 5, 5>; };
template <> struct RadixOf<54> { using type = RadixList<2, 3, 3, 3>; };
template <> struct RadixOf<56> { using type = RadixList<8, 7>; };
template <> struct RadixOf<60> { using type = RadixList<4, 3, 5>; };
template <> struct RadixOf<63> { using type = RadixList<3, 3, 7>; };
template <> struct RadixOf<64> { using type = RadixList<8, 8>; };
template <> struct RadixOf<70> { using type = RadixList<2, 5, 7>; };
template <> struct RadixOf<72> { using type = RadixList<8, 3, 3>; };
template <> struct RadixOf<75> { using type = RadixList<3, 5, 5>; };
template <> struct RadixOf<80> { using type = RadixList<8, 2, 5>; };
template <> struct RadixOf<128> { using type = RadixList<8, 4, 4>; };

template <bool INV, int BLOCK, int CAP, bool LINE_FAST, int N, int P, int R,
          int... Rest>
__device__ __forceinline__ void ct_stages(double2* buf, int nlines, int ls,
                                          int es, const short* lmap,
                                          const double2* __restrict__ tw) {
  fft_stage<R, INV, BLOCK, CAP, LINE_FAST>(buf, N, nlines, ls, es, lmap, P,
                                           tw);
  if constexpr (sizeof...(Rest) > 0)
    ct_stages<INV, BLOCK, CAP, LINE_FAST, N, P * R, Rest...>(buf, nlines, ls,
                                                             es, lmap, tw);
}

template <bool INV, int BLOCK, int CAP, bool LINE_FAST, int N, int... Rs>
__device__ __forceinline__ void ct_fft_impl(RadixList<Rs...>, double2* buf,
                                            int nlines, int ls, int es,
                                            const short* lmap,
                                            const double2* tw) {
  ct_stages<INV, BLOCK, CAP, LINE_FAST, N, 1, Rs...>(buf, nlines, ls, es,
                                                     lmap, tw);
}

// Transform of length N (compile-time schedule when N > 0 has one, the
// runtime schedule of L otherwise).
template <int N, bool INV, int BLOCK, int CAP, bool LINE_FAST>
__device__ __forceinline__ void fft_lines(double2* buf, int nlines, int ls,
                                          int es, const short* lmap,
                                          const FftLen& L) {
  if constexpr (N > 0 && !std::is_same<typename RadixOf<N>::type, void>::value)
    ct_fft_impl<INV, BLOCK, CAP, LINE_FAST, N>(typename RadixOf<N>::type{},
                                               buf, nlines, ls, es, lmap,
                                               L.tw);
  else
    lds_fft<INV, BLOCK, CAP, LINE_FAST>(buf, nlines, ls, es, lmap, L);
}

// ------------------------------------------- register 64-point transforms --
// Eight lanes own one line of 64 points; lane j holds x = j + 8m in v[m]
// (m = 0..7) on entry and X[j + 8t] in v[t] on exit:
//   Y_j[a] = sum_m x[j+8m] w8^{ma};  Y_j[a] *= w64^{ja};
//   exchange through 64 LDS slots of the line (slot(a, j) = 8a + (j+a)%8,
//   bank-swizzled);  X[b + 8t] = sum_j Y_j[b] w8^{jt} on lane b.
// Slots are base[s * es].  The lanes of a line are in one wave, so the
// exchange needs only wave-level ordering.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool INV>
__device__ __forceinline__ void fft64_8x8(double2 (&v)[8], double2* base,
                                          int es, int j,
                                          const double2* __restrict__ tw) {
  Dft<8, INV>::run(v, nullptr, 0);
#pragma unroll
  for (int a = 1; a < 8; ++a) v[a] = cmul(v[a], twid<INV>(tw, j * a));
  wave_sync();
#pragma unroll
  for (int a = 0; a < 8; ++a) base[(8 * a + ((j + a) & 7)) * es] = v[a];
  wave_sync();
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = base[(8 * j + ((i + j) & 7)) * es];
  wave_sync();
  Dft<8, INV>::run(v, nullptr, 0);
}

// fft64_8x8 with the lane's twiddles w[a] = w64^{j a} (a = 1..7, forward
// sign) held in registers, and exchange slot s at addr(s) (any one-to-one
// map of the 64 slots, e.g. a bank swizzle).
template <bool INV, class Addr>
__device__ __forceinline__ void fft64_8x8_rw(double2 (&v)[8], Addr addr,
                                             int j, const double2 (&w)[8]) {
  Dft<8, INV>::run(v, nullptr, 0);
#pragma unroll
  for (int a = 1; a < 8; ++a) v[a] = cmul(v[a], INV ? cconj(w[a]) : w[a]);
  wave_sync();
#pragma unroll
  for (int a = 0; a < 8; ++a) *addr(8 * a + ((j + a) & 7)) = v[a];
  wave_sync();
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = *addr(8 * j + ((i + j) & 7));
  wave_sync();
  Dft<8, INV>::run(v, nullptr, 0);
}

// 16-point DFT in registers: radix-2 split into two 8-point DFTs, with the
// roots w16^t = tw[8t] taken from the length-128 table.
template <bool INV>
__device__ __forceinline__ void dft16(double2 (&u)[16],
                                      const double2* __restrict__ tw128) {
  double2 e[8], o[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    e[t] = u[2 * t];
    o[t] = u[2 * t + 1];
  }
  Dft<8, INV>::run(e, nullptr, 0);
  Dft<8, INV>::run(o, nullptr, 0);
#pragma unroll
  for (int t = 1; t < 8; ++t) o[t] = cmul(o[t], twid<INV>(tw128, 8 * t));
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    u[t] = cadd(e[t], o[t]);
    u[t + 8] = csub(e[t], o[t]);
  }
}

// 128 points on eight lanes: lane j holds x = j + 8m in v[m] (m = 0..15)
// on entry and X[j + 8u] in v[u] on exit (the same layout), with
//   Y_j[a] = sum_m x[j+8m] w16^{ma} (registers),  Y_j[a] *= w128^{ja},
//   X[a + 16b] = sum_j Y_j[a] w8^{jb}  (lane a % 8, after an exchange).
// The exchange goes through 64 LDS slots base[s * es] in two rounds
// (a < 8, then a >= 8), so a line needs the same 64 slots as a 64-point
// one.  All lanes of a line are in one wave.
template <bool INV, class Addr>
__device__ __forceinline__ void fft128_16x8_a(double2 (&v)[16], Addr addr,
                                              int j,
                                              const double2* __restrict__ tw) {
  dft16<INV>(v, tw);
#pragma unroll
  for (int a = 1; a < 16; ++a) v[a] = cmul(v[a], twid<INV>(tw, j * a));
  double2 z0[8], z1[8];
  wave_sync();
#pragma unroll
  for (int a = 0; a < 8; ++a) *addr(8 * a + ((j + a) & 7)) = v[a];
  wave_sync();
#pragma unroll
  for (int i = 0; i < 8; ++i) z0[i] = *addr(8 * j + ((i + j) & 7));
  wave_sync();
#pragma unroll
  for (int a = 0; a < 8; ++a) *addr(8 * a + ((j + a) & 7)) = v[a + 8];
  wave_sync();
#pragma unroll
  for (int i = 0; i < 8; ++i) z1[i] = *addr(8 * j + ((i + j) & 7));
  wave_sync();
  Dft<8, INV>::run(z0, nullptr, 0);
  Dft<8, INV>::run(z1, nullptr, 0);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    v[2 * b] = z0[b];
    v[2 * b + 1] = z1[b];
  }
}
template <bool INV>
__device__ __forceinline__ void fft128_16x8(double2 (&v)[16], double2* base,
                                            int es, int j,
                                            const double2* __restrict__ tw) {
  fft128_16x8_a<INV>(v, [=](int s) { return base + s * es; }, j, tw);
}

// Register transform of one line of N points on eight lanes (N/8 values
// per lane, element j + 8m of the line in v[m] on lane j, same layout on
// exit); 64 exchange slots base[s * es].
template <int N>
struct RegFft;
template <>
struct RegFft<64> {
  static constexpr int V = 8;
  template <bool INV>
  __device__ __forceinline__ static void run(double2 (&v)[8], double2* base,
                                             int es, int j,
                                             const double2* tw) {
    fft64_8x8<INV>(v, base, es, j, tw);
  }
};
template <>
struct RegFft<128> {
  static constexpr int V = 16;
  template <bool INV>
  __device__ __forceinline__ static void run(double2 (&v)[16], double2* base,
                                             int es, int j,
                                             const double2* tw) {
    fft128_16x8<INV>(v, base, es, j, tw);
  }
};

}  // namespace jrx
