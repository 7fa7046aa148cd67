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

#include <algorithm>

namespace jrx {

GramSplit gram_split(int64_t P, int ng, int nb) {
  const int tiles = (ng + kQrRows - 1) / kQrRows;
  // about 2048 blocks in all; the wide kernels do four to sixteen times
  // the work per row, and each chunk adds an nb x nb partial to reduce
  const int64_t target = nb > kNarrowBands ? 256 : 2048;
  int64_t want = (target + P - 1) / P;  // chunks per problem
  if (want < 1) want = 1;
  if (want > tiles) want = tiles;
  const int tiles_per_chunk = (int)((tiles + want - 1) / want);
  GramSplit sp;
  sp.rows_per_chunk = tiles_per_chunk * kQrRows;
  sp.nchunk = (ng + sp.rows_per_chunk - 1) / sp.rows_per_chunk;
  return sp;
}


// ------------------------------------------------------------- FP64 MFMA --
// v_mfma_f64_16x16x4f64 on one wave: D(16x16) += A(16x4) B(4x16), lane l
// supplies A[l%16][l/16] and B[l/16][l%16]; accumulator register r of lane
// l holds D[4r + l/16][l%16].  Complex products are four real products.
typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma4(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ d4 d4zero() { return d4{0.0, 0.0, 0.0, 0.0}; }

// Small nb x nb matrix into LDS, zero-padded to NBP x NBP.
__device__ __forceinline__ void load_small(double2* dst, const double2* src,
                                           int nb, int NBP) {
  for (int e = threadIdx.x; e < NBP * NBP; e += kQrBlock) {
    const int i = e / NBP, j = e - i * NBP;
    dst[e] = (i < nb && j < nb) ? src[i * nb + j] : czero();
  }
}

// ------------------------------------------------- wide kernels (nb > 64) --
// Threads of the wide Gram kernel: 512 at 8 tiles (128 bands: 36 or 64
// pairs over 8 waves, 5 or 8 per wave); below, 256 (the compiler keeps the
// accumulators in AGPRs there and spills at 512).
constexpr int gram_wide_block(int T) { return T == 8 ? 512 : 256; }

// Tile pairs (I, J) of the wide Gram kernels: HERM takes I <= J (row-major
// over the upper triangle; the result is Hermitian), otherwise all T x T.
// NZ blocks share one chunk's pairs, PW pairs to a wave (at most 9).
template <int T, bool HERM>
struct WidePairs {
  static constexpr int NW = gram_wide_block(T) / 64;
  static constexpr int NPAIR = HERM ? T * (T + 1) / 2 : T * T;
  static constexpr int NZ = (NPAIR + 9 * NW - 1) / (9 * NW);
  static constexpr int PW = (NPAIR + NW * NZ - 1) / (NW * NZ);
};

__device__ inline void decode_pair(int q, int T, bool herm, int* I, int* J) {
  if (!herm) {
    *I = q / T;
    *J = q - *I * T;
    return;
  }
  int i = 0;
  while (q >= T - i) {
    q -= T - i;
    ++i;
  }
  *I = i;
  *J = i + q;
}

// Rows per LDS slab of the wide Gram kernel: 16 for a Hermitian result
// (one staged matrix), 8 for Q^H Qbar (two), so that the two slab buffers
// fit twice per CU.
template <bool HERM>
constexpr int wide_slab() { return HERM ? 16 : 8; }

// Partial Gram sums over a chunk's rows, one partial per chunk:
//   GMAJOR: S = W^H W (W planar g-major, the Hermitian pairs only);
//   else:   G = A^H B (A, B band-major complex; HERM when B is A).
// Row slabs are staged in LDS ([row][NBP + 1]), double-buffered: the next
// slab's global loads are in flight in registers while the MFMAs run on
// the current one, one barrier per slab.  Each wave accumulates its own
// tile pairs and writes them (and, for HERM, their mirror images).
template <int T, bool GMAJOR, bool HERM>
__global__ __launch_bounds__(gram_wide_block(T)) void gram_wide_kernel(
    const double* __restrict__ wr, const double* __restrict__ wi,
    const double2* __restrict__ qa, const double2* __restrict__ qb,
    double2* __restrict__ partial, int ng, int nb, int rows_per_chunk,
    int nchunk, const int* __restrict__ skip) {
  using PL = WidePairs<T, HERM>;
  constexpr int NBP = 16 * T, LD = NBP + 1, PW = PL::PW;
  constexpr int SR = wide_slab<HERM>();
  constexpr int NS = SR * NBP;                        // elements per slab
  constexpr int BLK = gram_wide_block(T);
  constexpr int NE = (NS + BLK - 1) / BLK;  // elements per thread
  constexpr int BUF = (HERM ? 1 : 2) * SR * LD;       // one slab buffer
  extern __shared__ double2 sm[];
  // (the wave index through readfirstlane: the pair tables are scalars)
  const int lane = threadIdx.x & 63,
            wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int li = lane & 15, lk = lane >> 4;
  const int64_t p = blockIdx.y;
  const int chunk = blockIdx.x;
  if (skip && skip[p]) return;  // uniform over the block
  int pI[PW], pJ[PW];
  bool live[PW];
#pragma unroll
  for (int t = 0; t < PW; ++t) {
    const int q = ((int)blockIdx.z * PL::NW + wave) * PW + t;
    live[t] = q < PL::NPAIR;  // uniform over the wave
    decode_pair(live[t] ? q : 0, T, HERM, &pI[t], &pJ[t]);
  }
  d4 re[PW], im[PW];
#pragma unroll
  for (int t = 0; t < PW; ++t) { re[t] = d4zero(); im[t] = d4zero(); }
  const int gb = chunk * rows_per_chunk;
  const int ge = min(ng, gb + rows_per_chunk);
  const double* Wr = wr + p * ng * nb;
  const double* Wi = wi + p * ng * nb;
  const double2* A = qa + p * nb * ng;
  const double2* B = qb + p * nb * ng;
  double2 ra[NE], rb[HERM ? 1 : NE];
  // slab element e: GMAJOR row e / NBP, band e % NBP (consecutive threads
  // read consecutive bands of a row); band-major band e / SR, row e % SR
  // (consecutive threads read consecutive rows of a band)
  auto fetch = [&](int g0) {
#pragma unroll
    for (int i = 0; i < NE; ++i) {
      const int e = threadIdx.x + i * BLK;
      const int r = GMAJOR ? e / NBP : e % SR;
      const int c = GMAJOR ? e % NBP : e / SR;
      const int g = g0 + r;
      const bool ok = e < NS && g < ge && c < nb;
      if (GMAJOR) {
        ra[i] = ok ? make_double2(Wr[(int64_t)g * nb + c], Wi[(int64_t)g * nb + c])
                   : czero();
      } else {
        ra[i] = ok ? A[(int64_t)c * ng + g] : czero();
        if (!HERM) rb[i] = ok ? B[(int64_t)c * ng + g] : czero();
      }
    }
  };
  auto store = [&](int buf) {
    double2* sa = sm + buf * BUF;
#pragma unroll
    for (int i = 0; i < NE; ++i) {
      const int e = threadIdx.x + i * BLK;
      if (e < NS) {
        const int r = GMAJOR ? e / NBP : e % SR;
        const int c = GMAJOR ? e % NBP : e / SR;
        sa[r * LD + c] = ra[i];
        if (!HERM) sa[SR * LD + r * LD + c] = rb[i];
      }
    }
  };
  if (gb < ge) {
    fetch(gb);
    store(0);
  }
  __syncthreads();
  int buf = 0;
  for (int g0 = gb; g0 < ge; g0 += SR) {
    const bool more = g0 + SR < ge;  // uniform over the block
    if (more) fetch(g0 + SR);
    const double2* sa = sm + buf * BUF;
    const double2* sb = HERM ? sa : sa + SR * LD;
#pragma unroll
    for (int r = 0; r < SR / 4; ++r) {
      const double2* ra4 = sa + (4 * r + lk) * LD + li;
      const double2* rb4 = sb + (4 * r + lk) * LD + li;
#pragma unroll
      for (int t = 0; t < PW; ++t) {
        if (!live[t]) continue;
        const double2 a = ra4[16 * pI[t]], b = rb4[16 * pJ[t]];
        // conj(a) b
        re[t] = mfma4(a.x, b.x, re[t]);
        re[t] = mfma4(a.y, b.y, re[t]);
        im[t] = mfma4(a.x, b.y, im[t]);
        im[t] = mfma4(-a.y, b.x, im[t]);
      }
    }
    // the other buffer was last read before the previous barrier
    if (more) store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  double2* out = partial + (p * nchunk + chunk) * nb * nb;
#pragma unroll
  for (int t = 0; t < PW; ++t) {
    if (!live[t]) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * pI[t] + 4 * r + lk, j = 16 * pJ[t] + li;
      if (i < nb && j < nb) {
        const double2 v = make_double2(re[t][r], im[t][r]);
        out[i * nb + j] = v;
        if (HERM && pI[t] < pJ[t]) out[j * nb + i] = cconj(v);
      }
    }
  }
}

// GMAJOR when wr is given (qa, qb unused), else band-major qa, qb (HERM when
// qa == qb).
int launch_gram_wide(const double* wr, const double* wi, const double2* qa,
                     const double2* qb, double2* partial, int64_t P, int ng,
                     int nb, const GramSplit& sp, hipStream_t s,
                     const int* skip) {
  const int T = (nb + 15) / 16;
  const bool gmajor = wr != nullptr, herm = gmajor || qa == qb;
  // two slab buffers of one (HERM) or two staged matrices
  const size_t shm = 2 * (herm ? 1 : 2) *
                     (size_t)(herm ? wide_slab<true>() : wide_slab<false>()) *
                     (16 * T + 1) * sizeof(double2);
#define JRX_GRAM_WIDE(TT, GM, HM)                                            \
  hipLaunchKernelGGL((gram_wide_kernel<TT, GM, HM>),                         \
                     dim3(sp.nchunk, (unsigned)P, WidePairs<TT, HM>::NZ),    \
                     dim3(gram_wide_block(TT)), shm, s, wr, wi, qa, qb,     \
                     partial, ng, nb,                                        \
                     sp.rows_per_chunk, sp.nchunk, skip)
#define JRX_GRAM_WIDE_T(TT)                              \
  if (gmajor) JRX_GRAM_WIDE(TT, true, true);             \
  else if (herm) JRX_GRAM_WIDE(TT, false, true);         \
  else JRX_GRAM_WIDE(TT, false, false);
  switch (T) {
    case 5: JRX_GRAM_WIDE_T(5) break;
    case 6: JRX_GRAM_WIDE_T(6) break;
    case 7: JRX_GRAM_WIDE_T(7) break;
    case 8: JRX_GRAM_WIDE_T(8) break;
    default:
      set_error("nb above 128 is not supported by this build");
      return JRX_ERR_UNSUPPORTED;
  }
#undef JRX_GRAM_WIDE_T
#undef JRX_GRAM_WIDE
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

constexpr int kCholWideBlock = 1024;
constexpr int kInvLanes = 8;  // threads per row in the inverse

// Element (i, j), i <= j, of an upper triangle packed row by row.
__device__ inline int up_index(int i, int j, int nb) {
  return i * nb - (i * (i - 1)) / 2 + (j - i);
}

// chol_inv_kernel for nb > 64: the upper triangle, packed, is the only
// matrix in LDS (132 KB at nb = 128), factored and inverted in place with
// one barrier per step.
// Factorisation, step j (right-looking, on the unscaled row j):
//   A[r][c] -= conj(A[j][r]) A[j][c] / A[j][j]   (j < r <= c),
// waves over rows r, lanes over columns c; row j itself is scaled to
// R[j][c] = A[j][c] / sqrt(A[j][j]) during step j + 1, which does not read
// it (the pivots are kept aside in piv).
// Inverse, column j (rows are independent given R):
//   X[j][j] = 1 / R[j][j],  X[i][j] = -sum_{k=i}^{j-1} X[i][k] R[k][j] / R[j][j],
// kInvLanes threads per row i sharing the sum; column j of R is copied
// aside (double-buffered) one step ahead, before X overwrites it.
__global__ __launch_bounds__(kCholWideBlock) void chol_inv_wide_kernel(
    const double2* __restrict__ S, double2* __restrict__ rinv, int nb,
    int* __restrict__ flag_out, const int* __restrict__ flag_in) {
  extern __shared__ double2 sm[];
  const int np = nb * (nb + 1) / 2;
  double2* A = sm;
  double2* col = sm + np;  // [2][nb]
  double* piv = reinterpret_cast<double*>(col + 2 * nb);  // [nb]
  double* red = piv + nb;                                  // [B / 64]
  const int64_t p = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int B = kCholWideBlock, NW = B / 64;
  double2* out = rinv + p * nb * nb;
  if (flag_in && flag_in[p]) {  // uniform over the block
    for (int e = tid; e < nb * nb; e += B) {
      const int i = e / nb, j = e - i * nb;
      out[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    }
    return;
  }
  const double2* Sp = S + p * nb * nb;
  for (int e = tid; e < nb * nb; e += B) {
    const int i = e / nb, j = e - i * nb;
    if (j >= i) A[up_index(i, j, nb)] = Sp[e];
  }
  __syncthreads();
  if (flag_out) {  // distance of the scaled Gram matrix from I
    double a = 0.0;
    for (int e = tid; e < nb * nb; e += B) {
      const int i = e / nb, j = e - i * nb;
      if (i < j) {  // (i, j) and (j, i)
        const double2 v = A[up_index(i, j, nb)];
        a += 2.0 * (v.x * v.x + v.y * v.y) /
             (A[up_index(i, i, nb)].x * A[up_index(j, j, nb)].x);
      }
    }
    const double d2 = block_sum(a, red);
    if (tid == 0) flag_out[p] = (d2 < kQrOnePass * kQrOnePass) ? 1 : 0;
  }
  // NaN for a non-positive pivot: sqrt of a negative number.
  auto root = [](double d) { return d > 0.0 ? sqrt(d) : sqrt(-1.0 - d * 0.0); };
  for (int j = 0; j <= nb; ++j) {
    if (j >= 1) {  // scale row j - 1 (its pivot is piv[j - 1])
      const int jm = j - 1, base = up_index(jm, jm, nb);
      const double sq = root(piv[jm]);
      for (int c = tid; c < nb - jm; c += B)
        A[base + c] = c == 0 ? make_double2(sq, 0.0) : cscale(A[base + c], 1.0 / sq);
    }
    if (j < nb) {
      const int base = up_index(j, j, nb);
      const double dd = A[base].x;
      if (tid == 0) piv[j] = dd;
      const double inv = 1.0 / dd;
      // two rows per pass, every load issued before the stores (nb <= 128:
      // columns r + lane and r + lane + 64 cover a row)
      for (int r = j + 1 + wave; r < nb; r += 2 * NW) {
        const int r2 = r + NW;
        const bool two = r2 < nb;
        double2 x[2][2], y[2][2], f[2];
        int rb[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int rr = h ? (two ? r2 : r) : r;
          const double2 ajr = A[base + rr - j];
          f[h] = make_double2(ajr.x * inv, -ajr.y * inv);  // conj / dd
          rb[h] = up_index(rr, rr, nb);
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            const int c = rr + lane + 64 * u;
            x[h][u] = c < nb ? A[rb[h] + c - rr] : czero();
            y[h][u] = c < nb ? A[base + c - j] : czero();
          }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          if (h && !two) break;
          const int rr = h ? r2 : r;
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            const int c = rr + lane + 64 * u;
            if (c < nb) A[rb[h] + c - rr] = csub(x[h][u], cmul(f[h], y[h][u]));
          }
        }
      }
    }
    __syncthreads();
  }
  // inverse, in place
  const int i = tid / kInvLanes, t = tid % kInvLanes;
  const int ib = i < nb ? up_index(i, i, nb) : 0;
  if (tid <= 0) col[0] = A[0];  // column 0 of R
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    const double2* cj = col + (j & 1) * nb;
    if (i < j) {
      double2 acc = czero(), acc2 = czero();
      int k = i + t;
      for (; k + kInvLanes < j; k += 2 * kInvLanes) {
        acc = cadd(acc, cmul(A[ib + k - i], cj[k]));
        acc2 = cadd(acc2, cmul(A[ib + k + kInvLanes - i], cj[k + kInvLanes]));
      }
      if (k < j) acc = cadd(acc, cmul(A[ib + k - i], cj[k]));
      acc = cadd(acc, acc2);
#pragma unroll
      for (int m = kInvLanes / 2; m > 0; m >>= 1) {
        acc.x += __shfl_xor(acc.x, m, kInvLanes);
        acc.y += __shfl_xor(acc.y, m, kInvLanes);
      }
      if (t == 0) A[ib + j - i] = cscale(acc, -1.0 / cj[j].x);
    } else if (i == j && t == 0) {
      A[ib] = make_double2(1.0 / cj[j].x, 0.0);
    }
    if (j + 1 < nb) {  // column j + 1 of R, untouched until the next step
      double2* cn = col + ((j + 1) & 1) * nb;
      for (int k = tid; k <= j + 1; k += B) cn[k] = A[up_index(k, j + 1, nb)];
    }
    __syncthreads();
  }
  for (int e = tid; e < nb * nb; e += B) {
    const int r = e / nb, c = e - r * nb;
    out[e] = (c >= r) ? A[up_index(r, c, nb)] : czero();
  }
}

// qr_bwd_small_kernel for nb > 64: one output entry per thread, operands
// read from global memory (L2-resident nb x nb matrices).
__global__ __launch_bounds__(kQrBlock) void qr_bwd_small_wide_kernel(
    const double2* __restrict__ G, const double2* __restrict__ rinv,
    double2* __restrict__ A, double2* __restrict__ Bm, int nb) {
  const int64_t p = blockIdx.y;
  const int e = blockIdx.x * kQrBlock + threadIdx.x;
  if (e >= nb * nb) return;
  const int i = e / nb, j = e - i * nb;
  const double2* Gp = G + p * nb * nb;
  const double2* Rp = rinv + p * nb * nb;
  A[p * nb * nb + e] = cconj(Rp[j * nb + i]);
  double2 acc = czero();
  for (int k = j; k < nb; ++k) {
    double2 m;  // copyltu(triu(-G))[i][k]
    if (i < k) m = make_double2(-Gp[i * nb + k].x, -Gp[i * nb + k].y);
    else if (i > k) m = make_double2(-Gp[k * nb + i].x, Gp[k * nb + i].y);
    else m = make_double2(-Gp[i * nb + i].x, 0.0);
    acc = cadd(acc, cmul(m, cconj(Rp[j * nb + k])));
  }
  Bm[p * nb * nb + e] = acc;
}

constexpr int kWideRows = 64;  // g rows per block of the wide apply kernels
constexpr int kWideK = 32;     // rinv rows per LDS chunk (apply_fwd)
constexpr int kWideKB = 16;    // band rows per LDS chunk (apply_bwd)

// apply_fwd_kernel for nb > 64: rinv staged in chunks of kWideK rows, one
// 16-row slab of W per wave; tiles below the diagonal of rinv are skipped.
template <int T>
__global__ __launch_bounds__(kQrBlock) void apply_fwd_wide_kernel(
    const double* __restrict__ wr, const double* __restrict__ wi,
    const double2* __restrict__ rinv, double2* __restrict__ q, int ng, int nb,
    const int* __restrict__ skip) {
  extern __shared__ double2 sm[];  // [kWideK][NBP]
  constexpr int NBP = 16 * T;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t p = blockIdx.y;
  if (skip && skip[p]) return;  // uniform over the block
  const double2* Rp = rinv + p * nb * nb;
  const double* Wr = wr + p * ng * nb;
  const double* Wi = wi + p * ng * nb;
  double2* Q = q + p * nb * ng;
  const int g0 = blockIdx.x * kWideRows + wave * 16;
  const int row = g0 + li;
  d4 qre[T], qim[T];
#pragma unroll
  for (int J = 0; J < T; ++J) { qre[J] = d4zero(); qim[J] = d4zero(); }
  for (int k0 = 0; k0 < nb; k0 += kWideK) {
    __syncthreads();  // the previous chunk is consumed
    for (int e = threadIdx.x; e < kWideK * NBP; e += kQrBlock) {
      const int kk = e / NBP, b = e - kk * NBP, k = k0 + kk;
      sm[e] = (k < nb && b < nb) ? Rp[k * nb + b] : czero();
    }
    __syncthreads();
#pragma unroll 2
    for (int m = 0; m < kWideK / 4; ++m) {
      const int col = k0 + 4 * m + lk;
      const bool ok = row < ng && col < nb;
      const double a_r = ok ? Wr[(int64_t)row * nb + col] : 0.0;
      const double a_i = ok ? Wi[(int64_t)row * nb + col] : 0.0;
#pragma unroll
      for (int J = 0; J < T; ++J) {
        if (16 * J + 15 < k0) continue;  // rinv[k][b] = 0 for k > b
        const double2 rv = sm[(4 * m + lk) * NBP + 16 * J + li];
        qre[J] = mfma4(rv.x, a_r, qre[J]);
        qre[J] = mfma4(-rv.y, a_i, qre[J]);
        qim[J] = mfma4(rv.x, a_i, qim[J]);
        qim[J] = mfma4(rv.y, a_r, qim[J]);
      }
    }
  }
#pragma unroll
  for (int J = 0; J < T; ++J)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int b = 16 * J + 4 * r + lk;
      if (b < nb && row < ng)
        Q[(int64_t)b * ng + row] = make_double2(qre[J][r], qim[J][r]);
    }
}

// apply_bwd_kernel for nb > 64: A and Bm staged in chunks of kWideKB band
// rows, one 16-row slab of the output per wave; A = rinv^H is lower
// triangular, so its tiles right of the chunk are skipped.
template <int T>
__global__ __launch_bounds__(kQrBlock) void apply_bwd_wide_kernel(
    const double2* __restrict__ qbar, const double2* __restrict__ q,
    const double2* __restrict__ A, const double2* __restrict__ Bm,
    double* __restrict__ dwr, double* __restrict__ dwi, int ng, int nb) {
  extern __shared__ double2 sm[];
  constexpr int NBP = 16 * T;
  double2* sa = sm;                   // [kWideKB][NBP]
  double2* sb = sm + kWideKB * NBP;   // [kWideKB][NBP]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t p = blockIdx.y;
  const double2* Ap = A + p * nb * nb;
  const double2* Bp = Bm + p * nb * nb;
  const double2* Qb = qbar + p * nb * ng;
  const double2* Q = q + p * nb * ng;
  double* Dr = dwr + p * ng * nb;
  double* Di = dwi + p * ng * nb;
  const int g0 = blockIdx.x * kWideRows + wave * 16;
  const int g = g0 + li;
  d4 re[T], im[T];
#pragma unroll
  for (int J = 0; J < T; ++J) { re[J] = d4zero(); im[J] = d4zero(); }
  for (int k0 = 0; k0 < nb; k0 += kWideKB) {
    __syncthreads();
    for (int e = threadIdx.x; e < kWideKB * NBP; e += kQrBlock) {
      const int kk = e / NBP, b = e - kk * NBP, k = k0 + kk;
      const bool ok = k < nb && b < nb;
      sa[e] = ok ? Ap[k * nb + b] : czero();
      sb[e] = ok ? Bp[k * nb + b] : czero();
    }
    __syncthreads();
#pragma unroll 2
    for (int m = 0; m < kWideKB / 4; ++m) {
      const int band = k0 + 4 * m + lk;
      const bool ok = g < ng && band < nb;
      const double2 vb = ok ? Qb[(int64_t)band * ng + g] : czero();
      const double2 vq = ok ? Q[(int64_t)band * ng + g] : czero();
#pragma unroll
      for (int J = 0; J < T; ++J) {
        if (16 * J < k0 + kWideKB) {  // A[i][j] = 0 for j > i
          const double2 a = sa[(4 * m + lk) * NBP + 16 * J + li];
          re[J] = mfma4(vb.x, a.x, re[J]);
          re[J] = mfma4(-vb.y, a.y, re[J]);
          im[J] = mfma4(vb.x, a.y, im[J]);
          im[J] = mfma4(vb.y, a.x, im[J]);
        }
        const double2 b = sb[(4 * m + lk) * NBP + 16 * J + li];
        re[J] = mfma4(vq.x, b.x, re[J]);
        re[J] = mfma4(-vq.y, b.y, re[J]);
        im[J] = mfma4(vq.x, b.y, im[J]);
        im[J] = mfma4(vq.y, b.x, im[J]);
      }
    }
  }
#pragma unroll
  for (int J = 0; J < T; ++J)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = g0 + 4 * r + lk, col = 16 * J + li;
      if (row < ng && col < nb) {
        Dr[(int64_t)row * nb + col] = re[J][r];
        Di[(int64_t)row * nb + col] = im[J][r];
      }
    }
}

// Partial Gram sums over the chunk's rows (one per block):
//   S = A^H A with A = W (PRE = false) or A = W pre (PRE = true).
// Waves take 16-row slabs in turn; only tile pairs I <= J are accumulated
// and the full Hermitian matrix is written.
template <int T, bool PRE>
__global__ __launch_bounds__(kQrBlock) void gram_w_kernel(
    const double* __restrict__ wr, const double* __restrict__ wi,
    const double2* __restrict__ pre, double2* __restrict__ partial, int ng,
    int nb, int rows_per_chunk, int nchunk, const int* __restrict__ skip) {
  extern __shared__ double2 sm[];
  constexpr int NBP = 16 * T;
  constexpr int NP = T * (T + 1) / 2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t p = blockIdx.y;
  const int chunk = blockIdx.x;
  if (skip && skip[p]) return;  // uniform over the block
  if (PRE) load_small(sm, pre + p * nb * nb, nb, NBP);
  __syncthreads();
  d4 sre[NP], sim[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) { sre[i] = d4zero(); sim[i] = d4zero(); }
  const int gb = chunk * rows_per_chunk;
  const int ge = min(ng, gb + rows_per_chunk);
  const double* Wr = wr + p * ng * nb;
  const double* Wi = wi + p * ng * nb;
  const int nm = (nb + 3) / 4;
  for (int g0 = gb + wave * 16; g0 < ge; g0 += 64) {
    double xr[T][4], xi[T][4];
    if (PRE) {
      d4 are[T], aim[T];
#pragma unroll
      for (int J = 0; J < T; ++J) { are[J] = d4zero(); aim[J] = d4zero(); }
      const int row = g0 + li;
      for (int m = 0; m < nm; ++m) {
        const int col = 4 * m + lk;
        const bool ok = row < ge && col < nb;
        const double a_r = ok ? Wr[(int64_t)row * nb + col] : 0.0;
        const double a_i = ok ? Wi[(int64_t)row * nb + col] : 0.0;
#pragma unroll
        for (int J = 0; J < T; ++J) {
          const double2 rv = sm[(4 * m + lk) * NBP + 16 * J + li];
          are[J] = mfma4(a_r, rv.x, are[J]);
          are[J] = mfma4(-a_i, rv.y, are[J]);
          aim[J] = mfma4(a_r, rv.y, aim[J]);
          aim[J] = mfma4(a_i, rv.x, aim[J]);
        }
      }
#pragma unroll
      for (int J = 0; J < T; ++J)
#pragma unroll
        for (int r = 0; r < 4; ++r) { xr[J][r] = are[J][r]; xi[J][r] = aim[J][r]; }
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = g0 + 4 * r + lk;
#pragma unroll
        for (int J = 0; J < T; ++J) {
          const int col = 16 * J + li;
          const bool ok = row < ge && col < nb;
          xr[J][r] = ok ? Wr[(int64_t)row * nb + col] : 0.0;
          xi[J][r] = ok ? Wi[(int64_t)row * nb + col] : 0.0;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int pi = 0;
#pragma unroll
      for (int I = 0; I < T; ++I)
#pragma unroll
        for (int J = I; J < T; ++J, ++pi) {
          sre[pi] = mfma4(xr[I][r], xr[J][r], sre[pi]);
          sre[pi] = mfma4(xi[I][r], xi[J][r], sre[pi]);
          sim[pi] = mfma4(xr[I][r], xi[J][r], sim[pi]);
          sim[pi] = mfma4(-xi[I][r], xr[J][r], sim[pi]);
        }
    }
  }
  // Sum the four waves' tiles in LDS (wave order 0..3), one partial per
  // block.
  double2* red = sm + (PRE ? NBP * NBP : 0);
  for (int w = 0; w < 4; ++w) {
    if (wave == w) {
      int pi = 0;
#pragma unroll
      for (int I = 0; I < T; ++I)
#pragma unroll
        for (int J = I; J < T; ++J, ++pi)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = 16 * I + 4 * r + lk, j = 16 * J + li;
            const double2 v = make_double2(sre[pi][r], sim[pi][r]);
            if (w == 0) {
              red[i * NBP + j] = v;
              if (I < J) red[j * NBP + i] = cconj(v);
            } else {
              red[i * NBP + j] = cadd(red[i * NBP + j], v);
              if (I < J) red[j * NBP + i] = cadd(red[j * NBP + i], cconj(v));
            }
          }
    }
    __syncthreads();
  }
  double2* out = partial + (p * nchunk + chunk) * nb * nb;
  for (int e = threadIdx.x; e < nb * nb; e += kQrBlock) {
    const int i = e / nb, j = e - i * nb;
    out[e] = red[i * NBP + j];
  }
}

int launch_gram_gmajor(const double* wr, const double* wi, const double2* pre,
                       double2* partial, int64_t P, int ng, int nb,
                       const GramSplit& sp, hipStream_t s, const int* skip,
                       double2* q1) {
  const int T = (nb + 15) / 16;
  if (nb > kNarrowBands) {
    if (!pre) return launch_gram_wide(wr, wi, nullptr, nullptr, partial, P, ng,
                                      nb, sp, s, skip);
    if (!q1) {
      set_error("the wide Gram pass with a preconditioner needs q1");
      return JRX_ERR_INVALID;
    }
    const int rc = launch_apply_fwd(wr, wi, pre, q1, P, ng, nb, s, skip);
    if (rc) return rc;
    return launch_gram_wide(nullptr, nullptr, q1, q1, partial, P, ng, nb, sp,
                            s, skip);
  }
  size_t shm = (pre ? 2 : 1) * (size_t)(16 * T) * (16 * T) * sizeof(double2);
  dim3 grid(sp.nchunk, (unsigned)P), block(kQrBlock);
#define JRX_GRAM_W(TT)                                                        \
  if (pre)                                                                    \
    hipLaunchKernelGGL((gram_w_kernel<TT, true>), grid, block, shm, s, wr, wi, \
                       pre, partial, ng, nb, sp.rows_per_chunk, sp.nchunk,    \
                       skip);                                                 \
  else                                                                        \
    hipLaunchKernelGGL((gram_w_kernel<TT, false>), grid, block, shm, s, wr,   \
                       wi, pre, partial, ng, nb, sp.rows_per_chunk,           \
                       sp.nchunk, skip);
  switch (T) {
    case 1: JRX_GRAM_W(1) break;
    case 2: JRX_GRAM_W(2) break;
    case 3: JRX_GRAM_W(3) break;
    case 4: JRX_GRAM_W(4) break;
    default:
      set_error("nb above 64 is not supported by this build");
      return JRX_ERR_UNSUPPORTED;
  }
#undef JRX_GRAM_W
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

// Partial sums of G = Q^H Qbar over the chunk (band-major inputs), one per
// block.  A wave takes 16-row slabs in turn, 4 rows at a time (row
// g0 + 4 r + lk, bands 16 I + li).
template <int T>
__global__ __launch_bounds__(kQrBlock) void gram_q_kernel(
    const double2* __restrict__ q, const double2* __restrict__ qb,
    double2* __restrict__ partial, int ng, int nb, int rows_per_chunk,
    int nchunk) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t p = blockIdx.y;
  const int chunk = blockIdx.x;
  d4 gre[T][T], gim[T][T];
#pragma unroll
  for (int I = 0; I < T; ++I)
#pragma unroll
    for (int J = 0; J < T; ++J) { gre[I][J] = d4zero(); gim[I][J] = d4zero(); }
  const int gb = chunk * rows_per_chunk;
  const int ge = min(ng, gb + rows_per_chunk);
  const double2* Q = q + p * nb * ng;
  const double2* B = qb + p * nb * ng;
  auto accum = [&](const double2* a, const double2* b) {
#pragma unroll
    for (int I = 0; I < T; ++I)
#pragma unroll
      for (int J = 0; J < T; ++J) {
        gre[I][J] = mfma4(a[I].x, b[J].x, gre[I][J]);
        gre[I][J] = mfma4(a[I].y, b[J].y, gre[I][J]);
        gim[I][J] = mfma4(a[I].x, b[J].y, gim[I][J]);
        gim[I][J] = mfma4(-a[I].y, b[J].x, gim[I][J]);
      }
  };
  if constexpr (T <= 2) {
    // units of 4 rows; the loads of the next unit are issued before the
    // MFMAs of the current one
    const int gw = gb + wave * 16;  // first slab of this wave
    const int nunit = gw < ge ? ((ge - gw + 63) / 64) * 4 : 0;
    double2 ca[T], cb[T];
    auto load = [&](int u, double2* a, double2* b) {
      const int g = gw + 64 * (u >> 2) + 4 * (u & 3) + lk;
#pragma unroll
      for (int I = 0; I < T; ++I) {
        const int band = 16 * I + li;
        const bool ok = u < nunit && g < ge && band < nb;
        a[I] = ok ? Q[(int64_t)band * ng + g] : czero();
        b[I] = ok ? B[(int64_t)band * ng + g] : czero();
      }
    };
    load(0, ca, cb);
    for (int u = 0; u < nunit; ++u) {
      double2 na[T], nbv[T];
      load(u + 1, na, nbv);
      accum(ca, cb);
#pragma unroll
      for (int I = 0; I < T; ++I) { ca[I] = na[I]; cb[I] = nbv[I]; }
    }
  } else {
    // whole slabs (more registers per unit than a second buffer allows)
    for (int g0 = gb + wave * 16; g0 < ge; g0 += 64) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int g = g0 + 4 * r + lk;
        double2 a[T], b[T];
#pragma unroll
        for (int I = 0; I < T; ++I) {
          const int band = 16 * I + li;
          const bool ok = g < ge && band < nb;
          a[I] = ok ? Q[(int64_t)band * ng + g] : czero();
          b[I] = ok ? B[(int64_t)band * ng + g] : czero();
        }
        accum(a, b);
      }
    }
  }
  // Sum the four waves' tiles in LDS (wave order 0..3), one partial per
  // block.
  extern __shared__ double2 red[];
  constexpr int NBP = 16 * T;
  for (int w = 0; w < 4; ++w) {
    if (wave == w) {
#pragma unroll
      for (int I = 0; I < T; ++I)
#pragma unroll
        for (int J = 0; J < T; ++J)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = 16 * I + 4 * r + lk, j = 16 * J + li;
            const double2 v = make_double2(gre[I][J][r], gim[I][J][r]);
            red[i * NBP + j] = w == 0 ? v : cadd(red[i * NBP + j], v);
          }
    }
    __syncthreads();
  }
  double2* out = partial + (p * nchunk + chunk) * nb * nb;
  for (int e = threadIdx.x; e < nb * nb; e += kQrBlock) {
    const int i = e / nb, j = e - i * nb;
    out[e] = red[i * NBP + j];
  }
}

int launch_gram_bmajor(const double2* q, const double2* qb, double2* partial,
                       int64_t P, int ng, int nb, const GramSplit& sp,
                       hipStream_t s) {
  const int T = (nb + 15) / 16;
  if (nb > kNarrowBands)
    return launch_gram_wide(nullptr, nullptr, q, qb, partial, P, ng, nb, sp, s,
                            nullptr);
  const size_t shm = (size_t)(16 * T) * (16 * T) * sizeof(double2);
  dim3 grid(sp.nchunk, (unsigned)P), block(kQrBlock);
  switch (T) {
    case 1: hipLaunchKernelGGL((gram_q_kernel<1>), grid, block, shm, s, q, qb, partial, ng, nb, sp.rows_per_chunk, sp.nchunk); break;
    case 2: hipLaunchKernelGGL((gram_q_kernel<2>), grid, block, shm, s, q, qb, partial, ng, nb, sp.rows_per_chunk, sp.nchunk); break;
    case 3: hipLaunchKernelGGL((gram_q_kernel<3>), grid, block, shm, s, q, qb, partial, ng, nb, sp.rows_per_chunk, sp.nchunk); break;
    case 4: hipLaunchKernelGGL((gram_q_kernel<4>), grid, block, shm, s, q, qb, partial, ng, nb, sp.rows_per_chunk, sp.nchunk); break;
    default:
      set_error("nb above 64 is not supported by this build");
      return JRX_ERR_UNSUPPORTED;
  }
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

// A block sums kRedOut consecutive outputs: its kRedGroups thread groups
// take chunks c = g, g + kRedGroups, ... in order, then the group sums are
// added in group order (a fixed order for every launch).
constexpr int kRedOut = 32, kRedGroups = kQrBlock / kRedOut;

__global__ __launch_bounds__(kQrBlock) void reduce_partials_kernel(
    const double2* __restrict__ partial, double2* __restrict__ out,
    int nchunk, int nout) {
  __shared__ double2 acc[kRedGroups][kRedOut];
  const int64_t p = blockIdx.y;
  const int ol = threadIdx.x % kRedOut, grp = threadIdx.x / kRedOut;
  const int o = blockIdx.x * kRedOut + ol;
  // four independent sums in flight, added in a fixed order
  double2 a = czero(), b = czero(), c2 = czero(), d = czero();
  if (o < nout) {
    const double2* src = partial + p * nchunk * nout + o;
    int c = grp;
    for (; c + 3 * kRedGroups < nchunk; c += 4 * kRedGroups) {
      a = cadd(a, src[(int64_t)c * nout]);
      b = cadd(b, src[(int64_t)(c + kRedGroups) * nout]);
      c2 = cadd(c2, src[(int64_t)(c + 2 * kRedGroups) * nout]);
      d = cadd(d, src[(int64_t)(c + 3 * kRedGroups) * nout]);
    }
    for (; c < nchunk; c += kRedGroups) a = cadd(a, src[(int64_t)c * nout]);
  }
  a = cadd(cadd(a, b), cadd(c2, d));
  acc[grp][ol] = a;
  __syncthreads();
  if (grp == 0 && o < nout) {
    for (int g = 1; g < kRedGroups; ++g) a = cadd(a, acc[g][ol]);
    out[p * nout + o] = a;
  }
}

int launch_reduce_partials(const double2* partial, double2* out, int64_t P,
                           int nchunk, int nb, hipStream_t s) {
  const int nout = nb * nb;
  dim3 grid(ceil_div(nout, kRedOut), (unsigned)P), block(kQrBlock);
  hipLaunchKernelGGL(reduce_partials_kernel, grid, block, 0, s, partial, out,
                     nchunk, nout);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

// One block per problem: the Cholesky factor of S = R^H R (R upper) and
// rinv = X = R^{-1}, in one sweep of nb + 3 steps with one barrier each.
// Step s:
//  (a) updates the trailing rows with the unscaled pivot row s,
//      A[r][c] -= conj(A[s][r]) A[s][c] / A[s][s];
//  (b) scales row s - 1 into R (R[j][j] = sqrt of the pivot, NaN for a
//      non-positive one);
//  (c) adds row x = s - 2 of R into the partial sums of X R = I: with
//      X[i][c] holding P[i][c] = sum_{k<x} X[i][k] R[k][c] for c >= x,
//      X[i][c] += X_ix R[x][c] (c > x), X_ix = -P[i][x] / R[x][x]
//      (1 / R[x][x] for i = x);
//  (d) writes column y = s - 3 of X from its complete partial sums.
// (a) writes rows > s of A, (b) row s - 1, (c) and (d) read rows <= s - 2;
// (c) writes columns > s - 2 of X and (d) column s - 3, so the parts of a
// step touch disjoint data.  A (then R) and X live in LDS ([nb][nb+1]
// each).
__global__ __launch_bounds__(kQrBlock) void chol_inv_kernel(
    const double2* __restrict__ S, double2* __restrict__ rinv, int nb,
    int* __restrict__ flag_out, const int* __restrict__ flag_in) {
  extern __shared__ double2 sm[];
  __shared__ double dg[kNarrowBands];  // R[j][j]
  const int ld = nb + 1;
  double2* A = sm;
  double2* X = sm + nb * ld;
  const int64_t p = blockIdx.x;
  const int tid = threadIdx.x;
  double2* out = rinv + p * nb * nb;
  if (flag_in && flag_in[p]) {  // uniform over the block
    for (int e = tid; e < nb * nb; e += kQrBlock) {
      const int i = e / nb, j = e - i * nb;
      out[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    }
    return;
  }
  const double2* Sp = S + p * nb * nb;
  for (int e = tid; e < nb * nb; e += kQrBlock) {
    const int i = e / nb, j = e - i * nb;
    A[i * ld + j] = Sp[e];
  }
  __syncthreads();
  if (flag_out) {  // distance of the scaled Gram matrix from I
    double a = 0.0;
    for (int e = tid; e < nb * nb; e += kQrBlock) {
      const int i = e / nb, j = e - i * nb;
      if (i != j) {
        const double2 v = A[i * ld + j];
        a += (v.x * v.x + v.y * v.y) / (A[i * ld + i].x * A[j * ld + j].x);
      }
    }
    // X is free until the sweep
    const double d2 = block_sum(a, reinterpret_cast<double*>(X));
    if (tid == 0)  // NaN or a non-positive diagonal compares false
      flag_out[p] = (d2 < kQrOnePass * kQrOnePass) ? 1 : 0;
  }
  for (int e = tid; e < nb * nb; e += kQrBlock) {
    const int i = e / nb, j = e - i * nb;
    X[i * ld + j] = czero();
  }
  __syncthreads();
  for (int st = 0; st < nb + 3; ++st) {
    if (st < nb) {  // (a)
      const double piv = A[st * ld + st].x;
      const double inv = 1.0 / piv;
      if (tid == 0)  // NaN for a non-positive pivot: sqrt of a negative number
        dg[st] = piv > 0.0 ? sqrt(piv) : sqrt(-1.0 - piv * 0.0);
      const int m = nb - st - 1;
      for (int e = tid; e < m * m; e += kQrBlock) {
        const int r = st + 1 + e / m, c = st + 1 + e % m;
        if (c >= r)
          A[r * ld + c] = csub(A[r * ld + c],
                               cscale(cmulc(A[st * ld + r], A[st * ld + c]), inv));
      }
    }
    if (st >= 1 && st <= nb) {  // (b)
      const int j = st - 1;
      const double d = dg[j];
      const double id = 1.0 / d;
      for (int c = j + tid; c < nb; c += kQrBlock)
        A[j * ld + c] = c == j ? make_double2(d, 0.0) : cscale(A[j * ld + c], id);
    }
    if (st >= 2 && st <= nb + 1) {  // (c)
      const int x = st - 2, w = nb - x - 1;
      const double ir = 1.0 / A[x * ld + x].x;
      for (int e = tid; e < (x + 1) * w; e += kQrBlock) {
        const int i = e / w, c = x + 1 + e % w;
        const double2 xi = i == x ? make_double2(ir, 0.0)
                                  : cscale(X[i * ld + x], -ir);
        X[i * ld + c] = cadd(X[i * ld + c], cmul(xi, A[x * ld + c]));
      }
    }
    if (st >= 3) {  // (d)
      const int y = st - 3;
      const double ir = 1.0 / A[y * ld + y].x;
      for (int i = tid; i <= y; i += kQrBlock)
        X[i * ld + y] = i == y ? make_double2(ir, 0.0)
                               : cscale(X[i * ld + y], -ir);
    }
    __syncthreads();
  }
  for (int e = tid; e < nb * nb; e += kQrBlock) {
    const int i = e / nb, j = e - i * nb;
    out[e] = (j >= i) ? X[i * ld + j] : czero();
  }
}

// chol_inv_kernel for nb <= 32 with the matrices in registers: thread
// (ty, tx) = (tid / 16, tid % 16) owns rows {ty, ty + 16} x columns
// {tx, tx + 16} of A and of X = R^{-1}.  The same sweep as chol_inv_kernel
// (right-looking Cholesky; X from the partial sums of X R = I), with the
// data that crosses threads passed through LDS buffers that are each written
// once, ahead of their use: rowbuf[s] = the unscaled pivot row s (written
// by its owners when step s - 1 has finished it), colbuf[y] = the final
// column y of X (written at the end of step y + 1).  Step s reads only
// buffers written before its barrier, so its loads need no ordering with
// its stores: one barrier per step, nb + 1 steps.
// 1 / sqrt(a): the hardware estimate refined by one Newton step for
// f(y) = 1/y^2 - a, y <- y (3 - a y^2) / 2; NaN for a <= 0 (or NaN), so
// that a non-positive pivot poisons the factor as the IEEE sqrt does.
__device__ __forceinline__ double rsqrt_newton(double a) {
  double y = __builtin_amdgcn_rsq(a);
  y = y * (1.5 - 0.5 * a * y * y);
  return a > 0.0 ? y : __builtin_nan("");
}

// With lmul (upper triangular, [P][nb][nb]) the output is lmul X instead
// of X (the product of the two CholeskyQR2 inverses, summed over k in
// ascending order as trmm_uu_kernel does).
__global__ __launch_bounds__(kQrBlock) void chol_inv32_kernel(
    const double2* __restrict__ S, double2* __restrict__ rinv, int nb,
    int* __restrict__ flag_out, const int* __restrict__ flag_in,
    const double2* __restrict__ lmul) {
  __shared__ double2 rowbuf[32][33];
  __shared__ double2 colbuf[32][33];
  __shared__ double red[kQrBlock / 64];
  const int64_t p = blockIdx.x;
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
  double2* out = rinv + p * nb * nb;
  const double2* L = lmul ? lmul + p * nb * nb : nullptr;
  if (flag_in && flag_in[p]) {  // uniform over the block: X = I
    for (int e = tid; e < nb * nb; e += kQrBlock) {
      const int i = e / nb, j = e - i * nb;
      out[e] = L ? L[e] : make_double2(i == j ? 1.0 : 0.0, 0.0);
    }
    return;
  }
  const double2* Sp = S + p * nb * nb;
  double2 A[4], X[4];
  int rr[2], cc[2];
  rr[0] = ty; rr[1] = ty + 16;
  cc[0] = tx; cc[1] = tx + 16;
  double dist = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = rr[q >> 1], c = cc[q & 1];
    const bool ok = r < nb && c < nb;
    A[q] = ok ? Sp[r * nb + c] : czero();
    X[q] = czero();
    if (ok && r != c)  // distance of the scaled Gram matrix from I
      dist += (A[q].x * A[q].x + A[q].y * A[q].y) /
              (Sp[r * nb + r].x * Sp[c * nb + c].x);
    if (ok && r == 0) rowbuf[0][c] = A[q];
  }
  if (flag_out) {
    const double d2 = block_sum(dist, red);
    if (tid == 0)  // NaN or a non-positive diagonal compares false
      flag_out[p] = (d2 < kQrOnePass * kQrOnePass) ? 1 : 0;
  }
  __syncthreads();
  for (int st = 0; st <= nb; ++st) {
    // Loads, all unconditional (indices clamped into the buffers) so that
    // they issue back to back; values outside the step's ranges are
    // selected away below.  Only buffers written before the last barrier
    // are read.
    const int sa = st < nb ? st : nb - 1;
    const int x = st - 2, xa = x < 0 ? 0 : (x < nb ? x : nb - 1);
    const int y = st - 1, ya = y < 0 ? 0 : (y < nb ? y : nb - 1);
    const double pst = rowbuf[sa][sa].x, px = rowbuf[xa][xa].x,
                 py = rowbuf[ya][ya].x;
    double2 sr[2], sc[2], rx[2], xi[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      sr[a] = rowbuf[sa][rr[a]];
      sc[a] = rowbuf[sa][cc[a]];
      rx[a] = rowbuf[xa][cc[a]];
      xi[a] = colbuf[xa][rr[a]];
    }
    const double ys = rsqrt_newton(pst);
    const double inv = ys * ys;                 // 1 / pivot
    const double idx = rsqrt_newton(px);        // 1 / R[x][x]
    const double iy = rsqrt_newton(py);         // 1 / R[y][y]
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int a = q >> 1, b = q & 1;
      const int r = rr[a], c = cc[b];
      // Cholesky update with pivot row st
      if (st < nb && r > st && c >= r && c < nb)
        A[q] = csub(A[q], cscale(cmulc(sr[a], sc[b]), inv));
      // row x of R into the partial sums of X
      if (x >= 0 && r <= x && c > x && c < nb)
        X[q] = cadd(X[q], cmul(xi[a], cscale(rx[b], idx)));
      // row st + 1 is final: publish it
      if (st + 1 < nb && r == st + 1 && c >= r && c < nb)
        rowbuf[st + 1][c] = A[q];
      // column y of X is complete: finalise and publish it
      if (y >= 0 && y < nb && c == y && r <= y) {
        X[q] = r == y ? make_double2(iy, 0.0) : cscale(X[q], -iy);
        colbuf[y][r] = X[q];
      }
    }
    __syncthreads();
  }
  if (L) {  // X into LDS (the loop's last barrier has retired colbuf)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = rr[q >> 1], c = cc[q & 1];
      colbuf[r][c] = (r < nb && c < nb && c >= r) ? X[q] : czero();
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = rr[q >> 1], c = cc[q & 1];
      if (r < nb && c < nb) {
        double2 acc = czero();
        for (int k = r; k <= c; ++k) acc = cadd(acc, cmul(L[r * nb + k], colbuf[k][c]));
        out[r * nb + c] = acc;
      }
    }
    return;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = rr[q >> 1], c = cc[q & 1];
    if (r < nb && c < nb) out[r * nb + c] = c >= r ? X[q] : czero();
  }
}

int launch_chol_inv(const double2* S, double2* rinv, int64_t P, int nb,
                    hipStream_t s, int* flag_out, const int* flag_in,
                    const double2* lmul) {
  if (lmul && nb > 32) {
    set_error("launch_chol_inv: lmul is for at most 32 bands");
    return JRX_ERR_INVALID;
  }
  if (nb > kNarrowBands) {
    const size_t shm = ((size_t)nb * (nb + 1) / 2 + 2 * nb) * sizeof(double2) +
                       (nb + kCholWideBlock / 64) * sizeof(double);
    hipLaunchKernelGGL(chol_inv_wide_kernel, dim3((unsigned)P),
                       dim3(kCholWideBlock), shm, s, S, rinv, nb, flag_out,
                       flag_in);
    JRX_LAUNCH_CHECK();
    return JRX_OK;
  }
  if (nb <= 32) {
    hipLaunchKernelGGL(chol_inv32_kernel, dim3((unsigned)P), dim3(kQrBlock), 0,
                       s, S, rinv, nb, flag_out, flag_in, lmul);
    JRX_LAUNCH_CHECK();
    return JRX_OK;
  }
  size_t shm = 2 * (size_t)nb * (nb + 1) * sizeof(double2);
  hipLaunchKernelGGL(chol_inv_kernel, dim3((unsigned)P), dim3(kQrBlock), shm,
                     s, S, rinv, nb, flag_out, flag_in);
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
    Ap[e] = i >= j ? cconj(Ri[j * ld + i]) : czero();
    double2 acc = czero();
    for (int k = j; k < nb; ++k) acc = cadd(acc, cmul(M[i * ld + k], cconj(Ri[j * ld + k])));
    Bp[e] = acc;
  }
}

int launch_qr_bwd_small(const double2* G, const double2* rinv, double2* A,
                        double2* Bm, int64_t P, int nb, hipStream_t s) {
  if (nb > kNarrowBands) {
    hipLaunchKernelGGL(qr_bwd_small_wide_kernel,
                       dim3(ceil_div(nb * nb, kQrBlock), (unsigned)P),
                       dim3(kQrBlock), 0, s, G, rinv, A, Bm, nb);
    JRX_LAUNCH_CHECK();
    return JRX_OK;
  }
  size_t shm = 2 * (size_t)nb * (nb + 1) * sizeof(double2);
  hipLaunchKernelGGL(qr_bwd_small_kernel, dim3((unsigned)P), dim3(kQrBlock),
                     shm, s, G, rinv, A, Bm, nb);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}


// Rows per block of the narrow apply kernels: 256, halved (down to 64)
// while the grid would hold fewer than 2048 blocks, so that problems with
// few k-points still spread over the chip.
inline int apply_rows(int ng, int64_t P) {
  int rows = 256;
  while (rows > 64 && (int64_t)ceil_div(ng, rows) * P < 2048) rows >>= 1;
  return rows;
}

// Q[p][b][g] = sum_i W[p][g][i] rinv[p][i][b], computed as Q^T = R^T W^T so
// that each accumulator row is 16 consecutive g of one band.  rinv is upper
// triangular (as the Cholesky inverses and their products are), so up to
// 48 bands the (column group, band tile) pairs below the diagonal are
// skipped.  A wave takes 16-row slabs in turn.
template <int T>
__global__ __launch_bounds__(kQrBlock) void apply_fwd_kernel(
    const double* __restrict__ wr, const double* __restrict__ wi,
    const double2* __restrict__ rinv, double2* __restrict__ q, int ng,
    int nb, int rows) {
  extern __shared__ double2 sm[];
  constexpr int NBP = 16 * T;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t p = blockIdx.y;
  load_small(sm, rinv + p * nb * nb, nb, NBP);
  __syncthreads();
  const double* Wr = wr + p * ng * nb;
  const double* Wi = wi + p * ng * nb;
  double2* Q = q + p * nb * ng;
  const int gb = blockIdx.x * rows;
  const int ge = min(ng, gb + rows);
  const int gw = gb + wave * 16;  // first slab of this wave
  if constexpr (T <= 3) {
    // units of 16 columns; the loads of the next unit are issued before
    // the MFMAs of the current one
    const int nunit = gw < ge ? ((ge - gw + 63) / 64) * T : 0;
    double ar[4], ai[4];
    auto load = [&](int u, double* xr, double* xi) {
      const int row = gw + 64 * (u / T) + li, mg = u % T;
#pragma unroll
      for (int mm = 0; mm < 4; ++mm) {
        const int col = 16 * mg + 4 * mm + lk;
        const bool ok = u < nunit && row < ge && col < nb;
        xr[mm] = ok ? Wr[(int64_t)row * nb + col] : 0.0;
        xi[mm] = ok ? Wi[(int64_t)row * nb + col] : 0.0;
      }
    };
    load(0, ar, ai);
    d4 qre[T], qim[T];
#pragma unroll
    for (int J = 0; J < T; ++J) { qre[J] = d4zero(); qim[J] = d4zero(); }
    for (int u = 0; u < nunit; ++u) {
      double nr[4], ni[4];
      load(u + 1, nr, ni);
      const int mg = u % T;
#pragma unroll
      for (int mm = 0; mm < 4; ++mm) {
        const int k = 16 * mg + 4 * mm + lk;
#pragma unroll
        for (int J = 0; J < T; ++J) {
          if (mg > J) continue;  // rinv[i][b] = 0 for i > b
          const double2 rv = sm[k * NBP + 16 * J + li];
          qre[J] = mfma4(rv.x, ar[mm], qre[J]);
          qre[J] = mfma4(-rv.y, ai[mm], qre[J]);
          qim[J] = mfma4(rv.x, ai[mm], qim[J]);
          qim[J] = mfma4(rv.y, ar[mm], qim[J]);
        }
      }
      if (mg == T - 1) {
        const int g = gw + 64 * (u / T) + li;
#pragma unroll
        for (int J = 0; J < T; ++J) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int b = 16 * J + 4 * r + lk;
            if (b < nb && g < ge)
              Q[(int64_t)b * ng + g] = make_double2(qre[J][r], qim[J][r]);
          }
          qre[J] = d4zero();
          qim[J] = d4zero();
        }
      }
#pragma unroll
      for (int mm = 0; mm < 4; ++mm) { ar[mm] = nr[mm]; ai[mm] = ni[mm]; }
    }
  } else {
    // one column group of 4 at a time (more registers per unit than a
    // second buffer allows)
    const int nm = (nb + 3) / 4;
    for (int g0 = gw; g0 < ge; g0 += 64) {
      d4 qre[T], qim[T];
#pragma unroll
      for (int J = 0; J < T; ++J) { qre[J] = d4zero(); qim[J] = d4zero(); }
      const int row = g0 + li;
      for (int m = 0; m < nm; ++m) {
        const int col = 4 * m + lk;
        const bool ok = row < ge && col < nb;
        const double a_r = ok ? Wr[(int64_t)row * nb + col] : 0.0;
        const double a_i = ok ? Wi[(int64_t)row * nb + col] : 0.0;
#pragma unroll
        for (int J = 0; J < T; ++J) {
          const double2 rv = sm[(4 * m + lk) * NBP + 16 * J + li];
          qre[J] = mfma4(rv.x, a_r, qre[J]);
          qre[J] = mfma4(-rv.y, a_i, qre[J]);
          qim[J] = mfma4(rv.x, a_i, qim[J]);
          qim[J] = mfma4(rv.y, a_r, qim[J]);
        }
      }
      const int g = g0 + li;
#pragma unroll
      for (int J = 0; J < T; ++J)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int b = 16 * J + 4 * r + lk;
          if (b < nb && g < ge)
            Q[(int64_t)b * ng + g] = make_double2(qre[J][r], qim[J][r]);
        }
    }
  }
}

int launch_apply_fwd(const double* wr, const double* wi, const double2* rinv,
                     double2* q, int64_t P, int ng, int nb, hipStream_t s,
                     const int* skip) {
  const int T = (nb + 15) / 16;
  if (nb > kNarrowBands) {
    const size_t shm = (size_t)kWideK * 16 * T * sizeof(double2);
    dim3 grid(ceil_div(ng, kWideRows), (unsigned)P), block(kQrBlock);
    switch (T) {
      case 5: hipLaunchKernelGGL((apply_fwd_wide_kernel<5>), grid, block, shm, s, wr, wi, rinv, q, ng, nb, skip); break;
      case 6: hipLaunchKernelGGL((apply_fwd_wide_kernel<6>), grid, block, shm, s, wr, wi, rinv, q, ng, nb, skip); break;
      case 7: hipLaunchKernelGGL((apply_fwd_wide_kernel<7>), grid, block, shm, s, wr, wi, rinv, q, ng, nb, skip); break;
      case 8: hipLaunchKernelGGL((apply_fwd_wide_kernel<8>), grid, block, shm, s, wr, wi, rinv, q, ng, nb, skip); break;
      default:
        set_error("nb above 128 is not supported by this build");
        return JRX_ERR_UNSUPPORTED;
    }
    JRX_LAUNCH_CHECK();
    return JRX_OK;
  }
  if (skip) {
    set_error("launch_apply_fwd: skip is for the wide kernels");
    return JRX_ERR_INVALID;
  }
  size_t shm = (size_t)(16 * T) * (16 * T) * sizeof(double2);
  const int rows = apply_rows(ng, P);
  dim3 grid(ceil_div(ng, rows), (unsigned)P), block(kQrBlock);
  switch (T) {
    case 1: hipLaunchKernelGGL((apply_fwd_kernel<1>), grid, block, shm, s, wr, wi, rinv, q, ng, nb, rows); break;
    case 2: hipLaunchKernelGGL((apply_fwd_kernel<2>), grid, block, shm, s, wr, wi, rinv, q, ng, nb, rows); break;
    case 3: hipLaunchKernelGGL((apply_fwd_kernel<3>), grid, block, shm, s, wr, wi, rinv, q, ng, nb, rows); break;
    case 4: hipLaunchKernelGGL((apply_fwd_kernel<4>), grid, block, shm, s, wr, wi, rinv, q, ng, nb, rows); break;
    default:
      set_error("nb above 64 is not supported by this build");
      return JRX_ERR_UNSUPPORTED;
  }
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

// Wbar[p][g][j] = sum_i Qbar[p][i][g] A[i][j] + Q[p][i][g] Bm[i][j], planar
// g-major outputs.  A = R^{-H} is lower triangular, so its (band group,
// column tile) pairs above the diagonal are skipped.  A wave takes 16-row
// slabs in turn and reads each slab in T units of 16 bands; the loads of the
// next unit are issued before the MFMAs of the current one.
template <int T>
__global__ __launch_bounds__(kQrBlock) void apply_bwd_kernel(
    const double2* __restrict__ qbar, const double2* __restrict__ q,
    const double2* __restrict__ A, const double2* __restrict__ Bm,
    double* __restrict__ dwr, double* __restrict__ dwi, int ng, int nb,
    int rows) {
  extern __shared__ double2 sm[];
  constexpr int NBP = 16 * T;
  double2* sa = sm;
  double2* sb = sm + NBP * NBP;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t p = blockIdx.y;
  load_small(sa, A + p * nb * nb, nb, NBP);
  load_small(sb, Bm + p * nb * nb, nb, NBP);
  __syncthreads();
  const double2* Qb = qbar + p * nb * ng;
  const double2* Q = q + p * nb * ng;
  double* Dr = dwr + p * ng * nb;
  double* Di = dwi + p * ng * nb;
  const int gb = blockIdx.x * rows;
  const int ge = min(ng, gb + rows);
  const int gw = gb + wave * 16;  // first slab of this wave
  const int nunit = gw < ge ? ((ge - gw + 63) / 64) * T : 0;
  double2 cb[4], cq[4];
  auto load = [&](int u, double2* vb, double2* vq) {
    const int g = gw + 64 * (u / T) + li, mg = u % T;
#pragma unroll
    for (int mm = 0; mm < 4; ++mm) {
      const int band = 16 * mg + 4 * mm + lk;
      const bool ok = u < nunit && g < ge && band < nb;
      vb[mm] = ok ? Qb[(int64_t)band * ng + g] : czero();
      vq[mm] = ok ? Q[(int64_t)band * ng + g] : czero();
    }
  };
  load(0, cb, cq);
  d4 re[T], im[T];
#pragma unroll
  for (int J = 0; J < T; ++J) { re[J] = d4zero(); im[J] = d4zero(); }
  for (int u = 0; u < nunit; ++u) {
    double2 nbv[4], nqv[4];
    load(u + 1, nbv, nqv);
    const int mg = u % T;
#pragma unroll
    for (int mm = 0; mm < 4; ++mm) {
      const int k = 16 * mg + 4 * mm + lk;
      const double2 vb = cb[mm], vq = cq[mm];
#pragma unroll
      for (int J = 0; J < T; ++J) {
        if (J <= mg) {  // A[i][j] = 0 for j > i
          const double2 a = sa[k * NBP + 16 * J + li];
          re[J] = mfma4(vb.x, a.x, re[J]);
          re[J] = mfma4(-vb.y, a.y, re[J]);
          im[J] = mfma4(vb.x, a.y, im[J]);
          im[J] = mfma4(vb.y, a.x, im[J]);
        }
        const double2 b = sb[k * NBP + 16 * J + li];
        re[J] = mfma4(vq.x, b.x, re[J]);
        re[J] = mfma4(-vq.y, b.y, re[J]);
        im[J] = mfma4(vq.x, b.y, im[J]);
        im[J] = mfma4(vq.y, b.x, im[J]);
      }
    }
    if (mg == T - 1) {
      const int g0 = gw + 64 * (u / T);
#pragma unroll
      for (int J = 0; J < T; ++J) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = g0 + 4 * r + lk, col = 16 * J + li;
          if (row < ge && col < nb) {
            Dr[(int64_t)row * nb + col] = re[J][r];
            Di[(int64_t)row * nb + col] = im[J][r];
          }
        }
        re[J] = d4zero();
        im[J] = d4zero();
      }
    }
#pragma unroll
    for (int mm = 0; mm < 4; ++mm) { cb[mm] = nbv[mm]; cq[mm] = nqv[mm]; }
  }
}

size_t qr_lds_bytes(int nb) {
  const size_t nbp = 16 * (size_t)((nb + 15) / 16), z = sizeof(double2);
  if (nb <= kNarrowBands)  // Gram (pre + sums), apply_bwd, chol_inv
    return std::max(2 * nbp * nbp * z, 2 * (size_t)nb * (nb + 1) * z);
  const size_t gram = 2 * 2 * (size_t)wide_slab<false>() * (nbp + 1) * z;
  const size_t fwd = (size_t)kWideK * nbp * z;
  const size_t bwd = 2 * (size_t)kWideKB * nbp * z;
  const size_t chol = ((size_t)nb * (nb + 1) / 2 + 2 * nb) * z +
                      (nb + kCholWideBlock / 64) * sizeof(double);
  return std::max(std::max(gram, fwd), std::max(bwd, chol));
}

int launch_apply_bwd(const double2* qbar, const double2* q, const double2* A,
                     const double2* Bm, double* dwr, double* dwi, int64_t P,
                     int ng, int nb, hipStream_t s) {
  const int T = (nb + 15) / 16;
  if (nb > kNarrowBands) {
    const size_t shm = 2 * (size_t)kWideKB * 16 * T * sizeof(double2);
    dim3 grid(ceil_div(ng, kWideRows), (unsigned)P), block(kQrBlock);
    switch (T) {
      case 5: hipLaunchKernelGGL((apply_bwd_wide_kernel<5>), grid, block, shm, s, qbar, q, A, Bm, dwr, dwi, ng, nb); break;
      case 6: hipLaunchKernelGGL((apply_bwd_wide_kernel<6>), grid, block, shm, s, qbar, q, A, Bm, dwr, dwi, ng, nb); break;
      case 7: hipLaunchKernelGGL((apply_bwd_wide_kernel<7>), grid, block, shm, s, qbar, q, A, Bm, dwr, dwi, ng, nb); break;
      case 8: hipLaunchKernelGGL((apply_bwd_wide_kernel<8>), grid, block, shm, s, qbar, q, A, Bm, dwr, dwi, ng, nb); break;
      default:
        set_error("nb above 128 is not supported by this build");
        return JRX_ERR_UNSUPPORTED;
    }
    JRX_LAUNCH_CHECK();
    return JRX_OK;
  }
  size_t shm = 2 * (size_t)(16 * T) * (16 * T) * sizeof(double2);
  const int rows = apply_rows(ng, P);
  dim3 grid(ceil_div(ng, rows), (unsigned)P), block(kQrBlock);
  switch (T) {
    case 1: hipLaunchKernelGGL((apply_bwd_kernel<1>), grid, block, shm, s, qbar, q, A, Bm, dwr, dwi, ng, nb, rows); break;
    case 2: hipLaunchKernelGGL((apply_bwd_kernel<2>), grid, block, shm, s, qbar, q, A, Bm, dwr, dwi, ng, nb, rows); break;
    case 3: hipLaunchKernelGGL((apply_bwd_kernel<3>), grid, block, shm, s, qbar, q, A, Bm, dwr, dwi, ng, nb, rows); break;
    case 4: hipLaunchKernelGGL((apply_bwd_kernel<4>), grid, block, shm, s, qbar, q, A, Bm, dwr, dwi, ng, nb, rows); break;
    default:
      set_error("nb above 64 is not supported by this build");
      return JRX_ERR_UNSUPPORTED;
  }
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

}  // namespace jrx
