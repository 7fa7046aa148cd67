// Batched QR orthonormalisation (CholeskyQR2) and its adjoint.
//
// A batch holds P problems; problem p is an (ng x nb) complex matrix.
// W arrives g-major and planar (w_re, w_im: [p][g][b], the reference's
// params layout); Q and Qbar are band-major interleaved complex
// ([p][b][g]), the layout the FFT passes read.  Small nb x nb matrices are
// row-major complex.
#pragma once
#include "common.h"

namespace jrx {

constexpr int kQrBlock = 256;
constexpr int kQrRows = 64;   // rows per LDS tile
constexpr int kMaxBands = 64;

struct GramSplit {
  int rows_per_chunk;  // multiple of kQrRows
  int nchunk;          // chunks per problem
};
GramSplit gram_split(int64_t P, int ng);

// partial[p][c] = sum over the rows of chunk c of conj(a_g)^T a_g,
// a_g = W[p][g][:] (times pre[p] on the right when pre != nullptr): one
// partial per chunk (the block's four waves are summed in LDS).
int launch_gram_gmajor(const double* wr, const double* wi, const double2* pre,
                       double2* partial, int64_t P, int ng, int nb,
                       const GramSplit& sp, hipStream_t s);

// partial[p][c] = sum over the rows of chunk c of
// conj(Q[p][:][g]) (x) Qb[p][:][g] (one partial per chunk).
int launch_gram_bmajor(const double2* q, const double2* qb, double2* partial,
                       int64_t P, int ng, int nb, const GramSplit& sp,
                       hipStream_t s);

// out[p] = sum_c partial[p][c] (fixed order).
int launch_reduce_partials(const double2* partial, double2* out, int64_t P,
                           int nchunk, int nb, hipStream_t s);

// S = R^H R (R upper, real positive diagonal); writes R^{-1}.  A
// non-positive pivot turns the result into NaN.
int launch_chol_inv(const double2* S, double2* rinv, int64_t P, int nb,
                    hipStream_t s);

// c = a * b for upper-triangular a, b.
int launch_trmm_uu(const double2* a, const double2* b, double2* c, int64_t P,
                   int nb, hipStream_t s);

// Q[p][b][g] = sum_i W[p][g][i] rinv[p][i][b].
int launch_apply_fwd(const double* wr, const double* wi, const double2* rinv,
                     double2* q, int64_t P, int ng, int nb, hipStream_t s);

// From G = Q^H Qbar: A = Rinv^H, Bm = copyltu(triu(-G)) Rinv^H.
int launch_qr_bwd_small(const double2* G, const double2* rinv, double2* A,
                        double2* Bm, int64_t P, int nb, hipStream_t s);

// Wbar[p][g][j] = sum_i Qbar[p][i][g] A[i][j] + Q[p][i][g] Bm[i][j]; planar
// outputs.
int launch_apply_bwd(const double2* qbar, const double2* q, const double2* A,
                     const double2* Bm, double* dwr, double* dwi, int64_t P,
                     int ng, int nb, hipStream_t s);

}  // namespace jrx
