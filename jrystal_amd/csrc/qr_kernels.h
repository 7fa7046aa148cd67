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
// Up to kNarrowBands bands one wave holds every 16 x 16 tile of a problem
// in registers; above it (up to kMaxBands) the wide kernels stage row
// slabs or band chunks in LDS and spread the tiles over the waves.
constexpr int kNarrowBands = 64;
constexpr int kMaxBands = 128;

struct GramSplit {
  int rows_per_chunk;  // multiple of kQrRows
  int nchunk;          // chunks per problem
};
GramSplit gram_split(int64_t P, int ng, int nb);

// Largest LDS allocation (bytes per workgroup) of the QR kernels for nb.
size_t qr_lds_bytes(int nb);

// partial[p][c] = sum over the rows of chunk c of conj(a_g)^T a_g,
// a_g = W[p][g][:] (times pre[p] on the right when pre != nullptr): one
// partial per chunk (the block's four waves are summed in LDS).
// With skip, problems p with skip[p] != 0 are left out (their partials are
// not written).  Above kNarrowBands with pre != nullptr, A = W pre is first
// written band-major into q1 ([P][nb][ng] complex scratch).
int launch_gram_gmajor(const double* wr, const double* wi, const double2* pre,
                       double2* partial, int64_t P, int ng, int nb,
                       const GramSplit& sp, hipStream_t s,
                       const int* skip = nullptr, double2* q1 = nullptr);

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
// flag_out (optional): flag_out[p] = 1 when S is close enough to diagonal
// for one CholeskyQR pass to be accurate, namely
//   sum_{i != j} |S_ij|^2 / (S_ii S_jj) < kQrOnePass^2,
// i.e. ||D S D - I||_F < kQrOnePass with D = diag(S)^-1/2, which bounds
// cond(D S D) by (1 + x) / (1 - x) = 3 and cond(W D) by sqrt(3); else 0.
// flag_in (optional): problems with flag_in[p] != 0 get R^{-1} = I.
// lmul (optional, nb <= 32): rinv = lmul R^{-1} for upper-triangular lmul
// [P][nb][nb] (trmm_uu fused into the factorisation).
constexpr double kQrOnePass = 0.5;
int launch_chol_inv(const double2* S, double2* rinv, int64_t P, int nb,
                    hipStream_t s, int* flag_out = nullptr,
                    const int* flag_in = nullptr,
                    const double2* lmul = nullptr);

// c = a * b for upper-triangular a, b.
int launch_trmm_uu(const double2* a, const double2* b, double2* c, int64_t P,
                   int nb, hipStream_t s);

// Q[p][b][g] = sum_i W[p][g][i] rinv[p][i][b] (rinv upper triangular);
// problems with skip[p] != 0 are left out.
int launch_apply_fwd(const double* wr, const double* wi, const double2* rinv,
                     double2* q, int64_t P, int ng, int nb, hipStream_t s,
                     const int* skip = nullptr);

// From G = Q^H Qbar: A = Rinv^H, Bm = copyltu(triu(-G)) Rinv^H.
int launch_qr_bwd_small(const double2* G, const double2* rinv, double2* A,
                        double2* Bm, int64_t P, int nb, hipStream_t s);

// Wbar[p][g][j] = sum_i Qbar[p][i][g] A[i][j] + Q[p][i][g] Bm[i][j]; planar
// outputs.
int launch_apply_bwd(const double2* qbar, const double2* q, const double2* A,
                     const double2* Bm, double* dwr, double* dwi, int64_t P,
                     int ng, int nb, hipStream_t s);

// Rayleigh-Ritz rotation (ritz.hip): V[p] = eigenvectors, ascending, of
// the Hermitised (S[p] / (2 f)) on each run of >= 2 consecutive bands with
// equal occupation f > 0, identity elsewhere; whole: one run of all nb
// bands (every f > 0).  V and evals (the ascending eigenvalues at each
// run's bands) are optional.  scratch: ritz_scratch_bytes(P, nb), 256-B
// aligned.
size_t ritz_scratch_bytes(int64_t P, int nb);
int launch_ritz(const double2* S, const double* occ, int64_t P, int nb,
                double2* V, double* evals, bool whole, void* scratch,
                hipStream_t s);
// H[p]_ij = q_i^H H q_j = (S_ij / (2 f_j) + conj(S_ji) / (2 f_i)) / 2.
int launch_subspace_h(const double2* S, const double* occ, int64_t P, int nb,
                      double2* H, hipStream_t s);

}  // namespace jrx
