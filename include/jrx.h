/* jrx.h -- C ABI of libjrx_hip.so, the MI355X (gfx950) implementation of
 * jrystal's direct-minimisation total-energy path.
 *
 * Conventions
 *   - Every array argument is a DEVICE pointer owned by the caller, except
 *     where marked "host".  Complex arrays are interleaved (re, im) float64.
 *   - Every call is asynchronous on the given HIP stream (hipStream_t passed
 *     as void*; NULL is the default stream).  Only jrx_plan_read_timing()
 *     and the one-off set-up calls (plan create/destroy) synchronise.
 *   - Return value 0 (JRX_OK) on success, otherwise a JRX_ERR_* code and a
 *     message retrievable with jrx_last_error() (per host thread).
 *   - A plan is not thread-safe; separate plans are independent.
 */
#ifndef JRX_H
#define JRX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  JRX_OK = 0,
  JRX_ERR_INVALID = 1,     /* bad shape / argument (Python: ValueError) */
  JRX_ERR_UNSUPPORTED = 2, /* size outside what the kernels handle       */
  JRX_ERR_HIP = 3,         /* HIP runtime error (Python: RuntimeError)   */
  JRX_ERR_NOMEM = 4
};

enum { JRX_XC_LDA_X = 0 };

const char* jrx_last_error(void);
const char* jrx_version(void);

/* ------------------------------------------------------------------------
 * Operator boundary of the reference: ifftn3d / fftn3d
 * (jrystal/_src/spmd/fft.py:37-64, single-device lowering
 * custom_partitioning.py:157-160 = jnp.fft.ifftn / fftn over the last three
 * axes).  x, y: (batch, nx, ny, nz) complex128; y = scale * transform(x),
 * unnormalised transform (pass scale = 1/(nx*ny*nz) for numpy's ifftn).
 * In-place (x == y) is allowed.  Sizes must be 7-smooth and <= 2048.
 * ---------------------------------------------------------------------- */
int jrx_fftn3d(const double* x, double* y, int64_t batch, int nx, int ny,
               int nz, double scale, void* stream);
int jrx_ifftn3d(const double* x, double* y, int64_t batch, int nx, int ny,
                int nz, double scale, void* stream);

/* ------------------------------------------------------------------------
 * Plan of the fused evaluation of the direct-minimisation objective
 *   E(w) = E_kin + E_H + E_ext + E_xc
 * replacing the jitted closure total_energy / value_and_grad of
 * jrystal/calc/calc_ground_state_energy_all_electrons.py:119-137,175-181
 * (pw.coeff -> pw.density_grid(_reciprocal) -> energy.kinetic / hartree /
 * external / xc_energy) for fixed occupations.
 * ---------------------------------------------------------------------- */
typedef struct jrx_plan jrx_plan;

typedef struct {
  int ns, nk, nb;          /* spins (1 or 2), k-points, bands (<= 64)      */
  int nx, ny, nz;          /* FFT box, 7-smooth                             */
  int natom;
  const uint8_t* mask;     /* host [nx*ny*nz], C order, nonzero = G kept    */
  const double* kpts;      /* host [nk*3], cartesian, 1/bohr                */
  const double* cell;      /* host [9], lattice vectors as rows, bohr       */
  const double* positions; /* host [natom*3], bohr                          */
  const double* charges;   /* host [natom]                                  */
  const double* occupation;/* host [ns*nk*nb]                               */
  int xc;                  /* JRX_XC_LDA_X                                  */
  int chunk_states;        /* states per FFT chunk, 0 = automatic           */
} jrx_problem;

int jrx_plan_create(jrx_plan** plan, const jrx_problem* problem);
int jrx_plan_destroy(jrx_plan* plan);
/* Workspace the caller passes to jrx_energy_grad (256-byte aligned). */
int jrx_plan_workspace_bytes(const jrx_plan* plan, size_t* bytes);
typedef struct {
  int ng;             /* number of mask points (G_n)                        */
  int active_planes;  /* x-planes of the box that hold mask points          */
  int chunk_states;   /* states per FFT chunk                               */
  size_t workspace_bytes;
} jrx_plan_info;
int jrx_plan_get_info(const jrx_plan* plan, jrx_plan_info* info);

/* One evaluation.  w_re, w_im: [ns][nk][ng][nb] (pw.param_init layout,
 * pw.py:29-91).  energies: device [5] = kinetic, hartree, external, xc,
 * total.  dw_re, dw_im: same layout as w, dE/dw_re and dE/dw_im (both NULL
 * for an energy-only evaluation).  rho: device [ns][nx][ny][nz] density
 * (pw.density_grid), may be NULL.  q: device [ns][nk][nb][ng] complex,
 * the orthonormal coefficients (pw.coeff before the box scatter), may be
 * NULL. */
int jrx_energy_grad(jrx_plan* plan, const double* w_re, const double* w_im,
                    double* energies, double* dw_re, double* dw_im,
                    double* rho, double* q, void* workspace,
                    size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Per-kernel timing of jrx_energy_grad with HIP events recorded on the
 * call's stream (no counterpart in the reference, which has no profiling;
 * used by bench.py for the roofline of the dominant kernel).
 * ---------------------------------------------------------------------- */
enum {
  JRX_K_QR_GRAM = 0,      /* Gram matrices of CholeskyQR2                  */
  JRX_K_QR_SMALL,         /* partial reductions, Cholesky, small products  */
  JRX_K_QR_APPLY,         /* Q = W R^-1                                    */
  JRX_K_PLANE_IFFT,       /* scatter + 2D inverse FFT of the x-planes      */
  JRX_K_XLINE_DENSITY,    /* x inverse FFT + occupation-weighted |psi|^2   */
  JRX_K_POTENTIAL,        /* density FFTs, energies, v(r)                  */
  JRX_K_XLINE_VPSI,       /* x FFTs around v(r) psi(r)                     */
  JRX_K_PLANE_FFT_GATHER, /* 2D forward FFT + gather + kinetic term        */
  JRX_K_QR_BWD_GRAM,      /* Q^H Qbar                                      */
  JRX_K_QR_BWD_APPLY,     /* Wbar = Qbar R^-H + Q copyltu(X) R^-H          */
  JRX_K_OTHER,
  JRX_K_COUNT
};
int jrx_plan_set_timing(jrx_plan* plan, int enable);
/* Synchronises on the recorded events, writes the summed milliseconds and
 * launch counts per class since the last read, and resets them. */
int jrx_plan_read_timing(jrx_plan* plan, double* ms, int* launches);

#ifdef __cplusplus
}
#endif

#endif /* JRX_H */
