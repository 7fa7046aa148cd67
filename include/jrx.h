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

/* Exchange-correlation terms of jrx_problem.xc, OR-ed (the reference's
 * '+'-joined xc strings, jrystal/_src/xc.py:222-253); 0 means lda_x.
 * Terms other than lda_x need an unpolarised problem (ns = 1). */
enum {
  JRX_XC_LDA_X = 1,     /* Dirac exchange                                 */
  JRX_XC_LDA_C_PW = 2,  /* Perdew-Wang 1992 correlation                   */
  JRX_XC_GGA_X_PBE = 4, /* PBE exchange                                   */
  JRX_XC_GGA_C_PBE = 8  /* PBE correlation                                */
};

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
  int xc;                  /* JRX_XC_* terms OR-ed; 0 = lda_x              */
  int chunk_states;        /* states per FFT chunk, 0 = automatic           */
  int fft_variant;         /* 0 = automatic (row buffer for cubic 64^3 and
                              128^3 boxes, else plane buffer), 1 = plane
                              buffer with runtime-radix kernels only, 2 =
                              plane buffer (register 64-point kernels where
                              they apply)                                   */
  int xline_blocks;        /* target workgroups of an x-line pass, 0 = 4096 */
  int plane_pad;           /* extra elements per x-plane of the plane
                              buffer, -1 = automatic                        */
  int qr_passes;           /* 0 = adaptive CholeskyQR: one pass for the
                              k-points whose Gram matrix is close to
                              diagonal (||D S D - I||_F < 0.5), two for the
                              rest; 2 = always two passes                   */
} jrx_problem;

int jrx_plan_create(jrx_plan** plan, const jrx_problem* problem);
int jrx_plan_destroy(jrx_plan* plan);
/* Workspace the caller passes to jrx_energy_grad (256-byte aligned). */
int jrx_plan_workspace_bytes(const jrx_plan* plan, size_t* bytes);
typedef struct {
  int ng;             /* number of mask points (G_n)                        */
  int active_planes;  /* x-planes of the box that hold mask points          */
  int chunk_states;   /* states per FFT chunk                               */
  int keep_planes;    /* 1: the adjoint sweep reuses the forward planes     */
  size_t workspace_bytes;
  int row_path;       /* 1: row-buffer FFT path (cubic 64^3 / 128^3 boxes)  */
  int rows;           /* (x, y) lines along z that hold mask points         */
} jrx_plan_info;
int jrx_plan_get_info(const jrx_plan* plan, jrx_plan_info* info);
/* Number of float64 values in w_re (= in w_im): ns*nk*ng*nb. */
int jrx_plan_param_count(const jrx_plan* plan, int64_t* count);

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

/* No counterpart in the reference.  jrx_energy_grad for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same kernels, with the workspace
 * taken from a static device array of JRX_STATIC_WORKSPACE_BYTES that each
 * GPU the library runs on holds until the process ends.  The call allocates
 * nothing; the plan itself is still made by jrx_plan_create.  Plans whose
 * jrx_plan_workspace_bytes exceeds the limit are rejected with
 * JRX_ERR_UNSUPPORTED (the 2-atom 32^3 / 8 k / 8-band cell needs about
 * 23 MB).  Every call on one GPU shares that GPU's array: calls that overlap
 * on one GPU (two streams, two host threads) race, and the array keeps the
 * last call's contents, so the plan-level readers of a workspace
 * (jrx_plan_band_norms, jrx_plan_precondition, ...) do not apply to it.
 * Per-kernel timing (jrx_plan_set_timing) is not recorded by this call. */
#define JRX_STATIC_WORKSPACE_BYTES ((size_t)64 << 20)
int jrx_energy_grad_static(jrx_plan* plan, const double* w_re,
                           const double* w_im, double* energies,
                           double* dw_re, double* dw_im, double* rho,
                           double* q, void* stream);

/* Preconditioned minimisation support (no counterpart in the reference,
 * whose driver runs optax Adam on the raw parameters).  After
 * jrx_energy_grad the workspace holds, per state, tb = sum_G |G+k|^2 |q|^2
 * (twice the band kinetic energy).
 * jrx_plan_precondition: out = K(|G+k|^2 / tb) * g (Teter-Payne-Allan
 *   1989), times factor[i] when factor is not NULL; g, out, factor in the
 *   [s][k][g][b] parameter layout.
 * jrx_plan_band_norms: copies tb [s][k][b] to out.
 * jrx_plan_orthonormal_params: writes Q of the last evaluation (W = QR,
 *   R upper triangular with positive diagonal) in the [s][k][g][b]
 *   parameter layout; the energy is invariant under W -> Q, so a minimiser
 *   may reset its parameters to Q to undo drift of the gauge R. */
int jrx_plan_precondition(jrx_plan* plan, const double* g_re,
                          const double* g_im, const double* factor,
                          double* out_re, double* out_im, void* workspace,
                          size_t workspace_bytes, void* stream);
int jrx_plan_band_norms(const jrx_plan* plan, double* out,
                        const void* workspace, size_t workspace_bytes,
                        void* stream);

/* Band energies: with the switch on, every energy+gradient evaluation also
 * forms dE/df[s,k,b] = <psi|H|psi> = 1/2 sum_G |G+k|^2 |Q_Gb|^2 +
 * (Omega/N) sum_r v(r) |psi_b(r)|^2, v = v_H + v_ext + v_xc, for every band
 * (occupied or not); jrx_plan_band_energies writes the last one to out[S]
 * (device).  The gradient of the energy with respect to the occupations of
 * the reference's parameterised occupations (occupation.py:261-278, used by
 * calc_ground_state_energy_all_electrons.py:139-181) needs it. */
int jrx_plan_set_band_energies(jrx_plan* plan, int on);
int jrx_plan_band_energies(const jrx_plan* plan, double* out,
                           const void* workspace, size_t workspace_bytes,
                           void* stream);
/* Replaces the plan's occupation [S] (device or host memory, stream-ordered
 * copy): the occupations of the reference's parameterised occupation methods
 * change every optimiser step. */
int jrx_plan_set_occupation(jrx_plan* plan, const double* occupation,
                            void* stream);
int jrx_plan_orthonormal_params(const jrx_plan* plan, double* out_re,
                                double* out_im, const void* workspace,
                                size_t workspace_bytes, void* stream);
/* jrx_plan_subspace_matrix: after jrx_energy_grad with a gradient, copies
 *   Q^H Qbar [s][k][i][j] (complex, nb x nb per (s,k)); Qbar_j = 2 f_j H q_j,
 *   so the block of bands with equal occupation f > 0 is 2 f times the
 *   subspace Hamiltonian q_i^H H q_j.
 * jrx_plan_rotate: out = Q V in the parameter layout, V [s][k][c][b]
 *   complex nb x nb.  A unitary V mixing only bands of equal occupation
 *   leaves the energy unchanged (Rayleigh-Ritz gauge of a minimiser). */
int jrx_plan_subspace_matrix(const jrx_plan* plan, double* out,
                             const void* workspace, size_t workspace_bytes,
                             void* stream);
int jrx_plan_rotate(const jrx_plan* plan, const double* v, double* out_re,
                    double* out_im, const void* workspace,
                    size_t workspace_bytes, void* stream);
/* jrx_plan_ritz_rotation: after jrx_energy_grad with a gradient, v [s][k]
 *   (complex nb x nb) = the Rayleigh-Ritz rotation of a gauge reset: on each
 *   run of >= 2 consecutive bands with equal occupation f > 0, the
 *   eigenvectors (ascending eigenvalues) of the Hermitised subspace
 *   Hamiltonian (Q^H Qbar) / (2 f) over the run, by cyclic Jacobi on the
 *   device; the identity elsewhere.  No counterpart in the reference (its
 *   minimiser is Adam); replaces a host eigendecomposition. */
int jrx_plan_ritz_rotation(const jrx_plan* plan, double* v, void* workspace,
                           size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Kohn-Sham band mode (the band-structure driver's objective).
 *
 * jrx_plan_set_fixed_density: rho [s][x][y][z] (device, the plan's ns
 *   channels) -> the effective potential v_eff = V_H + V_ext + v_xc of that
 *   density (Kohn-Sham form: full Hartree potential, v_xc = d(rho e_xc) /
 *   d rho), kept by the plan; from then on jrx_energy_grad evaluates
 *   E = sum_{s,k,b} f <psi|T + v_eff|psi> (energies = {kinetic part,
 *   potential part, 0, 0, E}) and its gradient, with no density pass.
 *   rho = NULL returns the plan to the ground-state objective.  The first
 *   call allocates the potential (ns * nx * ny * nz doubles).  Replaces
 *   hamiltonian_matrix_trace (jrystal/_src/hamiltonian.py:105-168) with the
 *   potential of potential.effective (potential.py:203-279, kohn_sham=True).
 * jrx_plan_set_kpoints: kpts [nk][3] (host or device, 1/bohr) replace the
 *   plan's k-points for later evaluations (the k-path loop of
 *   jrystal/calc/calc_band_structure_all_electrons.py:120-152).
 * jrx_plan_subspace_hamiltonian: after an evaluation (with or without a
 *   gradient in band mode; with a gradient otherwise), h [s][k] (complex
 *   nb x nb) = q_i^H H q_j, H = T + v of the evaluation; NaN where an
 *   occupation is <= 0.  Replaces hamiltonian_matrix (hamiltonian.py:171-240).
 * jrx_plan_subspace_eigenvalues: the same matrix's eigenvalues, ascending,
 *   evals [s][k][nb], and optionally its eigenvectors vecs [s][k] (complex
 *   nb x nb, columns), by cyclic Jacobi on the device; NaN eigenvalues for
 *   a (s,k) with an occupation <= 0.  Replaces jnp.linalg.eigvalsh of
 *   calc_band_structure_all_electrons.py:156-169. */
int jrx_plan_set_fixed_density(jrx_plan* plan, const double* rho,
                               void* workspace, size_t workspace_bytes,
                               void* stream);
int jrx_plan_set_kpoints(jrx_plan* plan, const double* kpts, void* stream);
int jrx_plan_subspace_hamiltonian(const jrx_plan* plan, double* h,
                                  const void* workspace,
                                  size_t workspace_bytes, void* stream);
int jrx_plan_subspace_eigenvalues(const jrx_plan* plan, double* evals,
                                  double* vecs, void* workspace,
                                  size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Norm-conserving pseudopotentials.
 *
 * jrx_plan_set_external_potential: v_rec [x][y][z] (device, complex) is a
 *   local potential on the G box in the normalisation of
 *   external_reciprocal (potential.py:121-166), i.e. the reference's
 *   potential_local_reciprocal (pseudopotential/local.py:34-112); it takes
 *   the place of the Coulomb nuclei in the external energy and potential
 *   (energy_local, local.py:158-187).  v_rec = NULL restores the nuclei.
 * jrx_plan_set_nonlocal: proj [nk][nproj][ng] (complex; host or device) are
 *   the non-local projectors at the mask points (the mask restriction of
 *   potential_nonlocal_psi_reciprocal, nloc.py:58-150).  Later evaluations
 *   add E_nl = sum_{s,k,b} f sum_p |F_pb|^2 / vol, F = sum_g proj_pg q_gb
 *   (energy_nonlocal, nloc.py:196-227) to energies[2] (external) and the
 *   total, and its gradient; in band mode the projectors act in the
 *   Hamiltonian as well.  nproj = 0 switches them off.  nproj * nb <= 4096
 *   and nb <= 128.  The first call allocates the projector table and its
 *   scratch (plan-owned: evaluations of one plan with projectors must not
 *   overlap on different streams).
 * jrx_plan_nonlocal_energies: out [s][k][b] = sum_p |F_pb|^2 / vol of the
 *   last evaluation (per band, without the occupation). */
int jrx_plan_set_external_potential(jrx_plan* plan, const double* v_rec,
                                    void* stream);
int jrx_plan_set_nonlocal(jrx_plan* plan, int nproj, const double* proj,
                          void* stream);
int jrx_plan_nonlocal_energies(const jrx_plan* plan, double* out,
                               void* stream);

/* ------------------------------------------------------------------------
 * Learned preconditioner and warm start (no counterpart in the reference).
 * A FP32 MLP, 8 feature slots -> tanh 64 -> tanh 64 -> 2 outputs, with
 * flat parameters (jrx_mlp_num_params() floats: W1[64][8], b1[64],
 * W2[64][64], b2[64], W3[2][64], b3[2]) and norm[16] = (mu[8], 1/sd[8])
 * applied to the raw features first.  Raw features of coefficient
 * (s,k,g,b): log1p(|G+k|^2), log1p(|G+k|^2 / tb), log(tb),
 * log(rho_mean), log1p(rho_var / rho_mean^2), log1p(gap), 0, 0 with tb the
 * band norm, rho_mean / rho_var the statistics of the total density of the
 * last evaluation and gap[s][k][b] a caller-given per-band value (the
 * band's energy below the highest occupied band; NULL = 0).  Output 0 is
 * the log preconditioner factor (clamped to [-8, 8]), output 1 the log
 * warm-start amplitude (clamped to [-30, 8]).
 * ---------------------------------------------------------------------- */
/* out[0] = mean, out[1] = variance of the total density over the grid. */
int jrx_plan_density_stats(const jrx_plan* plan, double* out,
                           void* workspace, size_t workspace_bytes,
                           void* stream);
/* x[i][0..7] = raw features of coefficient idx[i] (device int64 flat
 * [s][k][g][b] indices; out-of-range indices give zeros). */
int jrx_plan_features(const jrx_plan* plan, const int64_t* idx, int64_t n,
                      const double* stats, const double* gap, float* x,
                      const void* workspace, size_t workspace_bytes,
                      void* stream);
/* Inference on every coefficient: factor = exp(y0) (if factor is not
 * NULL), w *= exp(y1) (if w_re / w_im are not NULL). */
int jrx_plan_mlp_apply(const jrx_plan* plan, const float* params,
                       const float* norm, const double* stats,
                       const double* gap, double* factor, double* w_re,
                       double* w_im,
                       const void* workspace, size_t workspace_bytes,
                       void* stream);
/* Plain network kernels (training).  x [n][8] raw features, y [n][2]. */
size_t jrx_mlp_num_params(void);
int jrx_mlp_forward(const float* params, const float* norm, const float* x,
                    int64_t n, float* y, void* stream);
/* grad (+)= d(sum_i dy_i . y_i)/dparams (accumulate != 0 adds). */
size_t jrx_mlp_backward_workspace_bytes(int64_t n);
int jrx_mlp_backward(const float* params, const float* norm, const float* x,
                     const float* dy, int64_t n, float* grad, int accumulate,
                     void* workspace, size_t workspace_bytes, void* stream);
/* Snapshot loss: nsnap groups of snap_size samples; with p = exp(y0):
 * loss[s] = 1 - (sum p a)^2 / (t2[s] sum p^2 b); dy = d(mean_s loss)/dy. */
int jrx_mlp_cosine_loss(const float* y, const float* a, const float* b,
                        const float* t2, int64_t snap_size, int64_t nsnap,
                        float* dy, float* loss, void* stream);
/* loss[0] = mean (y1 - target)^2; dy its gradient (output 0 gets 0). */
size_t jrx_mlp_mse_workspace_bytes(int64_t n);
int jrx_mlp_mse_loss(const float* y, const float* target, int64_t n,
                     float* dy, float* loss, void* workspace,
                     size_t workspace_bytes, void* stream);
/* Adam step (optax rule, FP32); count = steps taken including this one. */
int jrx_mlp_adam(float* params, const float* grad, float* m, float* v,
                 int64_t n, double lr, double b1, double b2, double eps,
                 int64_t count, void* stream);
/* Minimiser support (no counterpart in the reference): one pass of the CG
 * update, h1 = K * factor * g1 (as jrx_plan_precondition) and
 * out[0..2] = (g1.h1, g0.h1, g1.d) summed over the real and imaginary
 * planes (deterministic).  dots_ws: device buffer of
 * jrx_cg_dots_workspace_bytes(). */
size_t jrx_cg_dots_workspace_bytes(void);
int jrx_plan_precondition_dots(jrx_plan* plan, const double* g1_re,
                               const double* g1_im, const double* g0_re,
                               const double* g0_im, const double* d_re,
                               const double* d_im, const double* factor,
                               double* h1_re, double* h1_im, double* out,
                               void* dots_ws, size_t dots_ws_bytes,
                               const void* workspace, size_t workspace_bytes,
                               void* stream);
/* Flat float64 vector helpers: out[0] = sum a*b (deterministic, device
 * scalar); out = alpha x + beta y (out may be y). */
size_t jrx_dot_workspace_bytes(void);
int jrx_vec_dot(const double* a, const double* b, int64_t n, double* out,
                void* workspace, size_t workspace_bytes, void* stream);
int jrx_vec_axpby(double alpha, const double* x, double beta, const double* y,
                  double* out, int64_t n, void* stream);

/* ------------------------------------------------------------------------
 * Stand-alone operators behind the reference's public functions (full-box
 * layout).  Complex arrays interleaved; "state" = (spin, kpt, band) flattened
 * in that order; N = nx*ny*nz.  Gradients follow the convention
 * dL/dRe z + i dL/dIm z and are scaled by grad_out (device scalar, or
 * per-state array where noted).
 * ---------------------------------------------------------------------- */
/* unitary_module.unitary_matrix (unitary_module.py:30-81) for P problems:
 * q [P][nb][ng] (band-major) orthonormal columns of W = w_re + i w_im
 * ([P][ng][nb]); rinv [P][nb][nb] = R^-1 (upper), kept for the backward. */
size_t jrx_qr_workspace_bytes(int64_t P, int ng, int nb);
int jrx_qr(const double* w_re, const double* w_im, int64_t P, int ng, int nb,
           double* q, double* rinv, void* workspace, size_t workspace_bytes,
           void* stream);
int jrx_qr_backward(const double* q, const double* qbar, const double* rinv,
                    int64_t P, int ng, int nb, double* dw_re, double* dw_im,
                    void* workspace, size_t workspace_bytes, void* stream);
/* utils.expand_coefficient / squeeze_coefficient (utils.py:257-308):
 * q [states][ng] <-> box [states][N]; mask_index [ng] = flat C-order box
 * offsets of the mask points. */
int jrx_expand_coefficient(const double* q, const int64_t* mask_index,
                           double* box, int64_t nstates, int ng, int64_t N,
                           void* stream);
int jrx_squeeze_coefficient(const double* box, const int64_t* mask_index,
                            double* q, int64_t nstates, int ng, int64_t N,
                            void* stream);
/* pw.density_grid after wave_grid (pw.py:273-284): out [ns][N] =
 * sum_{k,b} occ |psi|^2, or out [states][N] = |psi|^2 when occ is NULL. */
int jrx_density(const double* psi, const double* occ, double* out, int ns,
                int64_t states_per_spin, int64_t N, void* stream);
int jrx_density_backward(const double* psi, const double* occ,
                         const double* grad_rho, double* grad_psi, int ns,
                         int64_t states_per_spin, int64_t N, void* stream);
/* energy.kinetic (energy.py:138-182) on the box: gvec [N][3], kpts [nk][3];
 * out[0] = 1/2 sum occ sum_G |G+k|^2 |C|^2 (occ given) or out[state] =
 * 1/2 sum_G |G+k|^2 |C|^2.  grad_out: scalar (occ given) or per state. */
size_t jrx_kinetic_workspace_bytes(int64_t nstates, int64_t N);
int jrx_kinetic_box(const double* coeff, const double* gvec,
                    const double* kpts, const double* occ, int ns, int nk,
                    int nb, int64_t N, double* out, void* workspace,
                    size_t workspace_bytes, void* stream);
int jrx_kinetic_box_backward(const double* coeff, const double* gvec,
                             const double* kpts, const double* occ,
                             const double* grad_out, int ns, int nk, int nb,
                             int64_t N, double* grad_coeff, void* stream);
/* energy.hartree / external / xc_energy (energy.py:29-135,185-211) of a
 * density: rho_rec [ns][N] complex, rho [ns][N] real.  grad (may be NULL)
 * receives dE/drho_rec (complex) or dE/drho (real) times grad_out[0]. */
size_t jrx_reduce_workspace_bytes(int64_t n);
int jrx_hartree(const double* rho_rec, const double* gvec, int ns, int64_t N,
                double vol, int kohn_sham, double* out, double* grad,
                const double* grad_out, void* workspace,
                size_t workspace_bytes, void* stream);
int jrx_external(const double* rho_rec, const double* gvec,
                 const double* positions, const double* charges, int natom,
                 int ns, int64_t N, double* out, double* grad,
                 const double* grad_out, void* workspace,
                 size_t workspace_bytes, void* stream);
int jrx_xc_lda(const double* rho, int ns, int64_t N, double vol, double* out,
               double* grad, const double* grad_out, void* workspace,
               size_t workspace_bytes, void* stream);
/* energy.xc_energy (energy.py:185-211) for the '+'-joined xc strings of
 * xc.py:222-253 (xc_flags: an OR of JRX_XC_* terms) on an unpolarised
 * density rho [nx][ny][nz] with the G grid gvec [nx][ny][nz][3]:
 *   out[0] = vol/N sum_r eps(rho, sigma) rho,
 *   sigma = sum_d |ifftn(i G_d fftn(rho))|^2 (complex gradients, as
 *   sigma_r_fn of xc.py:88-106 forms them),
 * and, when grad != NULL, grad = grad_out[0] * dE_xc/drho (the adjoint of
 * that discretisation).  Workspace: jrx_xc_workspace_bytes, 256-B aligned. */
size_t jrx_xc_workspace_bytes(int nx, int ny, int nz, int xc_flags);
int jrx_xc(const double* rho, const double* gvec, int nx, int ny, int nz,
           double vol, int xc_flags, double* out, double* grad,
           const double* grad_out, void* workspace, size_t workspace_bytes,
           void* stream);
/* Element-wise helpers of the autograd glue: y = (x, 0); y = Re x (n
 * complex); y = alpha[0] * x (alpha a device scalar). */
int jrx_real_to_complex(const double* x, double* y, int64_t n, void* stream);
int jrx_complex_real(const double* x, double* y, int64_t n, void* stream);
int jrx_scale(const double* x, const double* alpha, double* y, int64_t n,
              void* stream);
/* occupation.simplex_projector's projection (occupation.py:295-366) of
 * `channels` spin channels of n entries each (in [c][n]): x = sigmoid(in)
 * when sigmoid != 0, else in; y = the point of {0 <= y <= 1, sum y =
 * totals[c]} reached by one uniform shift and clipping, with the
 * reference's sorted-cumulative-sum rule for the shift.  With ybar != NULL
 * also in_bar = d(sum ybar * y) / d in (active set fixed).  n <= 8192;
 * totals is a device array [channels]. */
int jrx_simplex_projection(const double* in, int64_t n, int channels,
                           const double* totals, int sigmoid, double* y,
                           const double* ybar, double* in_bar, void* stream);
/* One optax.adam step (opt_utils.py:153-168 -> optax.adam) in place on a
 * flat float64 parameter; m, v are the moment buffers, step is 1-based. */
int jrx_adam_update(double* param, const double* grad, double* m, double* v,
                    int64_t n, double lr, double b1, double b2, double eps,
                    double eps_root, int64_t step, void* stream);
/* ewald.ewald_coulomb_repulsion (ewald.py:21-86) with the caller's
 * translation list [nt][3] and G grid [N][3]. */
size_t jrx_ewald_workspace_bytes(int64_t n_translations, int64_t N);
int jrx_ewald(const double* positions, const double* charges, int natom,
              const double* gvec, int64_t N, double vol, double eta,
              const double* translations, int64_t nt, double* out,
              void* workspace, size_t workspace_bytes, void* stream);

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

/* ------------------------------------------------------------------------
 * Batched optimiser step over plans of different shapes (no single
 * counterpart in the reference: its driver steps one structure,
 * calc_ground_state_energy_all_electrons.py:175-220).  One step = for
 * every member, jrx_energy_grad into its gradient buffers, then one grouped
 * optax.adam update (opt_utils.py:153-168 rule) of all members' w_re and
 * w_im with the step count t kept on the device (incremented by the step,
 * so a captured step can be replayed).  Members run on `streams` auxiliary
 * streams (0 = all on the caller's stream) between fork / join events.
 * jrx_batch_capture records one step into a hipGraph; jrx_batch_replay
 * launches it.  All buffers are device buffers owned by the caller.
 * ---------------------------------------------------------------------- */
typedef struct jrx_batch jrx_batch;
typedef struct {
  jrx_plan* plan;
  double *w_re, *w_im;     /* parameters, updated in place                  */
  double *dw_re, *dw_im;   /* gradient buffers                              */
  double *m_re, *m_im;     /* Adam first moments                            */
  double *v_re, *v_im;     /* Adam second moments                           */
  double* energies;        /* [5], written by every step                    */
  void* workspace;         /* jrx_plan_workspace_bytes, 256-byte aligned    */
  size_t workspace_bytes;
} jrx_batch_member;
typedef struct {
  double lr, b1, b2, eps, eps_root;
} jrx_adam_config;
int jrx_batch_create(jrx_batch** batch, const jrx_batch_member* members,
                     int n, const jrx_adam_config* adam, int streams);
int jrx_batch_destroy(jrx_batch* batch);
int jrx_batch_set_step_count(jrx_batch* batch, int64_t count, void* stream);
/* Synchronises `stream`. */
int jrx_batch_get_step_count(const jrx_batch* batch, int64_t* count,
                             void* stream);
int jrx_batch_step(jrx_batch* batch, void* stream);
int jrx_batch_capture(jrx_batch* batch, void* stream);
int jrx_batch_replay(jrx_batch* batch, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* JRX_H */
