// Host-side launchers of the reciprocal/real-space energy kernels.
#pragma once
#include "common.h"

namespace jrx {

constexpr int kRedBlock = 256;

struct CellParams {
  int nx, ny, nz;
  double B[9];       // reciprocal basis rows, B = 2 pi A^{-T}
  double vol;
  int natom;
  const double* pos;     // device [natom*3]
  const double* charge;  // device [natom]
};

// out[i] = (sum_s rho[s*N + i], 0)
int launch_rho_to_complex(const double* rho, double2* out, int ns, int64_t N,
                          hipStream_t s);

// From rho_rec = fftn(rho_tot) build the reciprocal potential
//   vrec(G) = (4 pi vol / N) rho_rec / |G|^2 - 4 pi conj(S(G)) / (|G|^2+1e-10)
// (zero at G = 0; S(G) = sum_a Z_a e^{i G.R_a}) and per-block partials of
//   E_H   = (2 pi vol / N^2) sum_{G!=0} |rho_rec|^2 / |G|^2
//   E_ext = -(4 pi / N) sum_{G!=0} Re(S rho_rec) / (|G|^2 + 1e-10).
// vrec may alias rho_rec.  Returns the number of partials written.
// sterm (optional): S(G) / (|G|^2 + 1e-10) per box point, as
// launch_structure_term writes it; without it S(G) is formed per point.
int launch_potential_rec(const double2* rho_rec, double2* vrec,
                         double* eh_part, double* ee_part,
                         const CellParams& cp, int* nparts, hipStream_t s,
                         const double2* sterm = nullptr);

// out[i] = S(G_i) / (|G_i|^2 + 1e-10) (0 at G = 0) over cp's box: the
// geometry-only factor of the external term, formed once per plan.
int launch_structure_term(const CellParams& cp, double2* out, hipStream_t s);

// v_s(r) = Re(vbox(r)) + v_xc,s(r) and partials of E_xc (LDA exchange);
// ks: the spin-polarised v_xc in the Kohn-Sham form of the band driver
// (eps + rho_s d eps / d rho_s instead of d E_xc / d rho_s).
int launch_potential_real(const double2* vbox, const double* rho, double* v,
                          double* exc_part, int ns, int64_t N, double vol,
                          double cx, int* nparts, hipStream_t s,
                          bool ks = false);

// Partials of |G+k|^2 |q|^2 per (state, g-block), band-major q.
int launch_kinetic(const double2* q, const double* occ, const double* kpts,
                   const short* g_m, const CellParams& cp, int64_t nstates,
                   int ng, int nk, int nb, double* kin_part, int* nparts,
                   hipStream_t s);

// tb[st] = sum_j part[st * nper + j]: sum_G |G+k|^2 |q|^2 of each state
// (twice the band kinetic energy).
int launch_state_sums(const double* part, int nper, int64_t S, double* tb,
                      hipStream_t s);

// energies[0..4] = kinetic (1/2 sum occ tb), hartree, external, xc, total
// from the partials (fixed summation order).
int launch_finalize_energies(const double* tb, const double* occ, int64_t S,
                             const double* eh, const double* ee, int nrec,
                             const double* exc, int nexc, double* energies,
                             hipStream_t s);

int launch_fill(double* p, double v, int64_t n, hipStream_t s);

// Exchange-correlation stage beyond LDA exchange (xc.hip), unpolarised.
struct FftLen;
bool xc_is_gga(int flags);
// gb[3][N] = ifftn(i G_d rho_rec) (the density gradient boxes; cp's box)
// (gvec, optional: the G grid [N][3] to use instead of cp's basis)
int xc_stage_gradients(const double2* rho_rec, double2* gb,
                       const CellParams& cp, const FftLen& L0,
                       const FftLen& L1, const FftLen& L2, hipStream_t s,
                       const double* gvec = nullptr);
// v = Re(vbox) (vbox may be NULL) + dE_xc/drho and partials of E_xc for
// the JRX_XC_* terms in flags; for GGA terms gb holds the gradient boxes on
// entry and is used as scratch.
int xc_stage_potential(const double2* vbox, const double* rho, double* v,
                       double2* gb, double* exc_part, int* nparts,
                       const CellParams& cp, int flags, const FftLen& L0,
                       const FftLen& L1, const FftLen& L2, hipStream_t s,
                       const double* gvec = nullptr);

// Twiddle tables of the plan-less FFT entry points (a per-device cache).
int cached_fftlen(int n, FftLen* L);

}  // namespace jrx
