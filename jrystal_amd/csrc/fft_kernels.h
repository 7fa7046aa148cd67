// Host-side launchers of the FFT kernels (fft.hip).
#pragma once
#include "fft_core.h"

namespace jrx {

// Geometry of the G-sphere inside the FFT box, built once per plan.
// Mask points are numbered g = 0..ng-1 in C order of the box (x, y, z),
// so the points of one x-plane form a contiguous g-range.
struct BoxGeom {
  int nx, ny, nz, ng;
  int nxa;              // number of x-planes holding mask points
  const int* plane_g0;  // [nxa+1] first g of each active plane
  const int* plane_x;   // [nxa]   x index of each active plane
  const int* g_yz;      // [ng]    y*nz + z of point g
  const short* rows;    // active (y) rows of each plane, concatenated
  const int* row_off;   // [nxa+1] offsets into rows
  int max_rows;         // max active rows over planes
  const int* xmap;      // [nx]    active-plane index of x, or -1
  int ps;               // element stride of the x-planes of the plane
                        // buffer T (ny*nz plus a pad that spreads the planes
                        // read together by the x-line passes over channels)
};

// y = scale * DFT_axis(x) along one axis of a (B, n0, n1, n2) complex box.
int launch_line_fft(const double2* in, double2* out, int64_t B, int n0, int n1,
                    int n2, int axis, bool inverse, double scale,
                    const FftLen& L, hipStream_t s);

// Full 3D transform (three line passes).  In-place allowed.
int fft3d(const double2* in, double2* out, int64_t B, int n0, int n1, int n2,
          bool inverse, double scale, const FftLen& Lx, const FftLen& Ly,
          const FftLen& Lz, hipStream_t s);

// T[c][xi][y][z] = 2D inverse transform over (y,z) of the scattered plane
// x = plane_x[xi] of state (state0 + c); q is band-major [state][ng].
int launch_plane_ifft(const double2* q, double2* T, int64_t state0,
                      int nstates, const BoxGeom& geo, const FftLen& Ly,
                      const FftLen& Lz, hipStream_t s);

// Groups of states per x-line launch (parallelism across states).
constexpr int kMaxXGroups = 32;
int xlines_per_block(int nx);
// State groups of an x-line pass: enough groups to reach about
// target_blocks workgroups (0 = 4096), at most kMaxXGroups and nstates.
int xline_groups(const BoxGeom& geo, int nstates, int target_blocks = 0);

// rho_part[g](x,y,z) += sum_{c in group g} occ[state0+c] |IFFT_x T[c]|^2 /
// vol.  All states of the chunk belong to one spin.
int launch_xline_density(const double2* T, double* rho_part, int groups,
                         const double* occ, int64_t state0, int nstates,
                         double inv_vol, const BoxGeom& geo, const FftLen& Lx,
                         hipStream_t s);

// out[i] = sum_g part[g*N + i] (fixed order).
int launch_reduce_groups(const double* part, double* out, int G, int64_t N,
                         hipStream_t s);

// T[c] <- FFT_x( v * IFFT_x(T[c]) ) restricted to the active planes; v is
// the potential of the chunk's spin.
int launch_xline_vpsi(double2* T, const double* v, int groups, int nstates,
                      const BoxGeom& geo, const FftLen& Lx, hipStream_t s);

// qbar[state][g] = f (2/vol * FFT_yz(T)[g] + |G+k|^2 q[state][g]) and the
// per-(state, plane) partials of |G+k|^2 |q|^2 (summed per state later).
struct GatherParams {
  const double* occ;   // [states]
  const double* kpts;  // [nk*3]
  double B[9];         // reciprocal basis rows
  int nk, nb;          // state -> k index = (state / nb) % nk
  double two_over_vol;
  // optional: pot_partial[state * nxa + plane] = Re sum_g conj(q_g) h_g over
  // the plane's mask points, h = FFT(v psi) (band energies)
  double* pot_partial;
};
int launch_plane_fft_gather(const double2* T, const double2* q, double2* qbar,
                            double* kin_partial, int64_t state0, int nstates,
                            const BoxGeom& geo, const FftLen& Ly,
                            const FftLen& Lz, const GatherParams& gp,
                            hipStream_t s);

// Register-resident 64-point kernels (nx = ny = nz = 64); tw64 is the
// length-64 twiddle table.  Same contracts as the generic launchers.
bool use_reg64(const BoxGeom& geo);
int launch_plane_ifft64(const double2* q, double2* T, int64_t state0,
                        int nstates, const BoxGeom& geo, const double2* tw64,
                        hipStream_t s);
int launch_plane_fft_gather64(const double2* T, const double2* q,
                              double2* qbar, double* kin_partial,
                              int64_t state0, int nstates, const BoxGeom& geo,
                              const double2* tw64, const GatherParams& gp,
                              hipStream_t s);
int launch_xline_density64(const double2* T, double* rho_part, int groups,
                           const double* occ, int64_t state0, int nstates,
                           double inv_vol, const BoxGeom& geo,
                           const double2* tw64, hipStream_t s);
int launch_xline_vpsi64(double2* T, const double* v, int groups, int nstates,
                        const BoxGeom& geo, const double2* tw64,
                        hipStream_t s);

// Largest plane (ny*nz) the fused plane kernels hold in LDS.
constexpr int kPlaneCap = 6400;
// Lines per block of the x-line kernels.
constexpr int kXLines = 32;

}  // namespace jrx
