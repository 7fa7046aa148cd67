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
  // Rows: the (x, y) lines along z that hold mask points, numbered in C
  // order (x-major), so that the rows of active plane xa are
  // row_off[xa] .. row_off[xa+1] and the mask points of a row are
  // contiguous in g.
  int nrows;
  int rstride;          // elements per z-plane of one state in the row
                        // buffer R[state][z][row] (nrows + 1 rounded up to
                        // 8: element nrows is a pad that nothing reads)
  const int* row_g0;    // [nrows+1] first mask point of each row
  const int* g_row;     // [ng]      row of each mask point
  const int* rowidx;    // [nxa][ny] row of (active plane, y), or -1
};

// y = scale * DFT_axis(x) along one axis of a (B, n0, n1, n2) complex box.
// rin (forward axis-2 pass of one box only): x = sum_s rin[s] of ns real
// boxes instead of in.
int launch_line_fft(const double2* in, double2* out, int64_t B, int n0, int n1,
                    int n2, int axis, bool inverse, double scale,
                    const FftLen& L, hipStream_t s,
                    const double* rin = nullptr, int ns = 0);

// Full 3D transform (three line passes).  In-place allowed.
int fft3d(const double2* in, double2* out, int64_t B, int n0, int n1, int n2,
          bool inverse, double scale, const FftLen& Lx, const FftLen& Ly,
          const FftLen& Lz, hipStream_t s);
// Forward 3D transform of the real box sum_s rin[s] (ns boxes of
// n0 n1 n2 points, added in spin order) into the complex box out.
int fft3d_real_forward(const double* rin, int ns, double2* out, int n0,
                       int n1, int n2, const FftLen& Lx, const FftLen& Ly,
                       const FftLen& Lz, hipStream_t s);

// T[c][xi][y][z] = 2D inverse transform over (y,z) of the scattered plane
// x = plane_x[xi] of state (state0 + c); q is band-major [state][ng].
int launch_plane_ifft(const double2* q, double2* T, int64_t state0,
                      int nstates, const BoxGeom& geo, const FftLen& Ly,
                      const FftLen& Lz, hipStream_t s);

// Groups of states per x-line launch (parallelism across states).
constexpr int kMaxXGroups = 32;
int xlines_per_block(int nx, int nz);
// State groups of an x-line pass: enough groups to reach about
// target_blocks workgroups (0 = 4096), at most kMaxXGroups and nstates.
int xline_groups(const BoxGeom& geo, int nstates, int target_blocks = 0);

// rho_part[g](x,y,z) += sum_{c in group g} occ[state0+c] |IFFT_x T[c]|^2 /
// vol, or = with store (then every group's partial is written, empty groups
// with zeros).  All states of the chunk belong to one spin.
int launch_xline_density(const double2* T, double* rho_part, int groups,
                         const double* occ, int64_t state0, int nstates,
                         double inv_vol, const BoxGeom& geo, const FftLen& Lx,
                         bool store, hipStream_t s);

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
  const short* g_m;    // [ng*3] integer frequencies of the mask points
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
                           const double2* tw64, bool store, hipStream_t s);
int launch_xline_vpsi64(double2* T, const double* v, int groups, int nstates,
                        const BoxGeom& geo, const double2* tw64,
                        hipStream_t s);

// Row-buffer path (rowfft.hip) for cubic 64^3 and 128^3 boxes whose mask
// spans at most 64 x-planes.  R[c][z][row] (row stride geo.rstride) holds
// the z-transformed rows of chunk state c; densities and potentials on
// this path are z-major, [z][y][x].  tw is the length-n twiddle table.
bool use_rowpath(const BoxGeom& geo);
// Row blocks of the row passes (the gather writes that many kinetic and
// band-energy partials per state).
int row_chunks(const BoxGeom& geo);
// State groups of a z-plane pass.
int zplane_groups(const BoxGeom& geo, int nstates);
int launch_row_ifft(const double2* q, double2* R, int64_t state0, int nstates,
                    const BoxGeom& geo, const double2* tw, hipStream_t s);
// rho_part[g][z][y][x] += sum_{c in group g} occ |psi_c|^2 / vol, or =
// with store (then every group's plane is written, empty groups with zeros)
int launch_zplane_density(const double2* R, double* rho_part, int groups,
                          const double* occ, int64_t state0, int nstates,
                          double inv_vol, const BoxGeom& geo,
                          const double2* tw, bool store, hipStream_t s);
// R[c] <- rows of FFT_xy(v * IFFT_xy(R[c])) (in place), v [z][y][x]
int launch_zplane_vpsi(double2* R, const double* v, int groups, int nstates,
                       const BoxGeom& geo, const double2* tw, hipStream_t s);
// qbar and kinetic partials kin_partial[st * row_chunks + chunk]
int launch_row_fft_gather(const double2* R, const double2* q, double2* qbar,
                          double* kin_partial, int64_t state0, int nstates,
                          const BoxGeom& geo, const double2* tw,
                          const GatherParams& gp, hipStream_t s);
// [ns][z][y][x] -> [ns][x][y][z] for cubic n^3 boxes.
int launch_zmajor_to_xmajor(const double* in, double* out, int ns, int n,
                            hipStream_t s);

// Largest plane (ny*nz) the fused plane kernels hold in LDS.
constexpr int kPlaneCap = 6400;
// Lines per block of the x-line kernels.
constexpr int kXLines = 32;

}  // namespace jrx
