// Plan, workspace layout and orchestration of one energy(+gradient)
// evaluation; C ABI entry points.
#include <math.h>
#include <stdlib.h>

#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "energy_kernels.h"
#include "fft_kernels.h"
#include "mlp_core.h"
#include "nonlocal_kernels.h"
#include "qr_kernels.h"

namespace jrx {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

// Twiddle table exp(-2 pi i t / n), t in [0, n), in extended precision.
static std::vector<double2> host_twiddles(int n) {
  std::vector<double2> tw(n);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int t = 0; t < n; ++t) {
    const long double a = two_pi * (long double)t / (long double)n;
    tw[t] = make_double2((double)cosl(a), (double)(-sinl(a)));
  }
  return tw;
}

// Per-device cache of twiddle tables for the plan-less FFT entry points.
static std::mutex g_tw_mutex;
static std::map<std::pair<int, int>, double2*> g_tw_cache;

int cached_fftlen(int n, FftLen* L) {
  if (n < 1 || !fft_factorize(n, L)) {
    set_error("FFT size " + std::to_string(n) +
              " is not a product of 2, 3, 5 and 7");
    return JRX_ERR_INVALID;
  }
  if (n > 2048) {
    set_error("FFT size above 2048 is not supported");
    return JRX_ERR_UNSUPPORTED;
  }
  int dev = 0;
  JRX_CHECK_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_tw_mutex);
  auto it = g_tw_cache.find({dev, n});
  if (it == g_tw_cache.end()) {
    std::vector<double2> h = host_twiddles(n);
    double2* d = nullptr;
    JRX_CHECK_HIP(hipMalloc(&d, n * sizeof(double2)));
    JRX_CHECK_HIP(hipMemcpy(d, h.data(), n * sizeof(double2),
                            hipMemcpyHostToDevice));
    it = g_tw_cache.emplace(std::make_pair(dev, n), d).first;
  }
  L->tw = it->second;
  return JRX_OK;
}

static size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Blocks of the density-statistics reduction.
constexpr int kStatBlocks = 256;

// Workspace of jrx_energy_grad_static: a static device array, one copy per
// GPU, held until the process ends.  In host code `scratch` names a host
// stand-in, not the device array: kernels and copies get the address that
// hipGetSymbolAddress returns on each device (static_workspace below).
static __device__ __align__(256) char scratch[JRX_STATIC_WORKSPACE_BYTES];

static std::mutex g_scratch_mutex;
static std::map<int, void*> g_scratch_addr;  // device -> address of scratch

// Address of `scratch` on the current device, looked up once per device
// (called by jrx_plan_create with the plan's device current).
static int static_workspace(int device, void** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  auto it = g_scratch_addr.find(device);
  if (it == g_scratch_addr.end()) {
    void* ptr = nullptr;
    JRX_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr || ((uintptr_t)ptr & 255) != 0) {
      set_error("static workspace has no 256-byte aligned device address");
      return JRX_ERR_HIP;
    }
    it = g_scratch_addr.emplace(device, ptr).first;
  }
  *out = it->second;
  return JRX_OK;
}

}  // namespace jrx

using namespace jrx;

struct jrx_plan {
  int device = 0;
  int ns = 0, nk = 0, nb = 0, nx = 0, ny = 0, nz = 0, ng = 0, natom = 0;
  int64_t P = 0, S = 0, N = 0;
  double vol = 0, cx = 0;
  double B[9];
  std::vector<void*> allocs;
  BoxGeom geo;
  FftLen Lx, Ly, Lz;
  CellParams cp;
  short* d_g_m = nullptr;
  double* d_kpts = nullptr;
  double* d_occ = nullptr;
  double2* d_sterm = nullptr;  // S(G) / (|G|^2 + 1e-10) on the rho box
  int chunk = 0;   // states per FFT chunk (never straddles spins)
  bool force_generic = false;  // run the runtime-radix kernels at 64^3
  bool keep_planes = false;    // plane buffer holds all states
  int groups = 1;  // state groups of the x-line kernels
  int xline_blocks = 0;  // target workgroups of an x-line pass (0 = 4096)
  bool qr_adaptive = true;  // one CholeskyQR pass where W is well conditioned
  bool band_e = false;      // energy+gradient evaluations also form dE/df
  // Kohn-Sham band mode: evaluations use this potential (the plan's rho/v
  // layout, ns * N) and form sum f <psi|T + v|psi>; null = ground state
  double* d_vfix = nullptr;
  bool fixed = false;
  // norm-conserving pseudopotential: the non-local projectors
  // [nk][nproj][ng] and their scratch (plan-owned, so evaluations of one
  // plan with projectors are not re-entrant across streams); a custom local
  // potential lives in d_sterm (custom_vloc)
  int nproj = 0;
  double2* d_proj = nullptr;
  void* d_nlscratch = nullptr;
  double* d_enl = nullptr;  // per-state non-local energies
  size_t nl_cap = 0, nl_scratch_cap = 0;
  bool custom_vloc = false;
  void* static_ws = nullptr;  // device address of `scratch` on this device
  int xc = JRX_XC_LDA_X;    // JRX_XC_* terms
  bool rowpath = false;     // row-buffer FFT path (z-major rho and v)
  int row_chunks = 0;       // row blocks of the row passes
  CellParams cpz;           // the cell with box axes (z, y, x)
  GramSplit gsw, gsq;
  int kin_parts_max = 0, rec_parts = 0;
  // timing
  int timing = 0;
  std::vector<hipEvent_t> events;
  std::vector<std::pair<int, int>> records;  // (class, event pair index)
  // workspace offsets (bytes)
  size_t o_gb = 0;  // density-gradient boxes of GGA terms [3][N] complex
  size_t o_ritz = 0;  // scratch of the Rayleigh-Ritz rotation
  size_t o_stats;
  size_t o_q, o_qbar, o_T, o_rho, o_rpart, o_v, o_box, o_tb, o_S, o_R1, o_R2, o_Ri, o_A,
      o_Bm, o_part, o_kin, o_eh, o_ee, o_exc, o_qrflag, o_bepart, o_pot,
      total;
};

namespace {

template <typename T>
int upload(jrx_plan* p, const T* h, size_t n, T** out) {
  T* d = nullptr;
  JRX_CHECK_HIP(hipMalloc(&d, n * sizeof(T) + 16));
  p->allocs.push_back(d);
  if (n) JRX_CHECK_HIP(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
  *out = d;
  return JRX_OK;
}

int make_len(jrx_plan* p, int n, FftLen* L) {
  if (!fft_factorize(n, L)) {
    set_error("grid size " + std::to_string(n) +
              " is not a product of 2, 3, 5 and 7");
    return JRX_ERR_INVALID;
  }
  std::vector<double2> tw = host_twiddles(n);
  double2* d = nullptr;
  int rc = upload(p, tw.data(), tw.size(), &d);
  L->tw = d;
  return rc;
}

void det_inv3(const double* a, double* det, double* inv) {
  const double c00 = a[4] * a[8] - a[5] * a[7];
  const double c01 = a[5] * a[6] - a[3] * a[8];
  const double c02 = a[3] * a[7] - a[4] * a[6];
  *det = a[0] * c00 + a[1] * c01 + a[2] * c02;
  const double id = 1.0 / *det;
  inv[0] = c00 * id;
  inv[1] = (a[2] * a[7] - a[1] * a[8]) * id;
  inv[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  inv[3] = c01 * id;
  inv[4] = (a[0] * a[8] - a[2] * a[6]) * id;
  inv[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  inv[6] = c02 * id;
  inv[7] = (a[1] * a[6] - a[0] * a[7]) * id;
  inv[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

int plan_build(jrx_plan* p, const jrx_problem* d) {
  JRX_REQUIRE(d->ns == 1 || d->ns == 2, "ns must be 1 or 2");
  JRX_REQUIRE(d->nk >= 1 && d->nb >= 1, "nk and nb must be positive");
  JRX_REQUIRE((int64_t)d->ns * d->nk <= 65535, "ns*nk above 65535");
  JRX_REQUIRE(d->nx >= 1 && d->ny >= 1 && d->nz >= 1, "grid sizes must be positive");
  JRX_REQUIRE(d->mask && d->kpts && d->cell && d->occupation,
              "mask, kpts, cell and occupation are required");
  JRX_REQUIRE(d->natom >= 0 && (d->natom == 0 || (d->positions && d->charges)),
              "positions and charges are required");
  JRX_REQUIRE(d->xc >= 0 && d->xc <= 15,
              "xc must be an OR of JRX_XC_* terms (0 = lda_x)");
  p->xc = d->xc == 0 ? JRX_XC_LDA_X : d->xc;
  if (p->xc != JRX_XC_LDA_X && d->ns != 1) {
    set_error("xc terms other than lda_x need an unpolarised problem (ns = 1)");
    return JRX_ERR_UNSUPPORTED;
  }
  if (d->nb > kMaxBands) {
    set_error("nb above 128 is not supported by this build");
    return JRX_ERR_UNSUPPORTED;
  }
  JRX_REQUIRE(d->fft_variant >= 0 && d->fft_variant <= 2,
              "fft_variant must be 0 (automatic), 1 (runtime radix) or 2 "
              "(plane buffer)");
  JRX_CHECK_HIP(hipGetDevice(&p->device));
  {
    int rc_ws = static_workspace(p->device, &p->static_ws);
    if (rc_ws) return rc_ws;
  }
  p->ns = d->ns; p->nk = d->nk; p->nb = d->nb;
  p->nx = d->nx; p->ny = d->ny; p->nz = d->nz; p->natom = d->natom;
  p->P = (int64_t)d->ns * d->nk;
  p->S = p->P * d->nb;
  p->N = (int64_t)d->nx * d->ny * d->nz;
  // Cell: vol = |det A|, B = 2 pi A^{-T}
  double det, inv[9];
  det_inv3(d->cell, &det, inv);
  p->vol = fabs(det);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) p->B[3 * i + j] = 2.0 * M_PI * inv[3 * j + i];
  p->cx = 0.75 * cbrt(3.0 / M_PI);
  // Geometry of the mask.
  const int nx = d->nx, ny = d->ny, nz = d->nz;
  std::vector<int> plane_g0, plane_x, g_yz, row_off, row_g0, g_row;
  std::vector<short> rows, g_m;
  int ng = 0, max_rows = 0;
  auto fq = [](int j, int n) { return j < (n + 1) / 2 ? j : j - n; };
  for (int x = 0; x < nx; ++x) {
    int first = -1;
    std::vector<short> prow;
    for (int y = 0; y < ny; ++y) {
      bool any = false;
      for (int z = 0; z < nz; ++z) {
        if (d->mask[((size_t)x * ny + y) * nz + z]) {
          if (first < 0) first = ng;
          if (!any) row_g0.push_back(ng);  // first point of a new row
          g_row.push_back((int)row_g0.size() - 1);
          g_yz.push_back(y * nz + z);
          g_m.push_back((short)fq(x, nx));
          g_m.push_back((short)fq(y, ny));
          g_m.push_back((short)fq(z, nz));
          ++ng;
          any = true;
        }
      }
      if (any) prow.push_back((short)y);
    }
    if (first >= 0) {
      plane_g0.push_back(first);
      plane_x.push_back(x);
      row_off.push_back((int)rows.size());
      rows.insert(rows.end(), prow.begin(), prow.end());
      if ((int)prow.size() > max_rows) max_rows = (int)prow.size();
    }
  }
  JRX_REQUIRE(ng > 0, "mask selects no G vectors");
  JRX_REQUIRE(nx <= 2047 && ny <= 2047 && nz <= 2047, "grid too large");
  plane_g0.push_back(ng);
  row_off.push_back((int)rows.size());
  row_g0.push_back(ng);
  std::vector<int> xmap(nx, -1);
  for (size_t i = 0; i < plane_x.size(); ++i) xmap[plane_x[i]] = (int)i;
  // rows of the row-buffer path: rowidx[xa][y] = row of (plane xa, y)
  const int nrows = (int)rows.size();
  std::vector<int> rowidx(plane_x.size() * (size_t)ny, -1);
  for (size_t xa = 0; xa < plane_x.size(); ++xa)
    for (int r = row_off[xa]; r < row_off[xa + 1]; ++r)
      rowidx[xa * ny + rows[r]] = r;
  p->ng = ng;
  p->geo.nx = nx; p->geo.ny = ny; p->geo.nz = nz; p->geo.ng = ng;
  p->geo.nxa = (int)plane_x.size();
  p->geo.max_rows = max_rows;
  int rc;
  int* tmp = nullptr;
  if ((rc = upload(p, plane_g0.data(), plane_g0.size(), &tmp))) return rc;
  p->geo.plane_g0 = tmp;
  if ((rc = upload(p, plane_x.data(), plane_x.size(), &tmp))) return rc;
  p->geo.plane_x = tmp;
  if ((rc = upload(p, g_yz.data(), g_yz.size(), &tmp))) return rc;
  p->geo.g_yz = tmp;
  if ((rc = upload(p, row_off.data(), row_off.size(), &tmp))) return rc;
  p->geo.row_off = tmp;
  short* srows = nullptr;
  if ((rc = upload(p, rows.data(), rows.size(), &srows))) return rc;
  p->geo.rows = srows;
  if ((rc = upload(p, g_m.data(), g_m.size(), &p->d_g_m))) return rc;
  if ((rc = upload(p, xmap.data(), xmap.size(), &tmp))) return rc;
  p->geo.xmap = tmp;
  p->geo.nrows = nrows;
  // (at least one pad element per z-plane slice: the z-plane passes store
  // their off-mask lanes there)
  p->geo.rstride = (nrows + 1 + 7) / 8 * 8;
  if ((rc = upload(p, row_g0.data(), row_g0.size(), &tmp))) return rc;
  p->geo.row_g0 = tmp;
  if ((rc = upload(p, g_row.data(), g_row.size(), &tmp))) return rc;
  p->geo.g_row = tmp;
  if ((rc = upload(p, rowidx.data(), rowidx.size(), &tmp))) return rc;
  p->geo.rowidx = tmp;
  if ((rc = make_len(p, nx, &p->Lx))) return rc;
  if ((rc = make_len(p, ny, &p->Ly))) return rc;
  if ((rc = make_len(p, nz, &p->Lz))) return rc;
  if ((rc = upload(p, d->kpts, (size_t)d->nk * 3, &p->d_kpts))) return rc;
  if ((rc = upload(p, d->occupation, (size_t)p->S, &p->d_occ))) return rc;
  double *dpos = nullptr, *dchg = nullptr;
  if ((rc = upload(p, d->positions, (size_t)d->natom * 3, &dpos))) return rc;
  if ((rc = upload(p, d->charges, (size_t)d->natom, &dchg))) return rc;
  p->cp.nx = nx; p->cp.ny = ny; p->cp.nz = nz;
  for (int i = 0; i < 9; ++i) p->cp.B[i] = p->B[i];
  p->cp.vol = p->vol;
  p->cp.natom = d->natom;
  p->cp.pos = dpos;
  p->cp.charge = dchg;
  // FFT path: the row buffer for cubic 64^3 / 128^3 boxes (fft_variant 0),
  // else the plane buffer (register 64-point kernels where they apply, or
  // the runtime-radix kernels everywhere with fft_variant 1).
  // JRX_FFT_GENERIC=1 selects the runtime-radix kernels too (cross-checks).
  const char* gen = getenv("JRX_FFT_GENERIC");
  p->force_generic = (gen && gen[0] == '1') || d->fft_variant == 1;
  p->rowpath = d->fft_variant == 0 && !p->force_generic &&
               use_rowpath(p->geo);
  if (!p->rowpath && (int64_t)ny * nz > kPlaneCap) {
    set_error("ny*nz above 6400 is supported only on the row-buffer path "
              "(cubic 64^3 or 128^3 boxes whose mask spans at most 64 "
              "x-planes)");
    return JRX_ERR_UNSUPPORTED;
  }
  // LDS requirements against the device limit.
  int max_lds = 0;
  JRX_CHECK_HIP(hipDeviceGetAttribute(
      &max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, p->device));
  const int nb = d->nb;
  // (the row path's largest block holds an nz x 72 z-plane and twiddles)
  const size_t lds_plane =
      (p->rowpath ? (size_t)nz * 73 : (size_t)ny * nz) * sizeof(double2);
  // (the output transpose of jrx_plan_orthonormal_params: [nb][65])
  const size_t lds_params = (size_t)nb * 65 * sizeof(double2);
  size_t need = lds_plane;
  for (size_t v : {qr_lds_bytes(nb), lds_params}) need = v > need ? v : need;
  if (need > (size_t)max_lds) {
    set_error("problem needs " + std::to_string(need) +
              " bytes of LDS per workgroup; device allows " +
              std::to_string(max_lds));
    return JRX_ERR_UNSUPPORTED;
  }
  // Chunking of the FFT sweeps.
  // Plane stride of the plane buffer: ny*nz rounded up to 16 elements
  // (256 B), plus 16 when that is a multiple of 256 elements, so that the
  // x-planes an x-line pass reads together are not 2^k bytes apart.
  {
    const int64_t P2 = (int64_t)ny * nz;
    int64_t ps = (P2 + 15) / 16 * 16;
    if (d->plane_pad >= 0) ps = P2 + d->plane_pad;
    else if (ps % 256 == 0) ps += 16;
    JRX_REQUIRE(ps < (int64_t)1 << 30, "plane too large");
    p->geo.ps = (int)ps;
  }
  const size_t per_state =
      p->rowpath ? (size_t)nz * p->geo.rstride * sizeof(double2)
                 : (size_t)p->geo.nxa * p->geo.ps * sizeof(double2);
  // Chunking of the FFT sweeps.  When the plane buffer can hold every state
  // (default budget 16 GiB), the forward sweep's planes are kept and the
  // adjoint sweep starts from them instead of transforming Q again.
  int64_t chunk = d->chunk_states;
  if (chunk <= 0) {
    const size_t budget = (size_t)16 << 30;
    chunk = (int64_t)(budget / per_state);
    if (chunk < 1) chunk = 1;
  }
  const int64_t spin_states = (int64_t)d->nk * d->nb;
  p->keep_planes = chunk >= p->S && p->S <= 65535;
  if (p->keep_planes) chunk = p->S;
  else if (chunk > spin_states) chunk = spin_states;
  if (chunk > 65535) chunk = 65535;
  p->chunk = (int)chunk;
  JRX_REQUIRE(d->xline_blocks >= 0, "xline_blocks must be >= 0");
  JRX_REQUIRE(d->qr_passes == 0 || d->qr_passes == 2,
              "qr_passes must be 0 (adaptive) or 2");
  p->qr_adaptive = d->qr_passes == 0;
  p->xline_blocks = d->xline_blocks > 0 ? d->xline_blocks : 0;
  const int sweep_states = p->keep_planes ? (int)spin_states : p->chunk;
  p->groups = p->rowpath ? zplane_groups(p->geo, sweep_states)
                         : xline_groups(p->geo, sweep_states, p->xline_blocks);
  p->row_chunks = p->rowpath ? row_chunks(p->geo) : 0;
  // z-major cell of the row path: box axes (z, y, x)
  p->cpz = p->cp;
  p->cpz.nx = nz;
  p->cpz.nz = nx;
  for (int i = 0; i < 3; ++i) {
    p->cpz.B[i] = p->B[6 + i];
    p->cpz.B[6 + i] = p->B[i];
  }
  // The geometry-only factor of the external term, S(G) / (|G|^2 + 1e-10),
  // on the box rho lives in (it does not change between evaluations).
  JRX_CHECK_HIP(hipMalloc(&p->d_sterm, (size_t)p->N * sizeof(double2)));
  p->allocs.push_back(p->d_sterm);
  if ((rc = launch_structure_term(p->rowpath ? p->cpz : p->cp, p->d_sterm,
                                  nullptr)))
    return rc;
  JRX_CHECK_HIP(hipStreamSynchronize(nullptr));
  p->gsw = gram_split(p->P, ng, nb);
  p->gsq = p->gsw;
  const int64_t kin_gather =
      p->S * (p->rowpath ? p->row_chunks : p->geo.nxa);
  const int64_t kin_plain = p->S * ((ng + kRedBlock - 1) / kRedBlock);
  p->kin_parts_max = (int)(kin_gather > kin_plain ? kin_gather : kin_plain);
  p->rec_parts = (int)((p->N + kRedBlock - 1) / kRedBlock);
  // Workspace layout.
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align_up(off + bytes);
    return o;
  };
  const size_t mat = (size_t)p->P * nb * nb * sizeof(double2);
  p->o_q = take((size_t)p->S * ng * sizeof(double2));
  p->o_qbar = take((size_t)p->S * ng * sizeof(double2));
  p->o_T = take((size_t)p->chunk * per_state);
  p->o_rho = take((size_t)p->ns * p->N * sizeof(double));
  p->o_rpart = take((size_t)p->groups * p->N * sizeof(double));
  p->o_v = take((size_t)p->ns * p->N * sizeof(double));
  p->o_box = take((size_t)p->N * sizeof(double2));
  p->o_S = take(mat);
  p->o_R1 = take(mat);
  p->o_R2 = take(mat);
  p->o_Ri = take(mat);
  p->o_A = take(mat);
  p->o_Bm = take(mat);
  p->o_part = take((size_t)p->P * p->gsw.nchunk * nb * nb * sizeof(double2));
  p->o_kin = take((size_t)p->kin_parts_max * sizeof(double));
  p->o_tb = take((size_t)p->S * sizeof(double));
  p->o_eh = take((size_t)p->rec_parts * sizeof(double));
  p->o_ee = take((size_t)p->rec_parts * sizeof(double));
  p->o_exc = take((size_t)p->rec_parts * sizeof(double));
  p->o_stats = take((size_t)(2 * kStatBlocks + 8) * sizeof(double));
  if (xc_is_gga(p->xc)) p->o_gb = take(3 * (size_t)p->N * sizeof(double2));
  p->o_qrflag = take((size_t)p->P * sizeof(int));
  p->o_bepart = take((size_t)kin_gather * sizeof(double));
  p->o_pot = take((size_t)p->S * sizeof(double));
  p->o_ritz = take(ritz_scratch_bytes(p->P, nb));
  p->total = off;
  return JRX_OK;
}

}  // namespace

extern "C" {

const char* jrx_last_error(void) { return g_last_error.c_str(); }
const char* jrx_version(void) { return "jrx-hip 0.1 (gfx950, fp64)"; }

static int generic_fft(const double* x, double* y, int64_t batch, int nx,
                       int ny, int nz, double scale, void* stream, bool inv) {
  JRX_REQUIRE(batch >= 0 && nx > 0 && ny > 0 && nz > 0, "invalid FFT shape");
  JRX_REQUIRE(x && y, "null FFT buffer");
  if (batch == 0) return JRX_OK;
  FftLen Lx, Ly, Lz;
  int rc;
  if ((rc = cached_fftlen(nx, &Lx)) || (rc = cached_fftlen(ny, &Ly)) ||
      (rc = cached_fftlen(nz, &Lz)))
    return rc;
  return fft3d(reinterpret_cast<const double2*>(x),
               reinterpret_cast<double2*>(y), batch, nx, ny, nz, inv, scale,
               Lx, Ly, Lz, (hipStream_t)stream);
}

int jrx_fftn3d(const double* x, double* y, int64_t batch, int nx, int ny,
               int nz, double scale, void* stream) {
  return generic_fft(x, y, batch, nx, ny, nz, scale, stream, false);
}

int jrx_ifftn3d(const double* x, double* y, int64_t batch, int nx, int ny,
                int nz, double scale, void* stream) {
  return generic_fft(x, y, batch, nx, ny, nz, scale, stream, true);
}

int jrx_plan_create(jrx_plan** plan, const jrx_problem* problem) {
  JRX_REQUIRE(plan && problem, "null argument");
  jrx_plan* p = new jrx_plan();
  int rc = plan_build(p, problem);
  if (rc) {
    for (void* a : p->allocs) (void)hipFree(a);
    delete p;
    *plan = nullptr;
    return rc;
  }
  *plan = p;
  return JRX_OK;
}

int jrx_plan_destroy(jrx_plan* plan) {
  if (!plan) return JRX_OK;
  for (hipEvent_t e : plan->events) (void)hipEventDestroy(e);
  for (void* a : plan->allocs) (void)hipFree(a);
  delete plan;
  return JRX_OK;
}

int jrx_plan_workspace_bytes(const jrx_plan* plan, size_t* bytes) {
  JRX_REQUIRE(plan && bytes, "null argument");
  *bytes = plan->total;
  return JRX_OK;
}

int jrx_plan_param_count(const jrx_plan* plan, int64_t* count) {
  JRX_REQUIRE(plan && count, "null argument");
  *count = plan->S * plan->ng;
  return JRX_OK;
}

int jrx_plan_get_info(const jrx_plan* plan, jrx_plan_info* info) {
  JRX_REQUIRE(plan && info, "null argument");
  info->ng = plan->ng;
  info->active_planes = plan->geo.nxa;
  info->chunk_states = plan->chunk;
  info->keep_planes = plan->keep_planes ? 1 : 0;
  info->workspace_bytes = plan->total;
  info->row_path = plan->rowpath ? 1 : 0;
  info->rows = plan->geo.nrows;
  return JRX_OK;
}

// Teter-Payne-Allan diagonal preconditioner (PRB 40, 12255 (1989)) applied
// to a gradient in the parameter layout [s][k][g][b]:
//   x = |G+k|^2 / tb[state]  (plane-wave over band kinetic energy),
//   K(x) = (27 + 18x + 12x^2 + 8x^3) / (27 + 18x + 12x^2 + 8x^3 + 16x^4),
// optionally times a per-coefficient factor (learned preconditioner).
// |G+k|^2 of mask point g at k-point k.
__device__ inline double coef_gk2(const short* __restrict__ g_m,
                                  const double* __restrict__ kpts,
                                  const CellParams& cp, int g, int k) {
  const int mx = g_m[3 * g], my = g_m[3 * g + 1], mz = g_m[3 * g + 2];
  const double gx = mx * cp.B[0] + my * cp.B[3] + mz * cp.B[6] + kpts[3 * k];
  const double gy = mx * cp.B[1] + my * cp.B[4] + mz * cp.B[7] + kpts[3 * k + 1];
  const double gz = mx * cp.B[2] + my * cp.B[5] + mz * cp.B[8] + kpts[3 * k + 2];
  return gx * gx + gy * gy + gz * gz;
}

__global__ void tpa_kernel(const double* __restrict__ gr,
                           const double* __restrict__ gi,
                           const double* __restrict__ tb,
                           const short* __restrict__ g_m,
                           const double* __restrict__ kpts, CellParams cp,
                           int ng, int nk, int nb, int64_t n,
                           const double* __restrict__ factor,
                           double* __restrict__ outr,
                           double* __restrict__ outi) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int b = (int)(i % nb);
  const int64_t rest = i / nb;
  const int g = (int)(rest % ng);
  const int64_t sk = rest / ng;
  const int k = (int)(sk % nk);
  const int64_t st = sk * nb + b;
  const double t = tb[st];
  const double x = coef_gk2(g_m, kpts, cp, g, k) / (t > 1e-300 ? t : 1e-300);
  const double p = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x));
  double kx = p / (p + 16.0 * x * x * x * x);
  if (factor) kx *= factor[i];
  outr[i] = kx * gr[i];
  outi[i] = kx * gi[i];
}

extern "C" int jrx_plan_precondition(jrx_plan* p, const double* g_re,
                                     const double* g_im, const double* factor,
                                     double* out_re, double* out_im,
                                     void* workspace, size_t workspace_bytes,
                                     void* stream) {
  JRX_REQUIRE(p && g_re && g_im && out_re && out_im && workspace,
              "null argument");
  JRX_REQUIRE(workspace_bytes >= p->total, "workspace too small");
  const int64_t n = p->S * p->ng;
  const double* tb = reinterpret_cast<const double*>(
      static_cast<const char*>(workspace) + p->o_tb);
  hipLaunchKernelGGL(tpa_kernel, dim3(ceil_div(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, g_re, g_im, tb, p->d_g_m, p->d_kpts,
                     p->cp, p->ng, p->nk, p->nb, n, factor, out_re, out_im);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

// Fused CG step quantities: h1 = K(|G+k|^2/tb) * factor * g1 over the
// parameters (real and imaginary planes), and block partial sums of
// g1.h1, g0.h1 and g1.d.  Block b takes the contiguous slice
// [b*chunk, (b+1)*chunk) of each plane; fixed-order sums.
constexpr int kFuseBlocks = 1024;
__global__ void __launch_bounds__(256)
tpa_dots_kernel(const double* __restrict__ g1r, const double* __restrict__ g1i,
                const double* __restrict__ g0r, const double* __restrict__ g0i,
                const double* __restrict__ dr, const double* __restrict__ di,
                const double* __restrict__ tb, const short* __restrict__ g_m,
                const double* __restrict__ kpts, CellParams cp, int ng,
                int nk, int nb, int64_t n, int64_t chunk,
                const double* __restrict__ factor, double* __restrict__ hr,
                double* __restrict__ hi, double* __restrict__ part) {
  __shared__ double red[256];
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi_end = lo + chunk < n ? lo + chunk : n;
  double s11 = 0.0, s01 = 0.0, s1d = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi_end; i += 256) {
    const int b = (int)(i % nb);
    const int64_t rest = i / nb;
    const int g = (int)(rest % ng);
    const int64_t sk = rest / ng;
    const int k = (int)(sk % nk);
    const double t = tb[sk * nb + b];
    const double x = coef_gk2(g_m, kpts, cp, g, k) / (t > 1e-300 ? t : 1e-300);
    const double pp = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x));
    double kx = pp / (pp + 16.0 * x * x * x * x);
    if (factor) kx *= factor[i];
    const double ar = g1r[i], ai = g1i[i];
    const double h_r = kx * ar, h_i = kx * ai;
    hr[i] = h_r;
    hi[i] = h_i;
    s11 += ar * h_r + ai * h_i;
    s01 += g0r[i] * h_r + g0i[i] * h_i;
    s1d += ar * dr[i] + ai * di[i];
  }
  s11 = block_sum(s11, red);
  s01 = block_sum(s01, red);
  s1d = block_sum(s1d, red);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = s11;
    part[3 * blockIdx.x + 1] = s01;
    part[3 * blockIdx.x + 2] = s1d;
  }
}

__global__ void __launch_bounds__(256)
sum3_kernel(const double* __restrict__ part, int nblk,
            double* __restrict__ out) {
  __shared__ double red[256];
  double a = 0.0, b = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < nblk; i += 256) {
    a += part[3 * i];
    b += part[3 * i + 1];
    c += part[3 * i + 2];
  }
  a = block_sum(a, red);
  b = block_sum(b, red);
  c = block_sum(c, red);
  if (threadIdx.x == 0) {
    out[0] = a;
    out[1] = b;
    out[2] = c;
  }
}

extern "C" size_t jrx_cg_dots_workspace_bytes(void) {
  return (size_t)(3 * kFuseBlocks + 8) * sizeof(double);
}

extern "C" int jrx_plan_precondition_dots(
    jrx_plan* p, const double* g1_re, const double* g1_im,
    const double* g0_re, const double* g0_im, const double* d_re,
    const double* d_im, const double* factor, double* h1_re, double* h1_im,
    double* out, void* dots_ws, size_t dots_ws_bytes, const void* workspace,
    size_t workspace_bytes, void* stream) {
  JRX_REQUIRE(p && g1_re && g1_im && g0_re && g0_im && d_re && d_im &&
                  h1_re && h1_im && out && dots_ws && workspace,
              "null argument");
  JRX_REQUIRE(workspace_bytes >= p->total, "workspace too small");
  JRX_REQUIRE(dots_ws_bytes >= jrx_cg_dots_workspace_bytes(),
              "dots workspace too small (jrx_cg_dots_workspace_bytes)");
  const int64_t n = p->S * p->ng;
  int64_t nblk = (n + 256 * 16 - 1) / (256 * 16);
  if (nblk < 1) nblk = 1;
  if (nblk > kFuseBlocks) nblk = kFuseBlocks;
  const int64_t chunk = (n + nblk - 1) / nblk;
  const double* tb = reinterpret_cast<const double*>(
      static_cast<const char*>(workspace) + p->o_tb);
  double* part = static_cast<double*>(dots_ws);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(tpa_dots_kernel, dim3((unsigned)nblk), dim3(256), 0, s,
                     g1_re, g1_im, g0_re, g0_im, d_re, d_im, tb, p->d_g_m,
                     p->d_kpts, p->cp, p->ng, p->nk, p->nb, n, chunk, factor,
                     h1_re, h1_im, part);
  JRX_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum3_kernel, dim3(1), dim3(256), 0, s, part, (int)nblk,
                     out);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

extern "C" int jrx_plan_band_norms(const jrx_plan* p, double* out,
                                   const void* workspace,
                                   size_t workspace_bytes, void* stream) {
  JRX_REQUIRE(p && out && workspace, "null argument");
  JRX_REQUIRE(workspace_bytes >= p->total, "workspace too small");
  JRX_CHECK_HIP(hipMemcpyAsync(
      out, static_cast<const char*>(workspace) + p->o_tb,
      (size_t)p->S * sizeof(double), hipMemcpyDeviceToDevice,
      (hipStream_t)stream));
  return JRX_OK;
}

__global__ void band_energy_kernel(const double* __restrict__ tb,
                                   const double* __restrict__ pot,
                                   const double* __restrict__ enl,
                                   double two_over_vol, int64_t S,
                                   double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < S)
    out[i] = 0.5 * (tb[i] + two_over_vol * pot[i]) + (enl ? enl[i] : 0.0);
}

extern "C" int jrx_plan_set_band_energies(jrx_plan* p, int on) {
  JRX_REQUIRE(p, "null plan");
  p->band_e = on != 0;
  return JRX_OK;
}

extern "C" int jrx_plan_band_energies(const jrx_plan* p, double* out,
                                      const void* workspace,
                                      size_t workspace_bytes, void* stream) {
  JRX_REQUIRE(p && out && workspace, "null argument");
  JRX_REQUIRE(workspace_bytes >= p->total, "workspace too small");
  JRX_REQUIRE(p->band_e, "band energies are off (jrx_plan_set_band_energies)");
  const char* ws = static_cast<const char*>(workspace);
  hipLaunchKernelGGL(band_energy_kernel, dim3(ceil_div(p->S, 256)), dim3(256),
                     0, (hipStream_t)stream,
                     reinterpret_cast<const double*>(ws + p->o_tb),
                     reinterpret_cast<const double*>(ws + p->o_pot),
                     p->nproj > 0 ? p->d_enl : nullptr, 2.0 / p->vol, p->S,
                     out);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

extern "C" int jrx_plan_set_occupation(jrx_plan* p, const double* occupation,
                                       void* stream) {
  JRX_REQUIRE(p && occupation, "null argument");
  JRX_CHECK_HIP(hipMemcpyAsync(p->d_occ, occupation,
                               (size_t)p->S * sizeof(double),
                               hipMemcpyDefault, (hipStream_t)stream));
  return JRX_OK;
}

// Q [P][nb][ng] (complex, band-major) -> out_re / out_im [P][ng][nb], one
// 64-wide G tile of one (s,k) block per workgroup, transposed through LDS
// (tile [nb][65], dynamic).
__global__ void __launch_bounds__(256)
q_to_params_kernel(const double2* __restrict__ q, int ng, int nb,
                   double* __restrict__ outr, double* __restrict__ outi) {
  extern __shared__ double2 tile[];
  const int64_t pk = blockIdx.y;
  const int g0 = blockIdx.x * 64;
  const double2* src = q + pk * (int64_t)nb * ng;
  for (int i = threadIdx.x; i < nb * 64; i += 256) {
    const int b = i >> 6, gl = i & 63, g = g0 + gl;
    tile[b * 65 + gl] =
        g < ng ? src[(int64_t)b * ng + g] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nb * 64; i += 256) {
    const int gl = i / nb, b = i % nb, g = g0 + gl;
    if (g < ng) {
      const int64_t o = (pk * ng + g) * nb + b;
      outr[o] = tile[b * 65 + gl].x;
      outi[o] = tile[b * 65 + gl].y;
    }
  }
}

extern "C" int jrx_plan_orthonormal_params(const jrx_plan* p, double* out_re,
                                           double* out_im,
                                           const void* workspace,
                                           size_t workspace_bytes,
                                           void* stream) {
  JRX_REQUIRE(p && out_re && out_im && workspace, "null argument");
  JRX_REQUIRE(workspace_bytes >= p->total, "workspace too small");
  const double2* q = reinterpret_cast<const double2*>(
      static_cast<const char*>(workspace) + p->o_q);
  hipLaunchKernelGGL(q_to_params_kernel,
                     dim3((unsigned)ceil_div(p->ng, 64), (unsigned)p->P),
                     dim3(256), (size_t)p->nb * 65 * sizeof(double2),
                     (hipStream_t)stream, q, p->ng, p->nb, out_re, out_im);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

// out[p][g][b] = sum_c Q[p][c][g] V[p][c][b]: one 32-wide G tile of one
// (s,k) block per workgroup; the Q tile ([nb][kRotG + 1]) and, up to 64
// bands, V[p] are staged in LDS (dynamic); wider V is read through the
// caches.
constexpr int kRotG = 32;
__global__ void __launch_bounds__(256)
q_rotate_kernel(const double2* __restrict__ q, const double2* __restrict__ V,
                int ng, int nb, double* __restrict__ outr,
                double* __restrict__ outi) {
  extern __shared__ double2 tile[];
  const bool vlds = nb <= kNarrowBands;
  double2* vs = tile + nb * (kRotG + 1);
  const int64_t pk = blockIdx.y;
  const int g0 = blockIdx.x * kRotG;
  const double2* src = q + pk * (int64_t)nb * ng;
  for (int i = threadIdx.x; i < nb * kRotG; i += 256) {
    const int c = i / kRotG, gl = i % kRotG, g = g0 + gl;
    tile[c * (kRotG + 1) + gl] =
        g < ng ? src[(int64_t)c * ng + g] : make_double2(0.0, 0.0);
  }
  const double2* vp = V + pk * nb * nb;
  if (vlds)
    for (int i = threadIdx.x; i < nb * nb; i += 256) vs[i] = vp[i];
  __syncthreads();
  const double2* vsrc = vlds ? vs : vp;
  for (int i = threadIdx.x; i < nb * kRotG; i += 256) {
    const int gl = i / nb, b = i % nb, g = g0 + gl;
    if (g >= ng) continue;
    double re = 0.0, im = 0.0;
    for (int c = 0; c < nb; ++c) {
      const double2 a = tile[c * (kRotG + 1) + gl], v = vsrc[c * nb + b];
      re += a.x * v.x - a.y * v.y;
      im += a.x * v.y + a.y * v.x;
    }
    const int64_t o = (pk * ng + g) * nb + b;
    outr[o] = re;
    outi[o] = im;
  }
}

extern "C" int jrx_plan_subspace_matrix(const jrx_plan* p, double* out,
                                        const void* workspace,
                                        size_t workspace_bytes, void* stream) {
  JRX_REQUIRE(p && out && workspace, "null argument");
  JRX_REQUIRE(workspace_bytes >= p->total, "workspace too small");
  JRX_CHECK_HIP(hipMemcpyAsync(
      out, static_cast<const char*>(workspace) + p->o_S,
      (size_t)p->P * p->nb * p->nb * sizeof(double2),
      hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return JRX_OK;
}

extern "C" int jrx_plan_ritz_rotation(const jrx_plan* p, double* v,
                                      void* workspace, size_t workspace_bytes,
                                      void* stream) {
  JRX_REQUIRE(p && v && workspace, "null argument");
  JRX_REQUIRE(workspace_bytes >= p->total, "workspace too small");
  char* ws = static_cast<char*>(workspace);
  return launch_ritz(reinterpret_cast<const double2*>(ws + p->o_S), p->d_occ,
                     p->P, p->nb, reinterpret_cast<double2*>(v), nullptr,
                     false, ws + p->o_ritz, (hipStream_t)stream);
}

extern "C" int jrx_plan_rotate(const jrx_plan* p, const double* v,
                               double* out_re, double* out_im,
                               const void* workspace, size_t workspace_bytes,
                               void* stream) {
  JRX_REQUIRE(p && v && out_re && out_im && workspace, "null argument");
  JRX_REQUIRE(workspace_bytes >= p->total, "workspace too small");
  const double2* q = reinterpret_cast<const double2*>(
      static_cast<const char*>(workspace) + p->o_q);
  const size_t shm =
      ((size_t)p->nb * (kRotG + 1) +
       (p->nb <= kNarrowBands ? (size_t)p->nb * p->nb : 0)) *
      sizeof(double2);
  hipLaunchKernelGGL(q_rotate_kernel,
                     dim3((unsigned)ceil_div(p->ng, kRotG), (unsigned)p->P),
                     dim3(256), shm, (hipStream_t)stream, q,
                     reinterpret_cast<const double2*>(v), p->ng, p->nb,
                     out_re, out_im);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

// Partial sums of rho_tot and rho_tot^2 over contiguous chunks of the grid
// (rho_tot = sum over spins), then mean and variance in one block.
__global__ void __launch_bounds__(256)
rho_stats_part_kernel(const double* __restrict__ rho, int ns, int64_t N,
                      int64_t chunk, double* __restrict__ part) {
  __shared__ double red[256];
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < N ? lo + chunk : N;
  double s1 = 0.0, s2 = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += 256) {
    double r = 0.0;
    for (int sp = 0; sp < ns; ++sp) r += rho[sp * N + i];
    s1 += r;
    s2 += r * r;
  }
  s1 = block_sum(s1, red);
  s2 = block_sum(s2, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1;
    part[2 * blockIdx.x + 1] = s2;
  }
}

__global__ void __launch_bounds__(256)
rho_stats_final_kernel(const double* __restrict__ part, int nblk, int64_t N,
                       double* __restrict__ out) {
  __shared__ double red[256];
  double s1 = 0.0, s2 = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 256) {
    s1 += part[2 * b];
    s2 += part[2 * b + 1];
  }
  s1 = block_sum(s1, red);
  s2 = block_sum(s2, red);
  if (threadIdx.x == 0) {
    const double mean = s1 / (double)N;
    const double var = s2 / (double)N - mean * mean;
    out[0] = mean;
    out[1] = var > 0.0 ? var : 0.0;
  }
}

// Raw network features of the coefficients idx[i] (flat [s][k][g][b]
// index); an index outside the parameter vector gets zero features.
__global__ void coef_features_kernel(const int64_t* __restrict__ idx,
                                     int64_t n, int64_t total,
                                     const double* __restrict__ tb,
                                     const short* __restrict__ g_m,
                                     const double* __restrict__ kpts,
                                     CellParams cp, int ng, int nk, int nb,
                                     const double* __restrict__ stats,
                                     const double* __restrict__ gap,
                                     float* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int64_t c = idx[i];
  float f[kMlpIn];
  if (c >= 0 && c < total) {
    const int b = (int)(c % nb);
    const int64_t rest = c / nb;
    const int g = (int)(rest % ng);
    const int64_t sk = rest / ng;
    const int k = (int)(sk % nk);
    const int64_t st = sk * nb + b;
    mlp_features(coef_gk2(g_m, kpts, cp, g, k), tb[st], gap ? gap[st] : 0.0,
                 stats[0], stats[1], f);
  } else {
    for (int j = 0; j < kMlpIn; ++j) f[j] = 0.f;
  }
  for (int j = 0; j < kMlpIn; ++j) x[i * kMlpIn + j] = f[j];
}

// Network inference on every coefficient: factor[i] = exp(y0) and, when
// w is given, w[i] *= exp(y1) (warm start), y clamped before exp.
__global__ void __launch_bounds__(256)
coef_mlp_kernel(const float* __restrict__ params,
                const float* __restrict__ norm, int64_t n,
                const double* __restrict__ tb, const short* __restrict__ g_m,
                const double* __restrict__ kpts, CellParams cp, int ng,
                int nk, int nb, const double* __restrict__ stats,
                const double* __restrict__ gap,
                double* __restrict__ factor, double* __restrict__ wr,
                double* __restrict__ wi) {
  __shared__ float w[kMlpParams];
  mlp_stage(params, w);
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int b = (int)(i % nb);
  const int64_t rest = i / nb;
  const int g = (int)(rest % ng);
  const int64_t sk = rest / ng;
  const int k = (int)(sk % nk);
  float f[kMlpIn];
  const int64_t st = sk * nb + b;
  mlp_features(coef_gk2(g_m, kpts, cp, g, k), tb[st], gap ? gap[st] : 0.0,
               stats[0], stats[1], f);
  mlp_normalise(norm, f);
  float y[kMlpOut];
  mlp_eval(w, f, y);
  if (factor)
    factor[i] = exp((double)fminf(fmaxf(y[0], kLogFactorMin), kLogFactorMax));
  if (wr) {
    const double a =
        exp((double)fminf(fmaxf(y[1], kLogAmpMin), kLogAmpMax));
    wr[i] *= a;
    wi[i] *= a;
  }
}

extern "C" int jrx_plan_density_stats(const jrx_plan* p, double* out,
                                      void* workspace, size_t workspace_bytes,
                                      void* stream) {
  JRX_REQUIRE(p && out && workspace, "null argument");
  JRX_REQUIRE(workspace_bytes >= p->total, "workspace too small");
  char* ws = static_cast<char*>(workspace);
  const double* rho = reinterpret_cast<const double*>(ws + p->o_rho);
  double* part = reinterpret_cast<double*>(ws + p->o_stats);
  const int64_t chunk = (p->N + kStatBlocks - 1) / kStatBlocks;
  const int nblk = (int)((p->N + chunk - 1) / chunk);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(rho_stats_part_kernel, dim3(nblk), dim3(256), 0, s, rho,
                     p->ns, p->N, chunk, part);
  JRX_LAUNCH_CHECK();
  hipLaunchKernelGGL(rho_stats_final_kernel, dim3(1), dim3(256), 0, s, part,
                     nblk, p->N, out);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

extern "C" int jrx_plan_features(const jrx_plan* p, const int64_t* idx,
                                 int64_t n, const double* stats,
                                 const double* gap, float* x,
                                 const void* workspace,
                                 size_t workspace_bytes, void* stream) {
  JRX_REQUIRE(p && idx && stats && x && workspace && n >= 0,
              "null argument");
  JRX_REQUIRE(workspace_bytes >= p->total, "workspace too small");
  if (n == 0) return JRX_OK;
  const double* tb = reinterpret_cast<const double*>(
      static_cast<const char*>(workspace) + p->o_tb);
  hipLaunchKernelGGL(coef_features_kernel, dim3(ceil_div(n, 256)), dim3(256),
                     0, (hipStream_t)stream, idx, n, p->S * p->ng, tb,
                     p->d_g_m, p->d_kpts, p->cp, p->ng, p->nk, p->nb, stats,
                     gap, x);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

extern "C" int jrx_plan_mlp_apply(const jrx_plan* p, const float* params,
                                  const float* norm, const double* stats,
                                  const double* gap, double* factor,
                                  double* w_re, double* w_im,
                                  const void* workspace,
                                  size_t workspace_bytes, void* stream) {
  JRX_REQUIRE(p && params && norm && stats && workspace, "null argument");
  JRX_REQUIRE((w_re == nullptr) == (w_im == nullptr),
              "w_re and w_im must both be given or both be NULL");
  JRX_REQUIRE(factor || w_re, "nothing to compute");
  JRX_REQUIRE(workspace_bytes >= p->total, "workspace too small");
  const int64_t n = p->S * p->ng;
  const double* tb = reinterpret_cast<const double*>(
      static_cast<const char*>(workspace) + p->o_tb);
  hipLaunchKernelGGL(coef_mlp_kernel, dim3(ceil_div(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, params, norm, n, tb, p->d_g_m,
                     p->d_kpts, p->cp, p->ng, p->nk, p->nb, stats, gap,
                     factor, w_re, w_im);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int jrx_plan_set_timing(jrx_plan* plan, int enable) {
  JRX_REQUIRE(plan, "null plan");
  plan->timing = enable ? 1 : 0;
  return JRX_OK;
}

int jrx_plan_read_timing(jrx_plan* plan, double* ms, int* launches) {
  JRX_REQUIRE(plan && ms && launches, "null argument");
  for (int c = 0; c < JRX_K_COUNT; ++c) {
    ms[c] = 0.0;
    launches[c] = 0;
  }
  for (const auto& r : plan->records) {
    hipEvent_t a = plan->events[2 * r.second], b = plan->events[2 * r.second + 1];
    JRX_CHECK_HIP(hipEventSynchronize(b));
    float t = 0.f;
    JRX_CHECK_HIP(hipEventElapsedTime(&t, a, b));
    ms[r.first] += t;
    launches[r.first] += 1;
  }
  plan->records.clear();
  return JRX_OK;
}

static int timing_begin(jrx_plan* p, hipStream_t s, int* slot, bool on) {
  *slot = -1;
  if (!on || !p->timing) return JRX_OK;
  const int i = (int)p->records.size();
  while ((int)p->events.size() < 2 * (i + 1)) {
    hipEvent_t e;
    JRX_CHECK_HIP(hipEventCreate(&e));
    p->events.push_back(e);
  }
  JRX_CHECK_HIP(hipEventRecord(p->events[2 * i], s));
  *slot = i;
  return JRX_OK;
}

static int timing_end(jrx_plan* p, hipStream_t s, int slot, int cls) {
  if (slot < 0) return JRX_OK;
  JRX_CHECK_HIP(hipEventRecord(p->events[2 * slot + 1], s));
  p->records.push_back({cls, slot});
  return JRX_OK;
}

#define JRX_T(cls, x)                                   \
  do {                                                  \
    if ((rc = timing_begin(p, s, &slot, timed))) return rc; \
    if ((rc = (x))) return rc;                          \
    if ((rc = timing_end(p, s, slot, cls))) return rc;  \
  } while (0)

// v = dE/drho of rho (Hartree + external + xc, the vol/N weight included),
// both in the plan's layout (z-major on the row path, whose box axes are
// (z, y, x)).  The Hartree/external partials go to the workspace's eh/ee
// (*nrec of them) and the xc partials to exc (*nexc).  ks: the Kohn-Sham
// effective potential of the band driver, which differs from dE/drho only
// in the spin-polarised LDA exchange term.
static int potential_stage(jrx_plan* p, double* rho, double* v, char* ws,
                           int* nrec, int* nexc, hipStream_t s, bool timed,
                           bool ks = false) {
  int rc, slot;
  auto cptr = [&](size_t o) { return reinterpret_cast<double2*>(ws + o); };
  auto dptr = [&](size_t o) { return reinterpret_cast<double*>(ws + o); };
  double2* box = cptr(p->o_box);
  double* eh = dptr(p->o_eh);
  double* ee = dptr(p->o_ee);
  double* exc = dptr(p->o_exc);
  const CellParams& cpb = p->rowpath ? p->cpz : p->cp;
  const FftLen& L0 = p->rowpath ? p->Lz : p->Lx;
  const FftLen& L2 = p->rowpath ? p->Lx : p->Lz;
  JRX_T(JRX_K_POTENTIAL, fft3d_real_forward(rho, p->ns, box, cpb.nx, cpb.ny,
                                            cpb.nz, L0, p->Ly, L2, s));
  double2* gb = p->o_gb ? cptr(p->o_gb) : nullptr;
  if (gb)  // density gradients of the GGA terms, from fftn(rho)
    JRX_T(JRX_K_POTENTIAL,
          xc_stage_gradients(box, gb, cpb, L0, p->Ly, L2, s));
  JRX_T(JRX_K_POTENTIAL,
        launch_potential_rec(box, box, eh, ee, cpb, nrec, s, p->d_sterm));
  JRX_T(JRX_K_POTENTIAL, fft3d(box, box, 1, cpb.nx, cpb.ny, cpb.nz, true,
                               1.0 / (double)p->N, L0, p->Ly, L2, s));
  if (p->xc == JRX_XC_LDA_X)
    JRX_T(JRX_K_POTENTIAL, launch_potential_real(box, rho, v, exc, p->ns,
                                                 p->N, p->vol, p->cx, nexc,
                                                 s, ks));
  else
    JRX_T(JRX_K_POTENTIAL,
          xc_stage_potential(box, rho, v, gb, exc, nexc, cpb, p->xc, L0,
                             p->Ly, L2, s));
  return JRX_OK;
}

// energies = {sum f tb / 2, sum f pot / vol, 0, 0, their sum}: the band
// energies' trace sum f <psi|T + v|psi> of the Kohn-Sham band mode.
__global__ __launch_bounds__(256) void band_trace_kernel(
    const double* __restrict__ tb, const double* __restrict__ pot,
    const double* __restrict__ occ, int64_t S, double inv_vol,
    double* __restrict__ energies) {
  __shared__ double red[4];
  double a = 0.0, b = 0.0;
  for (int64_t i = threadIdx.x; i < S; i += 256) {
    a += occ[i] * tb[i];
    b += occ[i] * pot[i];
  }
  const double k = 0.5 * block_sum(a, red);
  const double u = inv_vol * block_sum(b, red);
  if (threadIdx.x == 0) {
    energies[0] = k;
    energies[1] = u;
    energies[2] = 0.0;
    energies[3] = 0.0;
    energies[4] = k + u;
  }
}

// One evaluation into the given workspace; `timed` allows the per-kernel
// events of jrx_plan_set_timing (their first use creates HIP events).
static int energy_grad_impl(jrx_plan* p, const double* w_re,
                            const double* w_im, double* energies,
                            double* dw_re, double* dw_im, double* rho_out,
                            double* q_out, void* workspace,
                            size_t workspace_bytes, void* stream, bool timed) {
  JRX_REQUIRE(p && w_re && w_im && energies, "null argument");
  JRX_REQUIRE((dw_re == nullptr) == (dw_im == nullptr),
              "dw_re and dw_im must both be given or both be NULL");
  JRX_REQUIRE(workspace && workspace_bytes >= p->total,
              "workspace too small (see jrx_plan_workspace_bytes)");
  JRX_REQUIRE(((uintptr_t)workspace & 255) == 0,
              "workspace must be 256-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  char* ws = static_cast<char*>(workspace);
  auto cptr = [&](size_t o) { return reinterpret_cast<double2*>(ws + o); };
  auto dptr = [&](size_t o) { return reinterpret_cast<double*>(ws + o); };
  double2* q = cptr(p->o_q);
  double2* qbar = cptr(p->o_qbar);
  double2* T = cptr(p->o_T);
  double* rho = dptr(p->o_rho);
  double* rpart = dptr(p->o_rpart);
  double* v = dptr(p->o_v);
  double2* Sm = cptr(p->o_S);
  double2* R1 = cptr(p->o_R1);
  double2* R2 = cptr(p->o_R2);
  double2* Ri = cptr(p->o_Ri);
  double2* Am = cptr(p->o_A);
  double2* Bm = cptr(p->o_Bm);
  double2* part = cptr(p->o_part);
  double* kin = dptr(p->o_kin);
  double* tb = dptr(p->o_tb);
  double* eh = dptr(p->o_eh);
  double* ee = dptr(p->o_ee);
  double* exc = dptr(p->o_exc);
  const int nb = p->nb, ng = p->ng;
  const int64_t P = p->P, S = p->S;
  int rc;
  int slot;
  const bool fixed = p->fixed;
  JRX_REQUIRE(!(fixed && rho_out),
              "Kohn-Sham band mode forms no density (rho_out must be NULL)");
  // --- QR (CholeskyQR2; adaptive: the second pass only for the k-points
  // whose first Gram matrix is far from diagonal) --------------------------
  int* one_pass = p->qr_adaptive
                      ? reinterpret_cast<int*>(
                            static_cast<char*>(workspace) + p->o_qrflag)
                      : nullptr;
  JRX_T(JRX_K_QR_GRAM, launch_gram_gmajor(w_re, w_im, nullptr, part, P, ng, nb, p->gsw, s));
  JRX_T(JRX_K_QR_SMALL, launch_reduce_partials(part, Sm, P, p->gsw.nchunk, nb, s));
  JRX_T(JRX_K_QR_SMALL, launch_chol_inv(Sm, R1, P, nb, s, one_pass));
  JRX_T(JRX_K_QR_GRAM, launch_gram_gmajor(w_re, w_im, R1, part, P, ng, nb, p->gsw, s, one_pass, q));
  JRX_T(JRX_K_QR_SMALL, launch_reduce_partials(part, Sm, P, p->gsw.nchunk, nb, s));
  if (nb <= 32) {  // Ri = R1^{-1} R2^{-1} formed by the factorisation
    JRX_T(JRX_K_QR_SMALL, launch_chol_inv(Sm, Ri, P, nb, s, nullptr, one_pass, R1));
  } else {
    JRX_T(JRX_K_QR_SMALL, launch_chol_inv(Sm, R2, P, nb, s, nullptr, one_pass));
    JRX_T(JRX_K_QR_SMALL, launch_trmm_uu(R1, R2, Ri, P, nb, s));
  }
  JRX_T(JRX_K_QR_APPLY, launch_apply_fwd(w_re, w_im, Ri, q, P, ng, nb, s));
  // --- density -------------------------------------------------------------
  const int64_t SS = (int64_t)p->nk * nb;  // states per spin
  const bool reg64 = use_reg64(p->geo) && !p->force_generic;
  const int64_t plane_elems =
      p->rowpath ? (int64_t)p->nz * p->geo.rstride
                 : (int64_t)p->geo.nxa * p->geo.ps;
  int nrec = 0, nexc = 0, nkin = 0;
  if (!fixed) {
    for (int sp = 0; sp < p->ns; ++sp) {
      const int64_t step = p->keep_planes ? SS : p->chunk;
      // the first chunk stores every group's partial when it runs all
      // p->groups groups; otherwise the partials start from zero
      const int first = (int)(SS < step ? SS : step);
      const bool store_first =
          (p->rowpath ? zplane_groups(p->geo, first)
                      : xline_groups(p->geo, first, p->xline_blocks)) ==
          p->groups;
      if (!store_first)
        JRX_T(JRX_K_OTHER, launch_fill(rpart, 0.0, (int64_t)p->groups * p->N, s));
      for (int64_t s0 = 0; s0 < SS; s0 += step) {
        const int nst = (int)((SS - s0) < step ? (SS - s0) : step);
        const int64_t st0 = sp * SS + s0;
        double2* Tc = p->keep_planes ? T + st0 * plane_elems : T;
        if (p->rowpath) {
          JRX_T(JRX_K_PLANE_IFFT,
                launch_row_ifft(q, Tc, st0, nst, p->geo, p->Lx.tw, s));
          JRX_T(JRX_K_XLINE_DENSITY,
                launch_zplane_density(Tc, rpart, zplane_groups(p->geo, nst),
                                      p->d_occ, st0, nst, 1.0 / p->vol,
                                      p->geo, p->Lx.tw,
                                      store_first && s0 == 0, s));
        } else if (reg64) {
          JRX_T(JRX_K_PLANE_IFFT, launch_plane_ifft64(q, Tc, st0, nst, p->geo,
                                                      p->Lx.tw, s));
          JRX_T(JRX_K_XLINE_DENSITY,
                launch_xline_density64(Tc, rpart, xline_groups(p->geo, nst, p->xline_blocks),
                                       p->d_occ, st0, nst, 1.0 / p->vol,
                                       p->geo, p->Lx.tw,
                                       store_first && s0 == 0, s));
        } else {
          JRX_T(JRX_K_PLANE_IFFT,
                launch_plane_ifft(q, Tc, st0, nst, p->geo, p->Ly, p->Lz, s));
          JRX_T(JRX_K_XLINE_DENSITY,
                launch_xline_density(Tc, rpart, xline_groups(p->geo, nst, p->xline_blocks),
                                     p->d_occ, st0, nst, 1.0 / p->vol, p->geo,
                                     p->Lx, store_first && s0 == 0, s));
        }
      }
      JRX_T(JRX_K_XLINE_DENSITY,
            launch_reduce_groups(rpart, rho + sp * p->N, p->groups, p->N, s));
    }
    // --- reciprocal energies and potential --------------------------------
    if ((rc = potential_stage(p, rho, v, ws, &nrec, &nexc, s, timed)))
      return rc;
  }
  const double* vv = fixed ? p->d_vfix : v;
  // the density pass left every state's planes in T (keep_planes)
  const bool planes_ready = p->keep_planes && !fixed;
  const bool pot_e = p->band_e || fixed;
  if (dw_re || fixed) {
    // --- adjoint sweep: qbar = f (2/vol FFT(v psi) + |G+k|^2 q) -------------
    GatherParams gp;
    gp.occ = p->d_occ;
    gp.kpts = p->d_kpts;
    for (int i = 0; i < 9; ++i) gp.B[i] = p->B[i];
    gp.nk = p->nk;
    gp.nb = nb;
    gp.two_over_vol = 2.0 / p->vol;
    gp.pot_partial = pot_e ? dptr(p->o_bepart) : nullptr;
    gp.g_m = p->d_g_m;
    const int64_t step = p->keep_planes ? SS : p->chunk;
    for (int sp = 0; sp < p->ns; ++sp) {
      for (int64_t s0 = 0; s0 < SS; s0 += step) {
        const int nst = (int)((SS - s0) < step ? (SS - s0) : step);
        const int64_t st0 = sp * SS + s0;
        double2* Tc = p->keep_planes ? T + st0 * plane_elems : T;
        if (p->rowpath) {
          if (!planes_ready)
            JRX_T(JRX_K_PLANE_IFFT,
                  launch_row_ifft(q, Tc, st0, nst, p->geo, p->Lx.tw, s));
          JRX_T(JRX_K_XLINE_VPSI,
                launch_zplane_vpsi(Tc, vv + sp * p->N,
                                   zplane_groups(p->geo, nst), nst, p->geo,
                                   p->Lx.tw, s));
          JRX_T(JRX_K_PLANE_FFT_GATHER,
                launch_row_fft_gather(Tc, q, qbar, kin, st0, nst, p->geo,
                                      p->Lx.tw, gp, s));
        } else if (reg64) {
          if (!planes_ready)
            JRX_T(JRX_K_PLANE_IFFT, launch_plane_ifft64(q, Tc, st0, nst,
                                                        p->geo, p->Lx.tw, s));
          JRX_T(JRX_K_XLINE_VPSI,
                launch_xline_vpsi64(Tc, vv + sp * p->N,
                                    xline_groups(p->geo, nst, p->xline_blocks), nst, p->geo,
                                    p->Lx.tw, s));
          JRX_T(JRX_K_PLANE_FFT_GATHER,
                launch_plane_fft_gather64(Tc, q, qbar, kin, st0, nst, p->geo,
                                          p->Lx.tw, gp, s));
        } else {
          if (!planes_ready)
            JRX_T(JRX_K_PLANE_IFFT, launch_plane_ifft(q, Tc, st0, nst, p->geo,
                                                      p->Ly, p->Lz, s));
          JRX_T(JRX_K_XLINE_VPSI,
                launch_xline_vpsi(Tc, vv + sp * p->N, xline_groups(p->geo, nst, p->xline_blocks),
                                  nst, p->geo, p->Lx, s));
          JRX_T(JRX_K_PLANE_FFT_GATHER,
                launch_plane_fft_gather(Tc, q, qbar, kin, st0, nst, p->geo,
                                        p->Ly, p->Lz, gp, s));
        }
      }
    }
    const int nper = p->rowpath ? p->row_chunks : p->geo.nxa;
    JRX_T(JRX_K_OTHER, launch_state_sums(kin, nper, S, tb, s));
    if (pot_e)
      JRX_T(JRX_K_OTHER, launch_state_sums(dptr(p->o_bepart), nper, S,
                                           dptr(p->o_pot), s));
    // --- non-local pseudopotential: energy per band and its qbar term ------
    if (p->nproj > 0)
      JRX_T(JRX_K_OTHER, launch_nonlocal(p->d_proj, q, p->d_occ, P, p->nk,
                                         p->nproj, ng, nb, p->vol, qbar,
                                         p->d_nlscratch, p->d_enl, s));
    // --- QR adjoint (Q^H Qbar alone without a gradient: the subspace
    // matrix of jrx_plan_subspace_eigenvalues) --------------------------------
    JRX_T(JRX_K_QR_BWD_GRAM, launch_gram_bmajor(q, qbar, part, P, ng, nb, p->gsq, s));
    JRX_T(JRX_K_QR_SMALL, launch_reduce_partials(part, Sm, P, p->gsq.nchunk, nb, s));
    if (dw_re) {
      JRX_T(JRX_K_QR_SMALL, launch_qr_bwd_small(Sm, Ri, Am, Bm, P, nb, s));
      JRX_T(JRX_K_QR_BWD_APPLY, launch_apply_bwd(qbar, q, Am, Bm, dw_re, dw_im, P, ng, nb, s));
    }
  } else {
    JRX_T(JRX_K_OTHER, launch_kinetic(q, p->d_occ, p->d_kpts, p->d_g_m, p->cp, S, ng,
                           p->nk, nb, kin, &nkin, s));
    JRX_T(JRX_K_OTHER, launch_state_sums(kin, (int)(nkin / S), S, tb, s));
    if (p->nproj > 0)
      JRX_T(JRX_K_OTHER, launch_nonlocal(p->d_proj, q, p->d_occ, P, p->nk,
                                         p->nproj, ng, nb, p->vol, nullptr,
                                         p->d_nlscratch, p->d_enl, s));
  }
  if (fixed) {
    hipLaunchKernelGGL(band_trace_kernel, dim3(1), dim3(256), 0, s, tb,
                       dptr(p->o_pot), p->d_occ, S, 1.0 / p->vol, energies);
    JRX_LAUNCH_CHECK();
  } else {
    JRX_T(JRX_K_OTHER, launch_finalize_energies(tb, p->d_occ, S, eh, ee, nrec,
                                                exc, nexc, energies, s));
  }
  if (p->nproj > 0)  // the non-local energy joins the external term
    JRX_T(JRX_K_OTHER, launch_nonlocal_finalize(p->d_enl, p->d_occ, S,
                                                energies, s));
  if (rho_out && p->rowpath) {
    if ((rc = launch_zmajor_to_xmajor(rho, rho_out, p->ns, p->nx, s)))
      return rc;
  } else if (rho_out) {
    JRX_CHECK_HIP(hipMemcpyAsync(rho_out, rho,
                                 (size_t)p->ns * p->N * sizeof(double),
                                 hipMemcpyDeviceToDevice, s));
  }
  if (q_out)
    JRX_CHECK_HIP(hipMemcpyAsync(q_out, q, (size_t)S * ng * sizeof(double2),
                                 hipMemcpyDeviceToDevice, s));
#undef JRX_T
  return JRX_OK;
}

int jrx_energy_grad(jrx_plan* p, const double* w_re, const double* w_im,
                    double* energies, double* dw_re, double* dw_im,
                    double* rho_out, double* q_out, void* workspace,
                    size_t workspace_bytes, void* stream) {
  return energy_grad_impl(p, w_re, w_im, energies, dw_re, dw_im, rho_out,
                          q_out, workspace, workspace_bytes, stream, true);
}

int jrx_energy_grad_static(jrx_plan* p, const double* w_re,
                           const double* w_im, double* energies,
                           double* dw_re, double* dw_im, double* rho_out,
                           double* q_out, void* stream) {
  JRX_REQUIRE(p, "null plan");
  if (p->total > JRX_STATIC_WORKSPACE_BYTES) {
    set_error("plan workspace of " + std::to_string(p->total) +
              " bytes exceeds the static workspace of " +
              std::to_string((size_t)JRX_STATIC_WORKSPACE_BYTES) +
              " bytes (use jrx_energy_grad with a caller workspace)");
    return JRX_ERR_UNSUPPORTED;
  }
  // The address was looked up when the plan was created, on its device.
  JRX_REQUIRE(p->static_ws, "static workspace address unknown");
  return energy_grad_impl(p, w_re, w_im, energies, dw_re, dw_im, rho_out,
                          q_out, p->static_ws, JRX_STATIC_WORKSPACE_BYTES,
                          stream, false);
}

int jrx_plan_set_fixed_density(jrx_plan* p, const double* rho,
                               void* workspace, size_t workspace_bytes,
                               void* stream) {
  JRX_REQUIRE(p, "null plan");
  if (!rho) {
    p->fixed = false;
    return JRX_OK;
  }
  JRX_REQUIRE(workspace && workspace_bytes >= p->total,
              "workspace too small (see jrx_plan_workspace_bytes)");
  JRX_REQUIRE(((uintptr_t)workspace & 255) == 0,
              "workspace must be 256-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  const size_t bytes = (size_t)p->ns * p->N * sizeof(double);
  if (!p->d_vfix) {
    double* d = nullptr;
    JRX_CHECK_HIP(hipMalloc(&d, bytes));
    p->allocs.push_back(d);
    p->d_vfix = d;
  }
  char* ws = static_cast<char*>(workspace);
  double* r = reinterpret_cast<double*>(ws + p->o_rho);
  int rc;
  if (p->rowpath) {
    // (x, y, z) -> the row path's (z, y, x): on its cubic box the axis swap
    // is its own inverse
    if ((rc = launch_zmajor_to_xmajor(rho, r, p->ns, p->nx, s))) return rc;
  } else {
    JRX_CHECK_HIP(hipMemcpyAsync(r, rho, bytes, hipMemcpyDeviceToDevice, s));
  }
  int nrec = 0, nexc = 0;
  if ((rc = potential_stage(p, r, p->d_vfix, ws, &nrec, &nexc, s, false,
                            true)))
    return rc;
  p->fixed = true;
  return JRX_OK;
}

int jrx_plan_set_kpoints(jrx_plan* p, const double* kpts, void* stream) {
  JRX_REQUIRE(p && kpts, "null argument");
  JRX_CHECK_HIP(hipMemcpyAsync(p->d_kpts, kpts, (size_t)p->nk * 3 * sizeof(double),
                               hipMemcpyDefault, (hipStream_t)stream));
  return JRX_OK;
}

int jrx_plan_subspace_eigenvalues(const jrx_plan* p, double* evals,
                                  double* vecs, void* workspace,
                                  size_t workspace_bytes, void* stream) {
  JRX_REQUIRE(p && evals && workspace, "null argument");
  JRX_REQUIRE(workspace_bytes >= p->total, "workspace too small");
  char* ws = static_cast<char*>(workspace);
  return launch_ritz(reinterpret_cast<const double2*>(ws + p->o_S), p->d_occ,
                     p->P, p->nb, reinterpret_cast<double2*>(vecs), evals,
                     true, ws + p->o_ritz, (hipStream_t)stream);
}

int jrx_plan_subspace_hamiltonian(const jrx_plan* p, double* h,
                                  const void* workspace,
                                  size_t workspace_bytes, void* stream) {
  JRX_REQUIRE(p && h && workspace, "null argument");
  JRX_REQUIRE(workspace_bytes >= p->total, "workspace too small");
  const char* ws = static_cast<const char*>(workspace);
  return launch_subspace_h(reinterpret_cast<const double2*>(ws + p->o_S),
                           p->d_occ, p->P, p->nb, reinterpret_cast<double2*>(h),
                           (hipStream_t)stream);
}

// sterm[i] = conj(v[x-major index of i]) * scale on the plan's box (on the
// row path the box axes are (z, y, x) of a cubic n^3 box).
__global__ void vloc_to_sterm_kernel(const double2* __restrict__ v,
                                     double2* __restrict__ sterm, int64_t N,
                                     int n, int zmajor, double scale) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int64_t src = i;
  if (zmajor) {
    const int64_t a = i / ((int64_t)n * n), rem = i - a * n * n;
    const int64_t b = rem / n, c = rem - b * n;  // i = (a, b, c) = (z, y, x)
    src = (c * n + b) * n + a;
  }
  const double2 t = v[src];
  sterm[i] = make_double2(t.x * scale, -t.y * scale);
}

static int realloc_dev(jrx_plan* p, void** ptr, size_t* cap, size_t bytes) {
  if (bytes <= *cap && *ptr) return JRX_OK;
  if (*ptr) {
    for (size_t i = 0; i < p->allocs.size(); ++i)
      if (p->allocs[i] == *ptr) {
        p->allocs.erase(p->allocs.begin() + i);
        break;
      }
    JRX_CHECK_HIP(hipFree(*ptr));
    *ptr = nullptr;
  }
  JRX_CHECK_HIP(hipMalloc(ptr, bytes));
  p->allocs.push_back(*ptr);
  *cap = bytes;
  return JRX_OK;
}

int jrx_plan_set_external_potential(jrx_plan* p, const double* v_rec,
                                    void* stream) {
  JRX_REQUIRE(p, "null plan");
  hipStream_t s = (hipStream_t)stream;
  if (!v_rec) {  // back to the Coulomb nuclei
    int rc = launch_structure_term(p->rowpath ? p->cpz : p->cp, p->d_sterm, s);
    if (rc) return rc;
    p->custom_vloc = false;
    return JRX_OK;
  }
  const double scale = -p->vol / (4.0 * M_PI * (double)p->N);
  hipLaunchKernelGGL(vloc_to_sterm_kernel, dim3(ceil_div(p->N, 256)),
                     dim3(256), 0, s, reinterpret_cast<const double2*>(v_rec),
                     p->d_sterm, p->N, p->nx, p->rowpath ? 1 : 0, scale);
  JRX_LAUNCH_CHECK();
  p->custom_vloc = true;
  return JRX_OK;
}

int jrx_plan_set_nonlocal(jrx_plan* p, int nproj, const double* proj,
                          void* stream) {
  JRX_REQUIRE(p && nproj >= 0, "invalid argument");
  if (nproj == 0) {
    p->nproj = 0;
    return JRX_OK;
  }
  JRX_REQUIRE(proj, "null projectors");
  JRX_REQUIRE(p->nb <= 128, "non-local projectors: more than 128 bands");
  JRX_REQUIRE((int64_t)nproj * p->nb <= 4096,
              "non-local projectors: projectors x bands above 4096");
  int rc;
  const size_t pbytes = (size_t)p->nk * nproj * p->ng * sizeof(double2);
  if ((rc = realloc_dev(p, reinterpret_cast<void**>(&p->d_proj), &p->nl_cap,
                        pbytes)))
    return rc;
  size_t sc = nl_scratch_bytes(p->P, nproj, p->nb, p->ng);
  if ((rc = realloc_dev(p, &p->d_nlscratch, &p->nl_scratch_cap, sc)))
    return rc;
  if (!p->d_enl) {
    size_t cap = 0;
    void* e = nullptr;
    if ((rc = realloc_dev(p, &e, &cap, (size_t)p->S * sizeof(double))))
      return rc;
    p->d_enl = static_cast<double*>(e);
  }
  JRX_CHECK_HIP(hipMemcpyAsync(p->d_proj, proj, pbytes, hipMemcpyDefault,
                               (hipStream_t)stream));
  p->nproj = nproj;
  return JRX_OK;
}

int jrx_plan_nonlocal_energies(const jrx_plan* p, double* out, void* stream) {
  JRX_REQUIRE(p && out, "null argument");
  JRX_REQUIRE(p->nproj > 0, "no non-local projectors (jrx_plan_set_nonlocal)");
  JRX_CHECK_HIP(hipMemcpyAsync(out, p->d_enl, (size_t)p->S * sizeof(double),
                               hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return JRX_OK;
}

}  // extern "C"
