// osbl_runtime.hip -- host runtime and C-ABI of libosbl_hip.so.
//
// A context owns the block's datasets on one device and a HIP stream; the
// time loop of the generated program (algorithm.py:423-477) is replayed as a
// sequence of stream-ordered launches per RK stage:
//   boundary conditions -> [constituent relations] -> characteristic sweeps
//   -> fused central/viscous residual + RK update -> boundary conditions.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include "osbl_kernels.h"

using namespace osbl;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

}  // namespace

int osbl::set_error(int code, const char* msg) {
  g_err = msg;
  return code;
}

namespace {

#define HIPCHK(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) return fail(OSBL_E_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

enum Family { F_BC = 0, F_PRIMS, F_SWEEP_X, F_SWEEP_Y, F_SWEEP_Z, F_UPDATE, F_FILTER, F_CENTRAL, F_COUNT };
const char* kFamilyNames[F_COUNT] = {"boundary", "constituent_relations", "char_sweep_x", "char_sweep_y",
                                     "char_sweep_z", "residual_rk_update", "filters", "central_convective"};

int rk_coefficients(int rk, double* a, double* b) {
  switch (rk) {
    case OSBL_RK3:  // rknew / rkold, rk_sbli.py:58-60
      a[0] = 2.0 / 3.0; a[1] = 5.0 / 12.0; a[2] = 3.0 / 5.0;
      b[0] = 1.0 / 4.0; b[1] = 3.0 / 20.0; b[2] = 3.0 / 5.0;
      return 3;
    case OSBL_RK3_LS:  // rk_LS.py:91-94
      a[0] = 0.0; a[1] = -5.0 / 9.0; a[2] = -153.0 / 128.0;
      b[0] = 1.0 / 3.0; b[1] = 15.0 / 16.0; b[2] = 8.0 / 15.0;
      return 3;
    case OSBL_RK3_LS_SSP: {  // rk_LS.py:74-89 (Gottlieb-Shu-Tadmor c = 0.924574)
      const double c = 0.924574;
      const double z1 = sqrt(36 * c * c * c * c + 36 * c * c * c - 135 * c * c + 84 * c - 12);
      const double z2 = 2 * c * c + c - 2;
      const double z3 = 12 * c * c * c * c - 18 * c * c * c + 18 * c * c - 11 * c + 2;
      const double z4 = 36 * c * c * c * c - 36 * c * c * c + 13 * c * c - 8 * c + 4;
      const double z5 = 69 * c * c * c - 62 * c * c + 28 * c - 8;
      const double z6 = 34 * c * c * c * c - 46 * c * c * c + 34 * c * c - 13 * c + 2;
      const double y = 3 * z2 - z1;
      b[0] = c;
      b[1] = (12 * c * (c - 1) * y - y * y) / (144 * c * (3 * c - 2) * (c - 1) * (c - 1));
      b[2] = (-24 * (3 * c - 2) * (c - 1) * (c - 1)) / (y * y - 12 * c * (c - 1) * y);
      a[0] = 0.0;
      a[1] = (-z1 * (6 * c * c - 4 * c + 1) + 3 * z3) / ((2 * c + 1) * z1 - 3 * (c + 2) * (2 * c - 1) * (2 * c - 1));
      a[2] = (-z1 * z4 + 108 * (2 * c - 1) * pow(c, 5) - 3 * (2 * c - 1) * z5) /
             (24 * z1 * c * pow(c - 1, 4) + 72 * c * z6 + 72 * pow(c, 6) * (2 * c - 13));
      return 3;
    }
    case OSBL_RK4_LS: {  // rk_LS.py:96-97
      const double A[5] = {0, -0.4178904745, -1.192151694643, -1.697784692471, -1.514183444257};
      const double B[5] = {0.1496590219993, 0.3792103129999, 0.8229550293869, 0.6994504559488, 0.1530572479681};
      for (int i = 0; i < 5; ++i) { a[i] = A[i]; b[i] = B[i]; }
      return 5;
    }
    default:
      return 0;
  }
}

bool is_central(int conv) { return conv == OSBL_CONV_CENTRAL_SKEW || conv == OSBL_CONV_CENTRAL_CONS; }

// Closure rows of the central stencils, laid out as Bnd::cw.  Carpenter's
// first-derivative rows are bc4 = inv(al4) ar4 with the R1, R2 constants of
// Carpenter_scheme.py:80-102 (solved here in long double); ReducedAccess rows
// reduced_access_scheme.py:34-40; second-derivative rows of both schemes
// Carpenter_scheme.py:69-77.
void closure_table(double* cw) {
  const long double r1 = -(2177.0L * sqrtl(295369.0L) - 1166427.0L) / 25488.0L;
  const long double r2 = (66195.0L * sqrtl(53.0L) * sqrtl(5573.0L) - 35909375.0L) / 101952.0L;
  long double a[4][10] = {
      {-(216 * r2 + 2160 * r1 - 2125) / 12960, (81 * r2 + 675 * r1 + 415) / 540, -(72 * r2 + 720 * r1 + 445) / 1440,
       -(108 * r2 + 756 * r1 + 421) / 1296, -0.5L, -(864 * r2 + 6480 * r1 + 305) / 4320,
       (216 * r2 + 1620 * r1 + 725) / 540, -(864 * r2 + 6480 * r1 + 3335) / 4320, 0, 0},
      {(81 * r2 + 675 * r1 + 415) / 540, -(4104 * r2 + 32400 * r1 + 11225) / 4320,
       (1836 * r2 + 14580 * r1 + 7295) / 2160, -(216 * r2 + 2160 * r1 + 655) / 4320, (864 * r2 + 6480 * r1 + 305) / 4320,
       0, -(864 * r2 + 6480 * r1 + 2315) / 1440, (108 * r2 + 810 * r1 + 415) / 270, 0, 0},
      {-(72 * r2 + 720 * r1 + 445) / 1440, (1836 * r2 + 14580 * r1 + 7295) / 2160,
       -(4104 * r2 + 32400 * r1 + 12785) / 4320, (81 * r2 + 675 * r1 + 335) / 540, -(216 * r2 + 1620 * r1 + 725) / 540,
       (864 * r2 + 6480 * r1 + 2315) / 1440, 0, -(864 * r2 + 6480 * r1 + 785) / 4320, -1.0L / 12, 0},
      {-(108 * r2 + 756 * r1 + 421) / 1296, -(216 * r2 + 2160 * r1 + 655) / 4320, (81 * r2 + 675 * r1 + 335) / 540,
       -(216 * r2 + 2160 * r1 - 12085) / 12960, (864 * r2 + 6480 * r1 + 3335) / 4320, -(108 * r2 + 810 * r1 + 415) / 270,
       (864 * r2 + 6480 * r1 + 785) / 4320, 0, 8.0L / 12, -1.0L / 12}};
  // Gauss-Jordan elimination of the 4x4 left block over the augmented rows
  for (int col = 0; col < 4; ++col) {
    int piv = col;
    for (int r = col + 1; r < 4; ++r)
      if (fabsl(a[r][col]) > fabsl(a[piv][col])) piv = r;
    for (int k = 0; k < 10; ++k) std::swap(a[col][k], a[piv][k]);
    const long double inv = 1.0L / a[col][col];
    for (int k = 0; k < 10; ++k) a[col][k] *= inv;
    for (int r = 0; r < 4; ++r) {
      if (r == col) continue;
      const long double f = a[r][col];
      for (int k = 0; k < 10; ++k) a[r][k] -= f * a[col][k];
    }
  }
  for (int r = 0; r < 4; ++r)
    for (int k = 0; k < 6; ++k) cw[CW_CARP + 6 * r + k] = (double)a[r][4 + k];
  const double ra[2][6] = {{-25.0 / 12.0, 48.0 / 12.0, -36.0 / 12.0, 16.0 / 12.0, -3.0 / 12.0, 0.0},
                           {-3.0 / 12.0, -10.0 / 12.0, 18.0 / 12.0, -6.0 / 12.0, 1.0 / 12.0, 0.0}};
  const double d2[2][6] = {{35.0 / 12.0, -104.0 / 12.0, 114.0 / 12.0, -56.0 / 12.0, 11.0 / 12.0, 0.0},
                           {11.0 / 12.0, -20.0 / 12.0, 6.0 / 12.0, 4.0 / 12.0, -1.0 / 12.0, 0.0}};
  for (int r = 0; r < 2; ++r)
    for (int k = 0; k < 6; ++k) {
      cw[CW_RA + 6 * r + k] = ra[r][k];
      cw[CW_D2 + 6 * r + k] = d2[r][k];
    }
}

// first derivative (without 1/Delta) of a line of N values at index c with
// closures lo/hi (1 + OSBL_BS_*, or 0), the host twin of the device stencil
double d1_line(const double* f, int N, int lo, int hi, int c, const double* cw) {
  if (lo) {
    const int nr = lo == 1 ? 4 : 2, np = lo == 1 ? 6 : 5;
    if (c < nr) {
      const double* w = cw + (lo == 1 ? CW_CARP : CW_RA) + 6 * c;
      double acc = 0.0;
      for (int k = 0; k < np; ++k) acc += w[k] * f[c + (k - c)];
      return acc;
    }
  }
  if (hi) {
    const int nr = hi == 1 ? 4 : 2, np = hi == 1 ? 6 : 5;
    const int r = N - 1 - c;
    if (r < nr) {
      const double* w = cw + (hi == 1 ? CW_CARP : CW_RA) + 6 * r;
      double acc = 0.0;
      for (int k = 0; k < np; ++k) acc += w[k] * f[c + (r - k)];
      return -acc;
    }
  }
  return (1.0 / 12.0) * (f[c - 2] - f[c + 2]) + (2.0 / 3.0) * (f[c + 1] - f[c - 1]);
}

bool modifies_stencil(int bc) {
  return bc != OSBL_BC_PERIODIC && bc != OSBL_BC_ZERO_GRADIENT && bc != OSBL_BC_SYMMETRY;
}

}  // namespace

struct osbl_ctx {
  osbl_config cfg{};
  int nd = 0, nv = 0;
  Grid g{};
  Phys ph{};
  int device = 0;
  hipStream_t stream = nullptr;
  double* Qbuf[2] = {nullptr, nullptr};
  int cur = 0;
  double* Qo = nullptr;
  double* R = nullptr;
  double* P = nullptr;
  double* stage_buf = nullptr;  // contiguous interior staging for host copies
  double* Theta = nullptr;      // Ducros sensor (adaptive TENO, WENO filter)
  double* QF = nullptr;         // SFD filtered solution
  bool sfd_set[5] = {};         // vfilt of variable m holds a value
  unsigned char* fmask = nullptr;  // binomial filter grid condition (interior cells)
  double binom_w[11]{};
  int walls = 0;                // wall faces, bit 2d+side
  double sfd_c = 0.0, sfd_E = 0.0;
  int nstages = 0;
  double A[5]{}, B[5]{};
  int seg = 32;
  int vsz = 32;             // z segment of the tiled viscous kernels
  int csz = 32;             // z segment of the tiled central convective kernel
  bool need_prims = false;
  bool tiled_visc = false;  // 3-D characteristic scheme with viscous terms
  bool tiled_gen = false;   // the same on the general path (closures/metric/mu(T) in x and y only)
  bool tiled_central = false;  // 3-D periodic central scheme: tiled convective kernel (+ tiled viscous kernel)
  bool general = false;     // closures, metric or Sutherland viscosity: the general update kernel
  bool suth = false;
  Bnd bnd{};
  double* cw_dev = nullptr;                 // closure rows
  double* met_dev[3] = {nullptr, nullptr, nullptr};  // D_dd then SD_ddd, n[d] each
  double* plane_dev[3][2] = {};             // per-face planes
  bool curv = false;                        // general curvilinear metric
  double* cmet = nullptr;                   // its D_ab (a*nd + b) and detJ, padded fields
  double* cflux = nullptr;                  // interface fluxes of one direction (nv padded fields)
  double* cpt = nullptr;                    // per-point detJ F_d, U, a|k| of one direction (nv + 2 fields)
  bool cmet_ready = false;                  // osbl_set_grid_coordinates has run
  int nplane[3][2] = {};
  // profiling
  bool prof = false;
  struct Rec { int fam; hipEvent_t a, b; };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  double fam_ms[F_COUNT]{};
  long long fam_n[F_COUNT]{};
  // replay of two RK iterations from a hipGraph (one per starting buffer)
  bool use_graphs = true;
  hipGraphExec_t graph2[2] = {nullptr, nullptr};
  // adjoint work buffers (allocated on first use)
  double* Abuf = nullptr;  // stage state with halos
  double* Mbuf = nullptr;  // adjoint weights with halos
  double* pk = nullptr;    // 3 packed arrays: mu, l, lo
  double* vw = nullptr;    // viscous VJP fields (padded)
  double* vy = nullptr;    // y part of a 2-D WENO5-Z VJP (packed)
};

namespace {

hipEvent_t get_event(osbl_ctx* c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// bracket a launch with events when profiling
template <class F>
void timed(osbl_ctx* c, int fam, F&& launch) {
  if (!c->prof) {
    launch();
    return;
  }
  hipEvent_t a = get_event(c), b = get_event(c);
  (void)hipEventRecord(a, c->stream);
  launch();
  (void)hipEventRecord(b, c->stream);
  c->recs.push_back({fam, a, b});
}

void drain_records(osbl_ctx* c) {
  if (c->recs.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (auto& r : c->recs) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      c->fam_ms[r.fam] += ms;
      c->fam_n[r.fam] += 1;
    }
    c->pool.push_back(r.a);
    c->pool.push_back(r.b);
  }
  c->recs.clear();
}

double* Qcur(osbl_ctx* c) { return c->Qbuf[c->cur]; }

// halo depth a boundary kernel fills: the largest halo of the schemes,
// -3/+4 for WENO5/TENO (weno.py:17-29), -2/+2 for Central(4) (bc_core.py:136-147)
// halo points a non-periodic face fills: the block's boundary halos are those
// of its schemes (bc_core.py:145-166), the characteristic reconstruction's
// (3 below, 4 above) when the block has one -- the WENO filter included
int bc_halos(const osbl_ctx* c, int side) {
  const bool central = is_central(c->cfg.conv) && !(c->cfg.filters & OSBL_FILTER_WENO);
  return central ? 2 : (side == 0 ? 3 : 4);
}

void apply_bcs(osbl_ctx* c, double* Q) {
  bool all_periodic = true;
  for (int d = 0; d < c->nd; ++d) all_periodic &= c->cfg.bc[d][0] == OSBL_BC_PERIODIC;
  if (all_periodic) {
    // one pass that gives every halo point its periodic image
    timed(c, F_BC, [&] { launch_halo_periodic(Q, c->nv, c->g, c->stream); });
    return;
  }
  // the reference's order (block.py:199-210): direction by direction, side 0 then side 1
  double* Q2 = (Q == c->Qbuf[0]) ? c->Qbuf[1] : c->Qbuf[0];
  for (int d = 0; d < c->nd; ++d) {
    if (c->cfg.bc[d][0] == OSBL_BC_PERIODIC) {
      timed(c, F_BC, [&] { launch_periodic_dir(Q, c->nv, c->g, d, c->stream); });
      continue;
    }
    for (int s = 0; s < 2; ++s) {
      const int bc = c->cfg.bc[d][s];
      if (bc == OSBL_BC_EXTRAP0) {
        timed(c, F_BC, [&] { launch_extrap0(Q, c->nv, c->g, d, s, c->stream); });
      } else {
        timed(c, F_BC, [&] {
          launch_bc_face(c->nd, Q, Q2, c->g, c->ph, d, s, bc, bc_halos(c, s), c->cfg.twall, c->plane_dev[d][s],
                         c->stream);
        });
      }
    }
  }
}

// spatial residual of the current state into R (characteristic part only),
// followed by the fused central/viscous + update kernel with the given mode
void stage(osbl_ctx* c, int mode, double cA, double cB, int first) {
  double* Q = Qcur(c);
  double* Qn = c->Qbuf[1 - c->cur];
  const int conv = c->cfg.conv;
  if (c->need_prims) {
    // the tiled general schedule forms the constituent relations inside its
    // viscous kernel; only the near-face rows it leaves to k_visc_shell read P
    const bool slab = c->tiled_gen && mode != 2;
    bool shell = false;
    for (int d = 0; d < 2; ++d)
      for (int sd = 0; sd < 2; ++sd) shell |= c->bnd.mod[d][sd] != 0;
    if (!slab || shell) {
      timed(c, F_PRIMS, [&] {
        if (slab) launch_prims_slab(Q, c->P, c->g, c->ph, c->suth ? 1 : 0, c->stream);
        else if (c->suth) launch_prims_mu(c->nd, Q, c->P, c->g, c->ph, c->stream);
        else launch_prims(c->nd, Q, c->P, c->g, c->ph, c->stream);
      });
    }
  }
  if (c->Theta) timed(c, F_PRIMS, [&] { launch_sensor(c->nd, Q, c->Theta, c->g, c->ph, c->bnd, c->stream); });
  if (c->curv) {
    // strong conservation form: per direction the interface fluxes about the
    // metric normal and their difference, divided by detJ after the last one
    for (int d = 0; d < c->nd; ++d) {
      timed(c, F_SWEEP_X + d, [&] {
        launch_curv_direction(c->nd, d, conv, c->cfg.averaging, Q, c->cmet, c->cpt, c->cflux, c->R, c->g, c->ph,
                              c->cfg.delta[d], c->stream);
      });
    }
    if (mode != 2) {
      timed(c, F_UPDATE, [&] {
        launch_update(c->nd, conv, false, mode, Q, c->P, c->R, c->Qo, Qn, c->g, c->ph, cA, cB, first, c->stream);
      });
    }
    return;
  }
  if (c->tiled_central) {
    // 3-D central scheme: the tiled convective kernel applies the RK update,
    // or, with viscous terms, assigns R and the tiled viscous kernel adds its
    // terms and applies the update (mode 2: residual only, left in R)
    const bool visc = c->cfg.viscous != OSBL_VISC_NONE;
    timed(c, F_CENTRAL, [&] {
      launch_conv3d(conv, visc ? 2 : mode, Q, c->R, c->Qo, Qn, c->g, c->ph, cA, cB, first, c->csz, c->stream);
    });
    if (visc)
      timed(c, F_UPDATE, [&] {
        launch_visc_update3d(mode, Q, c->R, c->Qo, Qn, c->g, c->ph, cA, cB, first, c->vsz, c->stream);
      });
    return;
  }
  if (c->tiled_visc && mode != 2) {
    // 3-D characteristic scheme with viscous terms: the x sweep assigns R, the
    // viscous kernel and the y sweep accumulate, and the z sweep adds the last
    // difference and applies the RK update, so the update needs no pass of its own
    timed(c, F_SWEEP_X, [&] {
      launch_sweep(c->nd, 0, conv, c->cfg.averaging, Q, c->R, c->g, c->ph, c->seg, nullptr, c->Theta, false,
                   c->stream);
    });
    timed(c, F_UPDATE, [&] { launch_visc_update3d(2, Q, c->R, c->Qo, Qn, c->g, c->ph, cA, cB, first, c->vsz, c->stream); });
    timed(c, F_SWEEP_Y, [&] {
      launch_sweep(c->nd, 1, conv, c->cfg.averaging, Q, c->R, c->g, c->ph, c->seg, nullptr, c->Theta, true,
                   c->stream);
    });
    timed(c, F_SWEEP_Z, [&] {
      launch_sweep_z_rk(conv, c->cfg.averaging, mode, Q, c->R, c->Qo, Qn, c->g, c->ph, c->seg, c->Theta, cA, cB,
                        first, c->stream);
    });
    return;
  }
  if (c->tiled_gen && mode != 2) {
    // the same schedule on the general path: tiled viscous terms with closures,
    // metric and mu(T) (z free of both), sweeps with the metric, RK in the z sweep
    timed(c, F_SWEEP_X, [&] {
      launch_sweep(c->nd, 0, conv, c->cfg.averaging, Q, c->R, c->g, c->ph, c->seg, c->bnd.dm[0], c->Theta, false,
                   c->stream);
    });
    timed(c, F_UPDATE, [&] {
      launch_visc_gen3d(Q, c->P, c->R, c->g, c->ph, c->bnd, c->suth ? 1 : 0, c->vsz, c->stream);
      if (c->ph.has_force) launch_body_force3d(Q, c->R, c->g, c->ph, c->stream);
    });
    timed(c, F_SWEEP_Y, [&] {
      launch_sweep(c->nd, 1, conv, c->cfg.averaging, Q, c->R, c->g, c->ph, c->seg, c->bnd.dm[1], c->Theta, true,
                   c->stream);
    });
    timed(c, F_SWEEP_Z, [&] {
      launch_sweep_z_rk(conv, c->cfg.averaging, mode, Q, c->R, c->Qo, Qn, c->g, c->ph, c->seg, c->Theta, cA, cB,
                        first, c->stream);
    });
    return;
  }
  if (!is_central(conv)) {
    for (int d = 0; d < c->nd; ++d) {
      timed(c, F_SWEEP_X + d, [&] {
        launch_sweep(c->nd, d, conv, c->cfg.averaging, Q, c->R, c->g, c->ph, c->seg, c->bnd.dm[d], c->Theta, d > 0,
                     c->stream);
      });
    }
  }
  const bool visc = c->cfg.viscous != OSBL_VISC_NONE;
  if (c->general) {
    timed(c, F_UPDATE, [&] {
      launch_update_gen(c->nd, conv, visc, mode, Q, c->P, c->R, c->Qo, Qn, c->g, c->ph, c->bnd, cA, cB, first,
                        c->suth ? 1 : 0, c->stream);
    });
  } else if (c->tiled_visc) {
    timed(c, F_UPDATE, [&] {
      launch_visc_update3d(mode, Q, c->R, c->Qo, Qn, c->g, c->ph, cA, cB, first, c->vsz, c->stream);
    });
  } else {
    timed(c, F_UPDATE, [&] {
      launch_update(c->nd, conv, visc, mode, Q, c->P, c->R, c->Qo, Qn, c->g, c->ph, cA, cB, first, c->stream);
    });
  }
}

int check_config(const osbl_config* cfg) {
  if (cfg->ndim < 1 || cfg->ndim > 3) return fail(OSBL_E_ARG, "ndim must be 1, 2 or 3 (got %d)", cfg->ndim);
  for (int d = 0; d < 3; ++d) {
    const int n = cfg->n[d];
    if (d < cfg->ndim) {
      if (n < 8) return fail(OSBL_E_ARG, "block0np%d = %d: need at least 8 points", d, n);
      if (!(cfg->delta[d] > 0.0)) return fail(OSBL_E_ARG, "Delta%dblock0 must be positive", d);
    } else if (n != 1) {
      return fail(OSBL_E_ARG, "block0np%d must be 1 for a %d-D block", d, cfg->ndim);
    }
  }
  if (cfg->conv < OSBL_CONV_CENTRAL_SKEW || cfg->conv > OSBL_CONV_TENO6)
    return fail(OSBL_E_ARG, "unknown convective scheme %d", cfg->conv);
  if (cfg->averaging != OSBL_AVG_SIMPLE && cfg->averaging != OSBL_AVG_ROE)
    return fail(OSBL_E_ARG, "unknown averaging %d", cfg->averaging);
  if (cfg->viscous < OSBL_VISC_NONE || cfg->viscous > OSBL_VISC_SUTHERLAND)
    return fail(OSBL_E_ARG, "unknown viscous model %d", cfg->viscous);
  if (cfg->viscous == OSBL_VISC_SUTHERLAND && !(cfg->refT > 0.0 && cfg->suthT >= 0.0))
    return fail(OSBL_E_ARG, "Sutherland viscosity needs RefT > 0 and SuthT >= 0");
  if (cfg->rk < OSBL_RK3 || cfg->rk > OSBL_RK4_LS) return fail(OSBL_E_ARG, "unknown RK scheme %d", cfg->rk);
  for (int d = 0; d < cfg->ndim; ++d) {
    for (int s = 0; s < 2; ++s) {
      const int bc = cfg->bc[d][s];
      if (bc < OSBL_BC_PERIODIC || bc > OSBL_BC_SYMMETRY)
        return fail(OSBL_E_ARG, "unknown boundary condition %d on direction %d side %d", bc, d, s);
      if (bc == OSBL_BC_INLET_PEXTRAP && s != 0)
        return fail(OSBL_E_ARG, "direction %d: the pressure-extrapolating inlet is implemented for side 0 only", d);
      if (bc == OSBL_BC_ADIABATIC_WALL || bc == OSBL_BC_ISOTHERMAL_WALL || bc == OSBL_BC_INLET_PEXTRAP) {
        if (cfg->ndim < 2 && bc != OSBL_BC_INLET_PEXTRAP)
          return fail(OSBL_E_ARG, "wall boundaries need a 2-D or 3-D block");
      }
      if (cfg->bscheme[d][s] != OSBL_BS_CARPENTER && cfg->bscheme[d][s] != OSBL_BS_REDUCED_ACCESS)
        return fail(OSBL_E_ARG, "unknown boundary scheme %d on direction %d side %d", cfg->bscheme[d][s], d, s);
    }
    if ((cfg->bc[d][0] == OSBL_BC_PERIODIC) != (cfg->bc[d][1] == OSBL_BC_PERIODIC))
      return fail(OSBL_E_ARG, "direction %d: a periodic boundary needs both sides periodic", d);
    if (cfg->stretch[d] && !(modifies_stencil(cfg->bc[d][0]) && modifies_stencil(cfg->bc[d][1])))
      return fail(OSBL_E_ARG, "direction %d: a stretched direction needs boundary conditions that close the "
                              "central stencils on both faces", d);
  }
  if (cfg->gama <= 1.0) return fail(OSBL_E_ARG, "gama must exceed 1");
  {
    // the body force lives in the general update kernels (blocks with
    // closures, a metric or Sutherland viscosity: the channel flows)
    bool force = false, general = cfg->viscous == OSBL_VISC_SUTHERLAND;
    for (int d = 0; d < cfg->ndim; ++d) {
      force |= cfg->body_force[d] != 0.0;
      general |= cfg->stretch[d] != 0 || modifies_stencil(cfg->bc[d][0]) || modifies_stencil(cfg->bc[d][1]);
    }
    if (force && (!general || cfg->curvilinear))
      return fail(OSBL_E_ARG, "the body force is implemented on blocks with closures, a metric or Sutherland "
                              "viscosity");
  }
  if (cfg->curvilinear) {
    if (cfg->ndim < 2) return fail(OSBL_E_ARG, "a curvilinear metric needs a 2-D or 3-D block");
    if (is_central(cfg->conv))
      return fail(OSBL_E_ARG, "curvilinear blocks are implemented for the WENO/TENO schemes");
    if (cfg->viscous != OSBL_VISC_NONE || cfg->filters || cfg->sensor)
      return fail(OSBL_E_ARG, "curvilinear blocks are implemented for the Euler equations without filters or "
                              "sensor");
    for (int d = 0; d < cfg->ndim; ++d)
      if (cfg->bc[d][0] != OSBL_BC_PERIODIC || cfg->stretch[d])
        return fail(OSBL_E_ARG, "curvilinear blocks are implemented with periodic boundaries in every direction "
                                "(direction %d is not)", d);
  }
  if (cfg->sensor) {
    if (cfg->conv != OSBL_CONV_TENO5 && cfg->conv != OSBL_CONV_TENO6)
      return fail(OSBL_E_ARG, "the adaptive cut-off (Ducros sensor) applies to the TENO schemes");
    if (cfg->ndim < 2) return fail(OSBL_E_ARG, "the Ducros sensor needs a 2-D or 3-D block");
  }
  if (cfg->viscous != OSBL_VISC_NONE && (cfg->Re <= 0.0 || cfg->Pr <= 0.0 || cfg->Minf <= 0.0))
    return fail(OSBL_E_ARG, "viscous terms need positive Re, Pr and Minf");
  if (cfg->filters & ~(OSBL_FILTER_WENO | OSBL_FILTER_SFD | OSBL_FILTER_BINOMIAL))
    return fail(OSBL_E_ARG, "unknown filter bits 0x%x", cfg->filters);
  if (cfg->filters && cfg->ndim < 2)
    return fail(OSBL_E_ARG, "the filters are defined for 2-D and 3-D blocks (their conservative variable lists)");
  if (cfg->filters & OSBL_FILTER_WENO) {
    if (cfg->wf_sensor != OSBL_WF_DUCROS && cfg->wf_sensor != OSBL_WF_CONSTANT)
      return fail(OSBL_E_ARG, "unknown WENO filter dissipation sensor %d", cfg->wf_sensor);
  }
  if (cfg->filters & OSBL_FILTER_SFD) {
    if (!(cfg->omega_filt != 0.0) || !(cfg->chi_filt / cfg->omega_filt != -1.0))
      return fail(OSBL_E_ARG, "SFD needs omega_filt != 0 and chi_filt/omega_filt != -1");
  }
  if (cfg->filters & OSBL_FILTER_BINOMIAL) {
    const int N = cfg->binomial_order;
    if (N < 2 || N % 2) return fail(OSBL_E_ARG, "the binomial filter is defined for even orders (got %d)", N);
    if (N > 8)
      return fail(OSBL_E_ARG, "binomial order %d reads %d halo points; this build keeps at most 4", N, N / 2);
    const bool central = is_central(cfg->conv) && !(cfg->filters & OSBL_FILTER_WENO);
    for (int d = 0; d < cfg->ndim; ++d)
      for (int s = 0; s < 2; ++s) {
        const int bc = cfg->bc[d][s];
        if (bc == OSBL_BC_PERIODIC) continue;  // the exchange fills every halo point
        // the reference's boundary kernels fill the block's boundary halos only
        const int fill = central ? 2 : (s == 0 ? 3 : 4);
        if (N / 2 > fill)
          return fail(OSBL_E_ARG, "binomial order %d reads %d halo points; face (%d, %d) fills %d", N, N / 2, d, s,
                      fill);
      }
  }
  return OSBL_OK;
}

}  // namespace

extern "C" {

int osbl_abi_version(void) { return OSBL_ABI_VERSION; }

const char* osbl_last_error(void) { return g_err.c_str(); }

int osbl_create(const osbl_config* cfg, osbl_ctx** out) {
  if (!cfg || !out) return fail(OSBL_E_ARG, "null argument");
  *out = nullptr;
  int rc = check_config(cfg);
  if (rc) return rc;
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (ndev < 1) return fail(OSBL_E_HIP, "no HIP device");
  auto* c = new osbl_ctx();
  c->cfg = *cfg;
  c->nd = cfg->ndim;
  c->nv = cfg->ndim + 2;
  HIPCHK(hipGetDevice(&c->device));
  // 2 halo points serve the central stencils, 4 the characteristic
  // reconstructions (3 below, 4 above) and binomial filters above order 4
  const bool narrow = is_central(cfg->conv) && cfg->viscous == OSBL_VISC_NONE && !(cfg->filters & OSBL_FILTER_WENO) &&
                      !((cfg->filters & OSBL_FILTER_BINOMIAL) && cfg->binomial_order > 4);
  const int halo = narrow ? 2 : 4;
  Grid& g = c->g;
  g.nd = c->nd;
  for (int d = 0; d < 3; ++d) {
    g.n[d] = d < c->nd ? cfg->n[d] : 1;
    g.h[d] = d < c->nd ? halo : 0;
  }
  g.pitch = ((OX + g.n[0] + g.h[0]) + 15) / 16 * 16;
  g.periodic_x = cfg->bc[0][0] == OSBL_BC_PERIODIC;
  g.rows = g.n[1] + 2 * g.h[1];
  g.planes = g.n[2] + 2 * g.h[2];
  g.plane = g.pitch * g.rows;
  long long fs = (long long)g.plane * g.planes;
  fs = (fs + 63) / 64 * 64;
  g.fs = fs;
  if (fs * (long long)c->nv >= (1LL << 31)) {
    delete c;
    return fail(OSBL_E_ARG, "block too large for 32-bit in-field offsets");
  }
  Phys& ph = c->ph;
  ph.gama = cfg->gama;
  ph.gm1 = cfg->gama - 1.0;
  ph.igm1 = 1.0 / ph.gm1;
  ph.Minf = cfg->Minf;
  ph.Re = cfg->Re;
  ph.Pr = cfg->Pr;
  ph.dt = cfg->dt;
  ph.eps = cfg->eps;
  ph.teno_ct = cfg->teno_ct;
  for (int d = 0; d < 3; ++d) ph.inv_delta[d] = d < c->nd ? 1.0 / cfg->delta[d] : 0.0;
  ph.inv_re = cfg->viscous != OSBL_VISC_NONE ? 1.0 / cfg->Re : 0.0;
  ph.cq = cfg->viscous != OSBL_VISC_NONE ? 1.0 / ((cfg->gama - 1.0) * cfg->Minf * cfg->Minf * cfg->Pr * cfg->Re) : 0.0;
  ph.suth = cfg->viscous == OSBL_VISC_SUTHERLAND ? cfg->suthT / cfg->refT : 0.0;
  ph.teno_a1 = cfg->teno_a1;
  ph.teno_a2 = cfg->teno_a2;
  ph.sensor_eps = cfg->sensor_eps;
  for (int d = 0; d < 3; ++d) {
    ph.force[d] = d < c->nd ? cfg->body_force[d] : 0.0;
    ph.has_force |= ph.force[d] != 0.0;
  }
  // the non-temporal hint on R and the RK registers: the central kernels gain
  // from it once Q, the two registers and R exceed about half of the 256 MB
  // last-level cache (at 64^3, 76 MB, plain accesses are 4 % faster; at 128^3
  // and 256^3 the hint is), the characteristic sweeps, whose Q re-reads from
  // L2 matter more than keeping R cached, at every measured size (2-D 512^2,
  // 36 MB: 3 % faster forward steps)
  ph.nt = (!is_central(cfg->conv) || 4.0 * c->nv * (double)fs * sizeof(double) > 128.0 * (1 << 20)) ? 1 : 0;
  c->nstages = rk_coefficients(cfg->rk, c->A, c->B);
  c->suth = cfg->viscous == OSBL_VISC_SUTHERLAND;
  bool closures = false, stretched = false;
  for (int d = 0; d < c->nd; ++d) {
    stretched |= cfg->stretch[d] != 0;
    for (int s = 0; s < 2; ++s)
      if (modifies_stencil(cfg->bc[d][s])) {
        c->bnd.mod[d][s] = 1 + cfg->bscheme[d][s];
        closures = true;
      }
  }
  c->general = closures || stretched || c->suth;
  c->tiled_visc = !c->general && c->nd == 3 && !is_central(cfg->conv) && cfg->viscous != OSBL_VISC_NONE;
  c->tiled_gen = c->general && c->nd == 3 && !is_central(cfg->conv) && cfg->viscous != OSBL_VISC_NONE &&
                 !c->bnd.mod[2][0] && !c->bnd.mod[2][1] && !cfg->stretch[2];
  // the tiled central kernels march columns along z in segments short
  // enough to give the GPU ~1024 blocks; a grid with too few tiles even then
  // (below ~32^3) keeps the per-point kernels
  c->vsz = visc3d_segment(c->g);
  c->csz = conv3d_segment(c->g);
  c->tiled_central = !c->general && c->nd == 3 && is_central(cfg->conv) && !cfg->curvilinear &&
                     conv3d_blocks(c->g, c->csz) >= 128;
  c->need_prims = !c->tiled_central && (is_central(cfg->conv) || (cfg->viscous != OSBL_VISC_NONE && !c->tiled_visc));
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  const size_t bundle = (size_t)fs * c->nv * sizeof(double);
  auto alloc = [&](double** p, size_t bytes) -> int {
    if (hipMalloc((void**)p, bytes) != hipSuccess) return fail(OSBL_E_NOMEM, "hipMalloc of %zu bytes failed", bytes);
    if (hipMemsetAsync(*p, 0, bytes, c->stream) != hipSuccess) return fail(OSBL_E_HIP, "hipMemsetAsync failed");
    return OSBL_OK;
  };
  if ((rc = alloc(&c->Qbuf[0], bundle)) || (rc = alloc(&c->Qbuf[1], bundle)) || (rc = alloc(&c->Qo, bundle)) ||
      (rc = alloc(&c->R, bundle)) ||
      (rc = alloc(&c->stage_buf, (size_t)g.n[0] * g.n[1] * g.n[2] * sizeof(double) * c->nv))) {
    osbl_destroy(c);
    return rc;
  }
  if (c->need_prims && (rc = alloc(&c->P, (size_t)fs * (c->nd + 3) * sizeof(double)))) {
    osbl_destroy(c);
    return rc;
  }
  c->curv = cfg->curvilinear != 0;
  if (c->curv && ((rc = alloc(&c->cmet, (size_t)fs * (c->nd * c->nd + 1) * sizeof(double))) ||
                  (rc = alloc(&c->cflux, bundle)) ||
                  (rc = alloc(&c->cpt, (size_t)fs * (c->nv + 2) * sizeof(double))))) {
    osbl_destroy(c);
    return rc;
  }
  const bool wf_ducros = (cfg->filters & OSBL_FILTER_WENO) && cfg->wf_sensor == OSBL_WF_DUCROS;
  if ((cfg->sensor || wf_ducros) && (rc = alloc(&c->Theta, (size_t)fs * sizeof(double)))) {
    osbl_destroy(c);
    return rc;
  }
  if ((cfg->filters & OSBL_FILTER_SFD) && (rc = alloc(&c->QF, bundle))) {
    osbl_destroy(c);
    return rc;
  }
  for (int d = 0; d < c->nd; ++d)
    for (int s = 0; s < 2; ++s)
      if (cfg->bc[d][s] == OSBL_BC_ADIABATIC_WALL || cfg->bc[d][s] == OSBL_BC_ISOTHERMAL_WALL)
        c->walls |= 1 << (2 * d + s);
  if (cfg->filters & OSBL_FILTER_SFD) {
    c->sfd_c = cfg->chi_filt / cfg->omega_filt;
    c->sfd_E = std::exp(-(cfg->chi_filt + cfg->omega_filt) * cfg->dt);
  }
  if (cfg->filters & OSBL_FILTER_BINOMIAL) {
    // binomial(N, t)/2^N (binomial_filter.py:36-42)
    const int N = cfg->binomial_order;
    double b = 1.0;
    for (int t = 0; t <= N; ++t) {
      c->binom_w[t] = b / std::ldexp(1.0, N);
      b = b * (N - t) / (t + 1);
    }
  }
  {
    double cw[CW_SIZE];
    closure_table(cw);
    if ((rc = alloc(&c->cw_dev, sizeof(cw)))) {
      osbl_destroy(c);
      return rc;
    }
    if (hipMemcpyAsync(c->cw_dev, cw, sizeof(cw), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
      osbl_destroy(c);
      return fail(OSBL_E_HIP, "closure table upload failed");
    }
    c->bnd.cw = c->cw_dev;
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    osbl_destroy(c);
    return fail(OSBL_E_HIP, "stream synchronize failed");
  }
  *out = c;
  return OSBL_OK;
}

// forget the captured iterations (their launch arguments no longer hold);
// the next osbl_run captures them again
static void drop_graphs(osbl_ctx* c) {
  for (int k = 0; k < 2; ++k)
    if (c->graph2[k]) {
      (void)hipGraphExecDestroy(c->graph2[k]);
      c->graph2[k] = nullptr;
    }
}

void osbl_destroy(osbl_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int k = 0; k < 2; ++k)
    if (c->graph2[k]) (void)hipGraphExecDestroy(c->graph2[k]);
  for (auto& r : c->recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (auto e : c->pool) (void)hipEventDestroy(e);
  (void)hipFree(c->Qbuf[0]);
  (void)hipFree(c->Qbuf[1]);
  (void)hipFree(c->Qo);
  (void)hipFree(c->R);
  (void)hipFree(c->P);
  (void)hipFree(c->stage_buf);
  (void)hipFree(c->Abuf);
  (void)hipFree(c->Mbuf);
  (void)hipFree(c->pk);
  (void)hipFree(c->vw);
  (void)hipFree(c->vy);
  (void)hipFree(c->cmet);
  (void)hipFree(c->cflux);
  (void)hipFree(c->cpt);
  (void)hipFree(c->cw_dev);
  (void)hipFree(c->Theta);
  (void)hipFree(c->QF);
  (void)hipFree(c->fmask);
  for (int d = 0; d < 3; ++d) {
    (void)hipFree(c->met_dev[d]);
    for (int s = 0; s < 2; ++s) (void)hipFree(c->plane_dev[d][s]);
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

static long long cells(const osbl_ctx* c) { return (long long)c->g.n[0] * c->g.n[1] * c->g.n[2]; }

int osbl_set_field_dev(osbl_ctx* c, int var, const double* dev) {
  if (!c || !dev) return fail(OSBL_E_ARG, "null argument");
  if (var < 0 || var >= c->nv) return fail(OSBL_E_ARG, "variable index %d out of range [0,%d)", var, c->nv);
  launch_unpack(dev, Qcur(c) + var * c->g.fs, c->g, c->stream);
  HIPCHK(hipGetLastError());
  return OSBL_OK;
}

int osbl_get_field_dev(osbl_ctx* c, int var, double* dev) {
  if (!c || !dev) return fail(OSBL_E_ARG, "null argument");
  if (var < 0 || var >= c->nv) return fail(OSBL_E_ARG, "variable index %d out of range [0,%d)", var, c->nv);
  launch_pack(Qcur(c) + var * c->g.fs, dev, c->g, c->stream);
  HIPCHK(hipGetLastError());
  return OSBL_OK;
}

int osbl_set_field(osbl_ctx* c, int var, const double* host) {
  if (!c || !host) return fail(OSBL_E_ARG, "null argument");
  if (var < 0 || var >= c->nv) return fail(OSBL_E_ARG, "variable index %d out of range [0,%d)", var, c->nv);
  HIPCHK(hipMemcpyAsync(c->stage_buf, host, cells(c) * sizeof(double), hipMemcpyHostToDevice, c->stream));
  int rc = osbl_set_field_dev(c, var, c->stage_buf);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return OSBL_OK;
}

int osbl_get_field(osbl_ctx* c, int var, double* host) {
  if (!c || !host) return fail(OSBL_E_ARG, "null argument");
  int rc = osbl_get_field_dev(c, var, c->stage_buf);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(host, c->stage_buf, cells(c) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return OSBL_OK;
}

static int check_ready(const osbl_ctx* c);

// halo layers the boundary conditions keep filled on each face: every layer
// of a periodic direction, the boundary halos of the schemes on the others
static void filled_halos(const osbl_ctx* c, int32_t* d_m, int32_t* d_p) {
  for (int d = 0; d < 3; ++d) {
    d_m[d] = d_p[d] = 0;
    if (d >= c->nd) continue;
    if (c->cfg.bc[d][0] == OSBL_BC_PERIODIC) {
      d_m[d] = -c->g.h[d];
      d_p[d] = c->g.h[d];
    } else {
      d_m[d] = -std::min(bc_halos(c, 0), c->g.h[d]);
      d_p[d] = std::min(bc_halos(c, 1), c->g.h[d]);
    }
  }
}

int osbl_get_field_halo(osbl_ctx* c, int var, int32_t* d_m, int32_t* d_p, double* host) {
  if (!c || !d_m || !d_p) return fail(OSBL_E_ARG, "null argument");
  if (var < 0 || var >= c->nv) return fail(OSBL_E_ARG, "variable index %d out of range [0,%d)", var, c->nv);
  filled_halos(c, d_m, d_p);
  if (!host) return OSBL_OK;
  if (int rc = check_ready(c)) return rc;
  apply_bcs(c, Qcur(c));
  const int m0 = -d_m[0], m1 = -d_m[1], m2 = -d_m[2];
  const size_t W = size_t(c->g.n[0] + m0 + d_p[0]);
  const size_t H = size_t(c->g.n[1] + m1 + d_p[1]);
  const int K = c->g.n[2] + m2 + d_p[2];
  const double* src = Qcur(c) + var * c->g.fs;
  for (int k = 0; k < K; ++k) {
    HIPCHK(hipMemcpy2DAsync(host + size_t(k) * H * W, W * sizeof(double), src + c->g.idx(-m0, -m1, k - m2),
                            size_t(c->g.pitch) * sizeof(double), W * sizeof(double), H, hipMemcpyDeviceToHost,
                            c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return OSBL_OK;
}

// metric and face planes the configuration asks for must be set before stepping
static int check_ready(const osbl_ctx* c) {
  if (c->curv && !c->cmet_ready)
    return fail(OSBL_E_STATE, "curvilinear block: call osbl_set_grid_coordinates first");
  for (int d = 0; d < c->nd; ++d) {
    if (c->cfg.stretch[d] && !c->bnd.dm[d])
      return fail(OSBL_E_STATE, "direction %d is stretched: call osbl_set_coordinates first", d);
    for (int s = 0; s < 2; ++s) {
      const int bc = c->cfg.bc[d][s];
      if ((bc == OSBL_BC_DIRICHLET || bc == OSBL_BC_INLET_PEXTRAP) && !c->plane_dev[d][s])
        return fail(OSBL_E_STATE, "face (%d, %d) needs its plane: call osbl_set_bc_plane first", d, s);
    }
  }
  return OSBL_OK;
}

// the filters after the full time step (algorithm.py:412-470), in the order
// WENO filter, SFD, binomial filter
static void filters(osbl_ctx* c) {
  const int f = c->cfg.filters;
  if (f & OSBL_FILTER_WENO) {
    double* Q = Qcur(c);
    const bool ducros = c->cfg.wf_sensor == OSBL_WF_DUCROS;
    if (ducros) timed(c, F_FILTER, [&] { launch_sensor(c->nd, Q, c->Theta, c->g, c->ph, c->bnd, c->stream); });
    for (int d = 0; d < c->nd; ++d) {
      timed(c, F_FILTER, [&] {
        launch_sweep(c->nd, d, CONV_WFILTER, OSBL_AVG_SIMPLE, Q, c->R, c->g, c->ph, c->seg, c->bnd.dm[d], nullptr,
                     d > 0, c->stream);
      });
    }
    timed(c, F_FILTER, [&] {
      launch_weno_filter_apply(c->nd, Q, c->R, c->Theta, c->g, c->ph, ducros, c->cfg.wf_mach != 0,
                               c->cfg.shock_filter_control, c->walls, c->stream);
    });
  }
  if (f & OSBL_FILTER_SFD)
    timed(c, F_FILTER, [&] { launch_sfd(c->nd, Qcur(c), c->QF, c->g, c->sfd_c, c->sfd_E, c->stream); });
  if (f & OSBL_FILTER_BINOMIAL) {
    timed(c, F_FILTER, [&] {
      launch_binomial(c->nd, Qcur(c), c->Qbuf[1 - c->cur], c->g, c->binom_w, c->cfg.binomial_order, c->cfg.sigma_filt,
                      c->fmask, c->stream);
    });
    c->cur ^= 1;
  }
}

static void iterate(osbl_ctx* c, int mode) {
  // with filters the interior changes after the last stage's boundary
  // conditions, so the iteration starts with them as the reference's does
  if (c->cfg.filters) apply_bcs(c, Qcur(c));
  for (int s = 0; s < c->nstages; ++s) {
    stage(c, mode, c->A[s], c->B[s], s == 0);
    c->cur ^= 1;
    apply_bcs(c, Qcur(c));
  }
  if (c->cfg.filters) filters(c);
}

// capture two iterations starting from buffer `start` into an executable graph
static int capture_pair(osbl_ctx* c, int mode, int start) {
  const int saved = c->cur;
  c->cur = start;
  hipGraph_t graph;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  iterate(c, mode);
  iterate(c, mode);
  HIPCHK(hipStreamEndCapture(c->stream, &graph));
  c->cur = saved;
  hipError_t e = hipGraphInstantiate(&c->graph2[start], graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    c->graph2[start] = nullptr;
    return fail(OSBL_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  return OSBL_OK;
}

int osbl_run(osbl_ctx* c, int64_t niter) {
  if (!c) return fail(OSBL_E_ARG, "null context");
  if (niter < 0) return fail(OSBL_E_ARG, "niter must be non-negative");
  if (int rc = check_ready(c)) return rc;
  const int mode = c->cfg.rk == OSBL_RK3 ? 0 : 1;
  if (niter > 0) apply_bcs(c, Qcur(c));
  if ((c->cfg.filters & OSBL_FILTER_SFD) && niter > 0) {
    // "Initialize the filter" before the time loop: vfilt = v (SFD.py:50-61)
    for (int m = 0; m < c->nv; ++m) {
      if (c->sfd_set[m]) continue;
      HIPCHK(hipMemcpyAsync(c->QF + m * c->g.fs, Qcur(c) + m * c->g.fs, c->g.fs * sizeof(double),
                            hipMemcpyDeviceToDevice, c->stream));
      c->sfd_set[m] = true;
    }
  }
  int64_t it = 0;
  // Launch-bound blocks replay two iterations from a graph (kernels only); with
  // profiling on, every launch is issued directly so it can be timed.
  if (c->use_graphs && !c->prof && niter >= 2) {
    const int start = c->cur;
    // buffers swap once per stage and once per binomial filter pass
    const int parity_after = (2 * (c->nstages + ((c->cfg.filters & OSBL_FILTER_BINOMIAL) ? 1 : 0))) & 1;
    if (!c->graph2[start]) {
      int rc = capture_pair(c, mode, start);
      if (rc) return rc;
    }
    for (; it + 2 <= niter; it += 2) {
      HIPCHK(hipGraphLaunch(c->graph2[c->cur], c->stream));
      c->cur ^= parity_after;
      if (parity_after && !c->graph2[c->cur]) {
        int rc = capture_pair(c, mode, c->cur);
        if (rc) return rc;
      }
    }
  }
  for (; it < niter; ++it) iterate(c, mode);
  HIPCHK(hipGetLastError());
  return OSBL_OK;
}

int osbl_set_coordinates(osbl_ctx* c, int d, const double* x) {
  if (!c || !x) return fail(OSBL_E_ARG, "null argument");
  if (d < 0 || d >= c->nd) return fail(OSBL_E_ARG, "direction %d out of range", d);
  if (!c->cfg.stretch[d]) return fail(OSBL_E_STATE, "direction %d is not declared stretched", d);
  const int N = c->g.n[d];
  double cw[CW_SIZE];
  closure_table(cw);
  // metric kernels of MetricsEquation (metric.py:241-273): detJ = CD(x_d, xi_d),
  // D_dd = 1/detJ, then SD_ddd = CD(D_dd, xi_d), each with the faces' closures
  std::vector<double> buf(2 * (size_t)N);
  const double inv = 1.0 / c->cfg.delta[d];
  const int lo = c->bnd.mod[d][0], hi = c->bnd.mod[d][1];
  for (int i = 0; i < N; ++i) buf[i] = 1.0 / (inv * d1_line(x, N, lo, hi, i, cw));
  for (int i = 0; i < N; ++i) buf[N + i] = inv * d1_line(buf.data(), N, lo, hi, i, cw);
  for (int i = 0; i < N; ++i)
    if (!std::isfinite(buf[i]) || !std::isfinite(buf[N + i]))
      return fail(OSBL_E_ARG, "direction %d: the coordinates do not give a finite metric at %d", d, i);
  if (!c->met_dev[d] && hipMalloc((void**)&c->met_dev[d], 2 * (size_t)N * sizeof(double)) != hipSuccess)
    return fail(OSBL_E_NOMEM, "metric allocation failed");
  HIPCHK(hipMemcpyAsync(c->met_dev[d], buf.data(), 2 * (size_t)N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->bnd.dm[d] = c->met_dev[d];
  c->bnd.sd[d] = c->met_dev[d] + N;
  drop_graphs(c);  // launch arguments changed
  return OSBL_OK;
}

int osbl_set_grid_coordinates(osbl_ctx* c, int halo, const double* coords) {
  if (!c || !coords) return fail(OSBL_E_ARG, "null argument");
  if (!c->curv) return fail(OSBL_E_STATE, "the block is not curvilinear");
  if (halo < 2 || halo > 64) return fail(OSBL_E_ARG, "coordinate halo %d: need 2..64 points", halo);
  const int nd = c->nd;
  // the central stencil reads 2 points beyond the interior: keep those
  int e[3], in[3];
  for (int d = 0; d < 3; ++d) {
    e[d] = d < nd ? c->g.n[d] + 4 : 1;
    in[d] = d < nd ? c->g.n[d] + 2 * halo : 1;
  }
  const int off = halo - 2;
  const long long per = (long long)e[0] * e[1] * e[2], per_in = (long long)in[0] * in[1] * in[2];
  std::vector<double> sub((size_t)(per * nd));
  for (int a = 0; a < nd; ++a)
    for (int k = 0; k < e[2]; ++k)
      for (int j = 0; j < e[1]; ++j)
        for (int i = 0; i < e[0]; ++i) {
          const long long src = (i + off) + (long long)in[0] * ((j + (nd > 1 ? off : 0)) +
                                                                 (long long)in[1] * (k + (nd > 2 ? off : 0)));
          sub[(size_t)(a * per + i + (long long)e[0] * (j + (long long)e[1] * k))] = coords[a * per_in + src];
        }
  double* X = nullptr;
  if (hipMalloc((void**)&X, sub.size() * sizeof(double)) != hipSuccess)
    return fail(OSBL_E_NOMEM, "hipMalloc of the coordinates failed");
  int rc = OSBL_OK;
  if (hipMemcpyAsync(X, sub.data(), sub.size() * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess)
    rc = fail(OSBL_E_HIP, "coordinate upload failed");
  if (!rc) launch_curv_metric(nd, X, c->cmet, c->g, c->ph, e, c->stream);
  if (!rc && (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess))
    rc = fail(OSBL_E_HIP, "metric kernels failed");
  (void)hipFree(X);
  if (!rc) c->cmet_ready = true;
  return rc;
}

int osbl_get_metric_field(osbl_ctx* c, int idx, double* host) {
  if (!c || !host) return fail(OSBL_E_ARG, "null argument");
  if (idx < 0 || idx > c->nd * c->nd) return fail(OSBL_E_ARG, "metric index %d out of range", idx);
  if (!c->curv) {
    // a direction stretched along itself: D_dd is a function of its own index
    const int d = idx / c->nd;
    if (idx == c->nd * c->nd || idx % c->nd != d || !c->cfg.stretch[d])
      return fail(OSBL_E_ARG, "metric index %d: not an array of this block (D_dd of its stretched directions)", idx);
    if (!c->bnd.dm[d]) return fail(OSBL_E_STATE, "direction %d is stretched: call osbl_set_coordinates first", d);
    std::vector<double> line((size_t)c->g.n[d]);
    HIPCHK(hipMemcpyAsync(line.data(), c->met_dev[d], line.size() * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    long long o = 0;
    for (int k = 0; k < c->g.n[2]; ++k)
      for (int j = 0; j < c->g.n[1]; ++j)
        for (int i = 0; i < c->g.n[0]; ++i) host[o++] = line[d == 0 ? i : (d == 1 ? j : k)];
    return OSBL_OK;
  }
  if (!c->cmet_ready) return fail(OSBL_E_STATE, "curvilinear block: call osbl_set_grid_coordinates first");
  launch_pack(c->cmet + idx * c->g.fs, c->stage_buf, c->g, c->stream);
  HIPCHK(hipMemcpyAsync(host, c->stage_buf, cells(c) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return OSBL_OK;
}

int osbl_get_sensor(osbl_ctx* c, double* host) {
  if (!c || !host) return fail(OSBL_E_ARG, "null argument");
  if (!c->Theta) return fail(OSBL_E_STATE, "the block has no shock sensor");
  launch_pack(c->Theta, c->stage_buf, c->g, c->stream);
  HIPCHK(hipMemcpyAsync(host, c->stage_buf, cells(c) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return OSBL_OK;
}

int osbl_nan_count(osbl_ctx* c, int var, int64_t* count) {
  if (!c || !count) return fail(OSBL_E_ARG, "null argument");
  if (var >= c->nv) return fail(OSBL_E_ARG, "variable %d out of range (the block has %d)", var, c->nv);
  auto* ctr = reinterpret_cast<unsigned long long*>(c->stage_buf);
  launch_count_nan(Qcur(c), var < 0 ? 0 : var, var < 0 ? c->nv : var + 1, ctr, c->g, c->stream);
  unsigned long long n = 0;
  HIPCHK(hipMemcpyAsync(&n, ctr, sizeof(n), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *count = (int64_t)n;
  return OSBL_OK;
}

int osbl_get_point(osbl_ctx* c, int var, const int32_t* index, double* value) {
  if (!c || !index || !value) return fail(OSBL_E_ARG, "null argument");
  if (var < 0 || var >= c->nv) return fail(OSBL_E_ARG, "variable %d out of range (the block has %d)", var, c->nv);
  int ijk[3] = {0, 0, 0};
  for (int d = 0; d < c->g.nd; ++d) {
    if (index[d] < 0 || index[d] >= c->g.n[d])
      return fail(OSBL_E_ARG, "index %d of direction %d outside the interior 0..%d", index[d], d, c->g.n[d] - 1);
    ijk[d] = index[d];
  }
  const double* src = Qcur(c) + var * c->g.fs + c->g.idx(ijk[0], ijk[1], ijk[2]);
  HIPCHK(hipMemcpyAsync(value, src, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return OSBL_OK;
}

int osbl_set_filter_mask(osbl_ctx* c, const uint8_t* mask) {
  if (!c) return fail(OSBL_E_ARG, "null context");
  if (!(c->cfg.filters & OSBL_FILTER_BINOMIAL)) return fail(OSBL_E_STATE, "the block has no binomial filter");
  HIPCHK(hipStreamSynchronize(c->stream));
  // the captured iterations hold the mask pointer (or its absence) as a kernel
  // argument: both branches below change it, so they are captured again
  drop_graphs(c);
  if (!mask) {
    HIPCHK(hipFree(c->fmask));
    c->fmask = nullptr;
    return OSBL_OK;
  }
  const size_t n = (size_t)cells(c);
  if (!c->fmask && hipMalloc((void**)&c->fmask, n) != hipSuccess) {
    c->fmask = nullptr;
    return fail(OSBL_E_NOMEM, "hipMalloc of %zu bytes failed", n);
  }
  HIPCHK(hipMemcpy(c->fmask, mask, n, hipMemcpyHostToDevice));
  return OSBL_OK;
}

int osbl_get_sfd_field(osbl_ctx* c, int var, double* host) {
  if (!c || !host) return fail(OSBL_E_ARG, "null argument");
  if (!(c->cfg.filters & OSBL_FILTER_SFD)) return fail(OSBL_E_STATE, "the block has no SFD filter");
  if (var < 0 || var >= c->nv) return fail(OSBL_E_ARG, "variable index %d out of range [0,%d)", var, c->nv);
  launch_pack(c->QF + var * c->g.fs, c->stage_buf, c->g, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(host, c->stage_buf, cells(c) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return OSBL_OK;
}

int osbl_set_sfd_field(osbl_ctx* c, int var, const double* host) {
  if (!c || !host) return fail(OSBL_E_ARG, "null argument");
  if (!(c->cfg.filters & OSBL_FILTER_SFD)) return fail(OSBL_E_STATE, "the block has no SFD filter");
  if (var < 0 || var >= c->nv) return fail(OSBL_E_ARG, "variable index %d out of range [0,%d)", var, c->nv);
  HIPCHK(hipMemcpyAsync(c->stage_buf, host, cells(c) * sizeof(double), hipMemcpyHostToDevice, c->stream));
  launch_unpack(c->stage_buf, c->QF + var * c->g.fs, c->g, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  c->sfd_set[var] = true;
  return OSBL_OK;
}

int osbl_set_bc_plane(osbl_ctx* c, int d, int side, const double* plane) {
  if (!c || !plane) return fail(OSBL_E_ARG, "null argument");
  if (d < 0 || d >= c->nd || side < 0 || side > 1) return fail(OSBL_E_ARG, "face (%d, %d) out of range", d, side);
  const int bc = c->cfg.bc[d][side];
  if (bc != OSBL_BC_DIRICHLET && bc != OSBL_BC_INLET_PEXTRAP)
    return fail(OSBL_E_STATE, "face (%d, %d) takes no plane of values", d, side);
  const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
  const size_t n = (size_t)c->nv * c->g.n[e1] * c->g.n[e2];
  if (!c->plane_dev[d][side] && hipMalloc((void**)&c->plane_dev[d][side], n * sizeof(double)) != hipSuccess)
    return fail(OSBL_E_NOMEM, "plane allocation failed");
  HIPCHK(hipMemcpyAsync(c->plane_dev[d][side], plane, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (bc == OSBL_BC_INLET_PEXTRAP)  // both state buffers carry the inflow halos
    for (int k = 0; k < 2; ++k) launch_fill_face_halo(c->Qbuf[k], c->nv, c->g, d, side, c->plane_dev[d][side], c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  drop_graphs(c);
  return OSBL_OK;
}

int osbl_use_graphs(osbl_ctx* c, int enable) {
  if (!c) return fail(OSBL_E_ARG, "null context");
  c->use_graphs = enable != 0;
  return OSBL_OK;
}

int osbl_rhs(osbl_ctx* c, double* residual_host) {
  if (!c || !residual_host) return fail(OSBL_E_ARG, "null argument");
  if (int rc = check_ready(c)) return rc;
  apply_bcs(c, Qcur(c));
  stage(c, 2, 0.0, 0.0, 0);
  HIPCHK(hipGetLastError());
  const long long n = cells(c);
  for (int m = 0; m < c->nv; ++m) launch_pack(c->R + m * c->g.fs, c->stage_buf + m * n, c->g, c->stream);
  HIPCHK(hipMemcpyAsync(residual_host, c->stage_buf, n * c->nv * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return OSBL_OK;
}

int osbl_sync(osbl_ctx* c) {
  if (!c) return fail(OSBL_E_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  return OSBL_OK;
}

void* osbl_stream(osbl_ctx* c) { return c ? (void*)c->stream : nullptr; }

int osbl_profile(osbl_ctx* c, int enable) {
  if (!c) return fail(OSBL_E_ARG, "null context");
  drain_records(c);
  if (enable && !c->prof) {
    for (int f = 0; f < F_COUNT; ++f) { c->fam_ms[f] = 0.0; c->fam_n[f] = 0; }
  }
  c->prof = enable != 0;
  return OSBL_OK;
}

int osbl_kernel_stats(osbl_ctx* c, int index, char* name, int name_len, double* total_ms, int64_t* launches) {
  if (!c) return fail(OSBL_E_ARG, "null context");
  drain_records(c);
  if (index < 0 || index >= F_COUNT) return F_COUNT;
  if (name && name_len > 0) {
    std::strncpy(name, kFamilyNames[index], name_len - 1);
    name[name_len - 1] = '\0';
  }
  if (total_ms) *total_ms = c->fam_ms[index];
  if (launches) *launches = c->fam_n[index];
  return F_COUNT;
}

// ---------------------------------------------------------------------------
// Algorithmic work model.  FP64 operations are counted from the reference's
// formulas (FMA = 2, divide, sqrt, abs and max = 1 each), per interior cell:
//  * one characteristic interface (shock_capturing.py:357-536): Roe average,
//    eigenvectors, 6-point projection of F and Q, wave speeds, LLF split and
//    R.recon -- 850 (3-D), 620 (2-D), 300 (1-D) -- plus two reconstructions
//    per characteristic field: WENO5-Z 76, WENO5-JS 70, TENO5 88, TENO6 100;
//    plus 2 per variable for the flux difference;
//  * central skew convective terms per direction: 13 five-point derivatives
//    and the residual assembly, 140 (3-D), 110 (2-D), 80 (1-D);
//  * constant-viscosity terms: 530 (3-D), 200 (2-D), 40 (1-D);
//  * constituent relations 6 per variable; RK update 6 per variable.
// Bytes: 8*(4*nvar) per cell and stage (read Q and the RK register, write
// both); per kernel family, the datasets the kernel must read and write once.
// ---------------------------------------------------------------------------
static double recon_side_flops(int conv) {
  switch (conv) {
    case OSBL_CONV_WENO_Z: return 76.0;
    case OSBL_CONV_WENO_JS: return 70.0;
    case OSBL_CONV_TENO5: return 88.0;
    default: return 100.0;
  }
}
static double interface_flops(const osbl_ctx* c) {
  const double base = c->nd == 3 ? 850.0 : (c->nd == 2 ? 620.0 : 300.0);
  return base + 2.0 * c->nv * recon_side_flops(c->cfg.conv) + 2.0 * c->nv;
}
static double viscous_flops(const osbl_ctx* c) {
  if (c->cfg.viscous == OSBL_VISC_NONE) return 0.0;
  return c->nd == 3 ? 530.0 : (c->nd == 2 ? 200.0 : 40.0);
}
static double central_flops(const osbl_ctx* c) { return c->nd == 3 ? 140.0 : (c->nd == 2 ? 110.0 : 80.0); }

int osbl_work_per_stage(osbl_ctx* c, double* bytes, double* flops) {
  if (!c) return fail(OSBL_E_ARG, "null context");
  const double n = (double)cells(c);
  if (bytes) *bytes = 8.0 * 4.0 * c->nv * n;
  if (flops) {
    double f = 6.0 * c->nv + 6.0 * c->nv + viscous_flops(c);
    f += is_central(c->cfg.conv) ? c->nd * central_flops(c) : c->nd * interface_flops(c);
    *flops = f * n;
  }
  return OSBL_OK;
}

int osbl_kernel_work(osbl_ctx* c, int family, double* bytes, double* flops) {
  if (!c) return fail(OSBL_E_ARG, "null context");
  if (family < 0 || family >= F_COUNT) return fail(OSBL_E_ARG, "kernel family %d out of range", family);
  const double n = (double)cells(c);
  const double nv = c->nv;
  double b = 0.0, f = 0.0;
  switch (family) {
    case F_BC: b = 16.0 * nv * (double)halo_points(c->g); break;
    case F_PRIMS: b = 8.0 * (nv + c->nd + 2) * n; f = 6.0 * nv * n; break;
    case F_SWEEP_X: b = 16.0 * nv * n; f = interface_flops(c) * n; break;
    case F_SWEEP_Y: b = 24.0 * nv * n; f = interface_flops(c) * n; break;
    case F_SWEEP_Z:
      // with the fused RK update: read Q, R and the register, write Q and the register
      b = (c->tiled_visc || c->tiled_gen ? 40.0 : 24.0) * nv * n;
      f = (interface_flops(c) + (c->tiled_visc || c->tiled_gen ? 6.0 * nv : 0.0)) * n;
      break;
    case F_FILTER: b = 16.0 * nv * n; break;
    case F_CENTRAL:
      // read Q; write R (viscous terms follow) or read the register and write Q and the register
      b = (c->cfg.viscous != OSBL_VISC_NONE ? 16.0 : 32.0) * nv * n;
      f = (c->nd * central_flops(c) + (c->cfg.viscous != OSBL_VISC_NONE ? 0.0 : 6.0 * nv)) * n;
      break;
    case F_UPDATE: {
      if (c->tiled_central) {  // viscous terms + RK update: read Q, R, the register; write Q and the register
        b = 40.0 * nv * n;
        f = (viscous_flops(c) + 6.0 * nv) * n;
        break;
      }
      if (c->tiled_visc || c->tiled_gen) {  // viscous terms only: read Q (or the CRs), RMW R
        b = 24.0 * nv * n;
        f = viscous_flops(c) * n;
        break;
      }
      // read Q (central) or R, the constituent relations when viscous, the RK register; write Q and register
      double rd = (is_central(c->cfg.conv) ? nv : nv) + (c->need_prims ? c->nd + 2 : 0) + nv;
      b = 8.0 * (rd + 2.0 * nv) * n;
      f = (viscous_flops(c) + 6.0 * nv + (is_central(c->cfg.conv) ? c->nd * central_flops(c) : 0.0)) * n;
      break;
    }
  }
  if (bytes) *bytes = b;
  if (flops) *flops = f;
  return OSBL_OK;
}

int osbl_num_stages(osbl_ctx* c) { return c ? c->nstages : fail(OSBL_E_ARG, "null context"); }

}  // extern "C"

// ---------------------------------------------------------------------------
// Fixed device scratch of the allocation-free adjoint step.  One array per
// device (each device that runs the library's kernels holds its own copy
// until the process ends); host code addresses it only through the pointer
// hipGetSymbolAddress returns for that device.  Every call on a device
// shares that device's array: calls that overlap on one device race.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = (size_t)OSBL_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {

// ---------------------------------------------------------------------------
// discrete adjoint (central skew or LLF WENO5-Z, constant viscosity, periodic, RK3)
// ---------------------------------------------------------------------------
// the work buffers of one adjoint call
struct AdjBufs {
  double* Abuf;   // stage state with halos (nv padded fields)
  double* Mbuf;   // adjoint weights with halos (nv padded fields)
  double* pk;     // 3 packed arrays: lam_Q2, lam_Q1, lam_Qo
  double* vw;     // viscous VJP fields (padded)
  double* P;      // constituent relations over the padded grid
  double* vy;     // y part of a 2-D WENO5-Z VJP (packed), or null: two launches
};

static int adjoint_check(const osbl_ctx* c, bool need_rk3) {
  if (c->cfg.conv != OSBL_CONV_CENTRAL_SKEW && c->cfg.conv != OSBL_CONV_WENO_Z)
    return fail(OSBL_E_ARG, "the adjoint covers the central skew-symmetric and WENO5-Z schemes "
                            "(TENO's cut-off is not differentiable)");
  if (c->cfg.viscous == OSBL_VISC_SUTHERLAND)
    return fail(OSBL_E_ARG, "the adjoint covers constant viscosity (not Sutherland's law)");
  if (c->general || c->curv || c->cfg.filters)
    return fail(OSBL_E_ARG, "the adjoint covers periodic blocks without metric or filters");
  for (int d = 0; d < c->nd; ++d)
    if (c->cfg.bc[d][0] != OSBL_BC_PERIODIC) return fail(OSBL_E_ARG, "the adjoint needs periodic boundaries");
  if (need_rk3 && c->cfg.rk != OSBL_RK3) return fail(OSBL_E_ARG, "the adjoint step covers RungeKutta(3)");
  return OSBL_OK;
}

// which buffers the context's scheme uses
static bool adj_needs_halo_bufs(const osbl_ctx* c) {
  return c->cfg.conv == OSBL_CONV_CENTRAL_SKEW || c->cfg.viscous == OSBL_VISC_CONST;
}
static bool adj_needs_prims(const osbl_ctx* c) { return adj_needs_halo_bufs(c); }

// allocating set-up: the context's own buffers, made on first use
static int adjoint_ready(osbl_ctx* c, bool need_rk3, AdjBufs* b) {
  int rc = adjoint_check(c, need_rk3);
  if (rc) return rc;
  if (!c->Abuf) {
    const size_t bundle = (size_t)c->g.fs * c->nv * sizeof(double);
    const size_t packed = (size_t)cells(c) * c->nv * sizeof(double);
    if (hipMalloc((void**)&c->Abuf, bundle) != hipSuccess || hipMalloc((void**)&c->Mbuf, bundle) != hipSuccess ||
        hipMalloc((void**)&c->pk, 3 * packed) != hipSuccess)
      return fail(OSBL_E_NOMEM, "adjoint buffers");
    HIPCHK(hipMemsetAsync(c->Abuf, 0, bundle, c->stream));
    HIPCHK(hipMemsetAsync(c->Mbuf, 0, bundle, c->stream));
  }
  if (c->cfg.viscous == OSBL_VISC_CONST && !c->vw) {
    const size_t bytes = (size_t)c->g.fs * visc_vjp_fields(c->nd) * sizeof(double);
    if (hipMalloc((void**)&c->vw, bytes) != hipSuccess) return fail(OSBL_E_NOMEM, "adjoint viscous scratch");
    HIPCHK(hipMemsetAsync(c->vw, 0, bytes, c->stream));
  }
  if (adj_needs_prims(c) && !c->P) {
    // 3-D characteristic blocks evaluate their viscous terms without a
    // constituent-relation array; the viscous VJP needs one
    const size_t bytes = (size_t)c->g.fs * (c->nd + 3) * sizeof(double);
    if (hipMalloc((void**)&c->P, bytes) != hipSuccess) return fail(OSBL_E_NOMEM, "adjoint constituent relations");
    HIPCHK(hipMemsetAsync(c->P, 0, bytes, c->stream));
  }
  if (c->cfg.conv == OSBL_CONV_WENO_Z && c->nd == 2 && !c->vy) {
    if (hipMalloc((void**)&c->vy, (size_t)cells(c) * c->nv * sizeof(double)) != hipSuccess)
      return fail(OSBL_E_NOMEM, "adjoint VJP staging");
  }
  *b = AdjBufs{c->Abuf, c->Mbuf, c->pk, c->vw, c->P, c->vy};
  return OSBL_OK;
}

// out (packed) += (dR/dQ at q)^T mu
static void vjp(osbl_ctx* c, const AdjBufs& b, const double* q, const double* mu, double* out) {
  const long long n = cells(c);
  launch_unpack_prims(c->nd, q, b.Abuf, b.P, c->g, c->ph, c->stream);
  for (int m = 0; m < c->nv; ++m) launch_unpack(mu + m * n, b.Mbuf + m * c->g.fs, c->g, c->stream);
  launch_halo_periodic(b.Mbuf, c->nv, c->g, c->stream);
  launch_rhs_vjp(c->nd, b.Abuf, b.P, b.Mbuf, out, c->g, c->ph, c->stream);
}

// out (packed) += (dV/dQ at q)^T (alpha A + beta B) for the viscous terms V
static void visc_vjp(osbl_ctx* c, const AdjBufs& b, const double* q, const double* A, const double* B, double alpha,
                     double beta, double* out) {
  launch_visc_vjp(c->nd, q, A, B, alpha, beta, b.P, b.Mbuf, b.vw, out, c->g, c->ph, c->stream);
}

// out1 = c (A + B) + (dR/dQ at Qs)^T (alpha A + beta B), out2 = A + B (optional)
static void adj_stage(osbl_ctx* c, const AdjBufs& b, const double* Qs, const double* A, const double* B, double alpha,
                      double beta, double cc, double* out1, double* out2) {
  if (c->cfg.conv == OSBL_CONV_WENO_Z)
    launch_adj_stage_char(c->nd, c->cfg.averaging, Qs, A, B, alpha, beta, cc, out1, out2, c->g, c->ph, false,
                          c->stream, b.vy);
  else
    launch_adj_stage(c->nd, Qs, A, B, alpha, beta, cc, out1, out2, c->g, c->ph, c->stream);
  if (c->cfg.viscous == OSBL_VISC_CONST) visc_vjp(c, b, Qs, A, B, alpha, beta, out1);
}

// one RK3 step backwards: lam <- (dQ_{n+1}/dQ_n)^T lam
static int adjoint_step(osbl_ctx* c, const AdjBufs& bufs, const double* stages_dev, double* lam) {
  const long long n = cells(c) * c->nv;
  const double* Qn = stages_dev;
  const double* Q1 = stages_dev + n;
  const double* Q2 = stages_dev + 2 * n;
  double* l2 = bufs.pk;
  double* l1 = bufs.pk + n;
  double* lo = bufs.pk + 2 * n;
  const double dt = c->ph.dt;
  const double* a = c->A;  // rknew
  const double* b = c->B;  // rkold
  // Q3 = Qo2 + dt a2 R(Q2), Qo2 = lam's other path:  lam_Q2 = VJP(Q2; dt a2 lam)
  adj_stage(c, bufs, Q2, lam, nullptr, dt * a[2], 0.0, 0.0, l2, nullptr);
  // Q2 = Qo1 + dt a1 R(Q1), Qo2 = Qo1 + dt b1 R(Q1):
  //   lam_Q1 = VJP(Q1; dt a1 lam_Q2 + dt b1 lam_Qo2),  lam_Qo1 = lam_Q2 + lam_Qo2
  adj_stage(c, bufs, Q1, l2, lam, dt * a[1], dt * b[1], 0.0, l1, lo);
  // Q1 = Qn + dt a0 R(Qn), Qo1 = Qn + dt b0 R(Qn):
  //   lam_Qn = lam_Q1 + lam_Qo1 + VJP(Qn; dt a0 lam_Q1 + dt b0 lam_Qo1)
  adj_stage(c, bufs, Qn, l1, lo, dt * a[0], dt * b[0], 1.0, lam, nullptr);
  HIPCHK(hipGetLastError());
  return OSBL_OK;
}

// per-device address of the scratch array, looked up once per device
constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

// the scratch array's device address on device dev (made current for the lookup)
static int scratch_address(int dev, double** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(OSBL_E_ARG, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) HIPCHK(hipSetDevice(cur));
    if (e != hipSuccess) return fail(OSBL_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return OSBL_OK;
}

__global__ void __launch_bounds__(256) k_zero(double* __restrict__ p, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// the adjoint step's buffers carved from the scratch array (256-byte
// aligned pieces); the pieces the allocating set-up zeroes are cleared here
// by a fill kernel, since the array keeps the previous call's contents
static int scratch_bufs(osbl_ctx* c, AdjBufs* b) {
  double* base = nullptr;
  int rc = scratch_address(c->device, &base);
  if (rc) return rc;
  const long long fs = c->g.fs, packed = cells(c) * c->nv;
  long long off = 0;
  auto take = [&](long long ndoubles) {
    const long long at = off;
    off += (ndoubles + 31) / 32 * 32;
    return at;
  };
  const long long o_pk = take(3 * packed);
  const bool halo = adj_needs_halo_bufs(c);
  const long long o_A = halo ? take(fs * c->nv) : -1;
  const long long o_M = halo ? take(fs * c->nv) : -1;
  const bool visc = c->cfg.viscous == OSBL_VISC_CONST;
  const long long o_vw = visc ? take(fs * visc_vjp_fields(c->nd)) : -1;
  const long long o_P = adj_needs_prims(c) ? take(fs * (c->nd + 3)) : -1;
  const size_t need = (size_t)off * sizeof(double);
  if (need > kScratchBytes)
    return fail(OSBL_E_ARG, "osbl_adjoint_step_dev_noalloc: this block needs %zu bytes of scratch, more than the "
                            "%zu-byte fixed scratch (use osbl_adjoint_step_dev)", need, kScratchBytes);
  auto at = [&](long long o) { return o < 0 ? nullptr : base + o; };
  *b = AdjBufs{at(o_A), at(o_M), at(o_pk), at(o_vw), at(o_P), nullptr};
  // zero everything from the first halo buffer on (A, M, the viscous
  // fields, P): the allocating path's buffers start zeroed
  if (halo || visc) {
    const long long from = halo ? o_A : o_vw;
    const long long n = off - from;
    const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 8192);
    k_zero<<<blocks, 256, 0, c->stream>>>(base + from, n);
    HIPCHK(hipGetLastError());
  }
  return OSBL_OK;
}

}  // namespace

extern "C" {

int osbl_rhs_vjp_dev(osbl_ctx* c, const double* q_dev, const double* mu_dev, double* out_dev) {
  if (!c || !q_dev || !mu_dev || !out_dev) return fail(OSBL_E_ARG, "null argument");
  AdjBufs b;
  int rc = adjoint_ready(c, false, &b);
  if (rc) return rc;
  if (c->cfg.conv == OSBL_CONV_WENO_Z)
    launch_adj_stage_char(c->nd, c->cfg.averaging, q_dev, mu_dev, nullptr, 1.0, 0.0, 0.0, out_dev, nullptr, c->g,
                          c->ph, true, c->stream);
  else vjp(c, b, q_dev, mu_dev, out_dev);
  if (c->cfg.viscous == OSBL_VISC_CONST) visc_vjp(c, b, q_dev, mu_dev, nullptr, 1.0, 0.0, out_dev);
  HIPCHK(hipGetLastError());
  return OSBL_OK;
}

int osbl_step_stages_dev(osbl_ctx* c, double* stages_dev) {
  if (!c || !stages_dev) return fail(OSBL_E_ARG, "null argument");
  const long long n = cells(c);
  const int mode = c->cfg.rk == OSBL_RK3 ? 0 : 1;
  apply_bcs(c, Qcur(c));
  for (int s = 0; s < c->nstages; ++s) {
    launch_pack_all(Qcur(c), stages_dev + (long long)s * c->nv * n, c->g, c->nv, c->stream);
    stage(c, mode, c->A[s], c->B[s], s == 0);
    c->cur ^= 1;
    apply_bcs(c, Qcur(c));
  }
  HIPCHK(hipGetLastError());
  return OSBL_OK;
}

int osbl_adjoint_step_dev(osbl_ctx* c, const double* stages_dev, double* lam) {
  if (!c || !stages_dev || !lam) return fail(OSBL_E_ARG, "null argument");
  AdjBufs b;
  int rc = adjoint_ready(c, true, &b);
  if (rc) return rc;
  return adjoint_step(c, b, stages_dev, lam);
}

int osbl_adjoint_step_dev_noalloc(osbl_ctx* c, const double* stages_dev, double* lam) {
  if (!c || !stages_dev || !lam) return fail(OSBL_E_ARG, "null argument");
  int rc = adjoint_check(c, true);
  if (rc) return rc;
  AdjBufs b;
  if ((rc = scratch_bufs(c, &b))) return rc;
  return adjoint_step(c, b, stages_dev, lam);
}

}  // extern "C"
