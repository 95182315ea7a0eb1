/*
 * cpu_ref.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A CPU restatement of the kernel sequence that OpenSBLI's OPSC target emits
 * for the right-hand side and Runge-Kutta time advance.  It keeps the
 * structure of the generated code: OPS-style storage (one array per field,
 * x fastest, 5 halo points on every side, opsc.py:693-722), constituent
 * relations evaluated into arrays over their halo-extended ranges
 * (scheme.py:159-185, shock_capturing.py:119-131), one work array per
 * derivative or reconstructed flux, a residual kernel that assigns
 * (scheme.py:284-297, shock_capturing.py:21-34) followed by a viscous kernel
 * that accumulates, then the RK kernels (rk_sbli.py:102-133, rk_LS.py:139-166)
 * and the boundary exchanges (periodic.py:42-56, extrapolation.py:29-58).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library.  It is the checker, never the thing measured or shipped.
 *
 * Parity notes.  SymPy 1.1 rewrites expressions (factor, horner) before OPSC
 * prints them; this restatement evaluates the same formulas in a plain order,
 * which changes results only at the rounding level.  The sqrt(2.0) literal of
 * the eigenvector matrices (euler_eigensystem.py:109) is printed by SymPy to
 * 15 significant digits; here it is the exact double.  Both are "parity
 * unpinned" at the last-ulp level and irrelevant at the 1e-10 bar.
 */
#include "cpu_ref.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define H 5           /* OPS halo depth for every dataset, scheme.py:47-60 */
#define MAXV 5
#define NWK 64

struct ref_state {
  ref_config c;
  int nd, nvar;
  int n[3], h[3];
  ptrdiff_t sx, sy, sz, ntot, st[3];
  double* q[MAXV];
  double* qold[MAXV];  /* Q_RKold (3-register) or tempRK (low storage) */
  double* res[MAXV];
  double *u[3], *p, *T, *a, *mu, *theta;
  double* wk[NWK];
  int mod[3][2];      /* 1 + REF_BS_* where the face's BC replaces the central stencils, else 0 */
  double* dm[3];      /* metric D_dd of a stretched direction (n[d] values), else NULL */
  double* sd[3];      /* SD_ddd = CD(D_dd, xi_d) */
  double* plane[3][2];/* per-face conservative plane (Dirichlet values / inflow halo state) */
  double* qf[MAXV];   /* SFD filtered solution vfilt */
  int sfd_set[MAXV];  /* vfilt of variable m holds a value (set, or initialised from the state) */
  double* kappa;      /* WENO filter dissipation sensor (interior only; halos stay 0) */
  double* msens;      /* Yee Mach sensor of the WENO filter */
  double* tmp[MAXV];  /* binomial filter output */
  unsigned char* fmask; /* binomial filter grid condition (interior cells) or NULL */
  double* met[10];    /* curvilinear: D_ab at a*nd + b and detJ at nd*nd, whole padded arrays */
  double* U[3];       /* curvilinear: contravariant velocity U_a = D_ab u_b (a constituent relation) */
};

static inline ptrdiff_t IX(const ref_state* s, int i, int j, int k) {
  return (ptrdiff_t)(i + s->h[0]) + s->sx * ((ptrdiff_t)(j + s->h[1]) + s->sy * (ptrdiff_t)(k + s->h[2]));
}

typedef struct { int lo[3], hi[3]; } range3;

/* interior range, extended by ext[d] points on both sides of every active direction */
static range3 rng(const ref_state* s, int e0, int e1, int e2) {
  range3 r;
  int e[3] = {e0, e1, e2};
  for (int d = 0; d < 3; ++d) {
    if (d < s->nd) { r.lo[d] = -e[d]; r.hi[d] = s->n[d] + e[d]; }
    else { r.lo[d] = 0; r.hi[d] = 1; }
  }
  return r;
}
static range3 rng_all(const ref_state* s, int e) { return rng(s, e, e, e); }

#define LOOP(R)                                                                        \
  _Pragma("omp parallel for collapse(2) schedule(static) if(s->ntot > 200000)")        \
  for (int k = (R).lo[2]; k < (R).hi[2]; ++k)                                          \
    for (int j = (R).lo[1]; j < (R).hi[1]; ++j)                                        \
      for (int i = (R).lo[0]; i < (R).hi[0]; ++i)

/* ------------------------------------------------------------------------- */
/* Finite-difference weights (scheme.py:81-85 via sympy finite_diff_weights   */
/* on points -2..2, opensblifunctions.py:485-521).                            */
/* ------------------------------------------------------------------------- */
static const double W1[5] = {1.0 / 12.0, -2.0 / 3.0, 0.0, 2.0 / 3.0, -1.0 / 12.0};
static const double W2[5] = {-1.0 / 12.0, 4.0 / 3.0, -5.0 / 2.0, 4.0 / 3.0, -1.0 / 12.0};

void ref_central_weights(double* w1, double* w2) {
  for (int i = 0; i < 5; ++i) { w1[i] = W1[i]; w2[i] = W2[i]; }
}

/* ------------------------------------------------------------------------- */
/* Near-boundary replacements of the central stencils at faces whose BC is a  */
/* ModifyCentralDerivative (bc_core.py:15-16; opensblifunctions.py:523-534).  */
/* Rows are the side-0 stencils on f[0..5] at grid points 0, 1, ...; side 1   */
/* mirrors them on f[N-1], f[N-2], ... with the first derivative negated.     */
/* ------------------------------------------------------------------------- */
/* Carpenter 4th-order one-sided closure, bc4 = inv(al4) * ar4 with the R1, R2
   constants of Carpenter_scheme.py:80-102, evaluated here in long double. */
static double CARP1[4][6];
static int carp_ready = 0;
static void carpenter_init(void) {
  if (carp_ready) return;
  const long double R1 = -(2177.0L * sqrtl(295369.0L) - 1166427.0L) / 25488.0L;
  const long double R2 = (66195.0L * sqrtl(53.0L) * sqrtl(5573.0L) - 35909375.0L) / 101952.0L;
  long double al[4][4] = {
      {-(216 * R2 + 2160 * R1 - 2125) / 12960, (81 * R2 + 675 * R1 + 415) / 540, -(72 * R2 + 720 * R1 + 445) / 1440,
       -(108 * R2 + 756 * R1 + 421) / 1296},
      {(81 * R2 + 675 * R1 + 415) / 540, -(4104 * R2 + 32400 * R1 + 11225) / 4320,
       (1836 * R2 + 14580 * R1 + 7295) / 2160, -(216 * R2 + 2160 * R1 + 655) / 4320},
      {-(72 * R2 + 720 * R1 + 445) / 1440, (1836 * R2 + 14580 * R1 + 7295) / 2160,
       -(4104 * R2 + 32400 * R1 + 12785) / 4320, (81 * R2 + 675 * R1 + 335) / 540},
      {-(108 * R2 + 756 * R1 + 421) / 1296, -(216 * R2 + 2160 * R1 + 655) / 4320, (81 * R2 + 675 * R1 + 335) / 540,
       -(216 * R2 + 2160 * R1 - 12085) / 12960}};
  long double ar[4][6] = {
      {-1.0L / 2, -(864 * R2 + 6480 * R1 + 305) / 4320, (216 * R2 + 1620 * R1 + 725) / 540,
       -(864 * R2 + 6480 * R1 + 3335) / 4320, 0, 0},
      {(864 * R2 + 6480 * R1 + 305) / 4320, 0, -(864 * R2 + 6480 * R1 + 2315) / 1440, (108 * R2 + 810 * R1 + 415) / 270,
       0, 0},
      {-(216 * R2 + 1620 * R1 + 725) / 540, (864 * R2 + 6480 * R1 + 2315) / 1440, 0,
       -(864 * R2 + 6480 * R1 + 785) / 4320, -1.0L / 12, 0},
      {(864 * R2 + 6480 * R1 + 3335) / 4320, -(108 * R2 + 810 * R1 + 415) / 270, (864 * R2 + 6480 * R1 + 785) / 4320, 0,
       8.0L / 12, -1.0L / 12}};
  /* Gauss-Jordan with partial pivoting on [al | ar] */
  for (int c = 0; c < 4; ++c) {
    int piv = c;
    for (int r = c + 1; r < 4; ++r)
      if (fabsl(al[r][c]) > fabsl(al[piv][c])) piv = r;
    if (piv != c) {
      for (int k = 0; k < 4; ++k) { long double t = al[c][k]; al[c][k] = al[piv][k]; al[piv][k] = t; }
      for (int k = 0; k < 6; ++k) { long double t = ar[c][k]; ar[c][k] = ar[piv][k]; ar[piv][k] = t; }
    }
    long double inv = 1.0L / al[c][c];
    for (int k = 0; k < 4; ++k) al[c][k] *= inv;
    for (int k = 0; k < 6; ++k) ar[c][k] *= inv;
    for (int r = 0; r < 4; ++r) {
      if (r == c) continue;
      long double f = al[r][c];
      for (int k = 0; k < 4; ++k) al[r][k] -= f * al[c][k];
      for (int k = 0; k < 6; ++k) ar[r][k] -= f * ar[c][k];
    }
  }
  for (int r = 0; r < 4; ++r)
    for (int k = 0; k < 6; ++k) CARP1[r][k] = (double)ar[r][k];
  carp_ready = 1;
}
/* ReducedAccess first-derivative closure (reduced_access_scheme.py:34-40) */
static const double RA1[2][6] = {{-25.0 / 12.0, 48.0 / 12.0, -36.0 / 12.0, 16.0 / 12.0, -3.0 / 12.0, 0.0},
                                  {-3.0 / 12.0, -10.0 / 12.0, 18.0 / 12.0, -6.0 / 12.0, 1.0 / 12.0, 0.0}};
/* second-derivative closure of both schemes (Carpenter_scheme.py:69-77,
   reduced_access_scheme.py:75-83) */
static const double BC2[2][6] = {{35.0 / 12.0, -104.0 / 12.0, 114.0 / 12.0, -56.0 / 12.0, 11.0 / 12.0, 0.0},
                                 {11.0 / 12.0, -20.0 / 12.0, 6.0 / 12.0, 4.0 / 12.0, -1.0 / 12.0, 0.0}};

int ref_closure_weights(int scheme, int order, double* w) {
  carpenter_init();
  if (order == 2) {
    for (int r = 0; r < 2; ++r)
      for (int k = 0; k < 6; ++k) w[6 * r + k] = BC2[r][k];
    return 2;
  }
  if (scheme == REF_BS_CARPENTER) {
    for (int r = 0; r < 4; ++r)
      for (int k = 0; k < 6; ++k) w[6 * r + k] = CARP1[r][k];
    return 4;
  }
  for (int r = 0; r < 2; ++r)
    for (int k = 0; k < 6; ++k) w[6 * r + k] = RA1[r][k];
  return 2;
}

/* First derivative D_xi f (without 1/Delta) at the point x, grid index c of N
   along a line of stride st, with the closures lo/hi (1 + REF_BS_*, or 0).
   A closure row replaces the stencil only at its own grid index (the
   Equality(idx, loc) conditions of opensblifunctions.py:523-534, side 0
   tested first); halo points keep the central stencil. */
static double d1_gen(const double* f, ptrdiff_t x, ptrdiff_t st, int N, int lo, int hi, int c) {
  if (lo) {
    const int nr = lo - 1 == REF_BS_CARPENTER ? 4 : 2, np = lo - 1 == REF_BS_CARPENTER ? 6 : 5;
    if (c >= 0 && c < nr) {
      const double* w = lo - 1 == REF_BS_CARPENTER ? CARP1[c] : RA1[c];
      double acc = 0.0;
      for (int k = 0; k < np; ++k) acc += w[k] * f[x + (k - c) * st];
      return acc;
    }
  }
  if (hi) {
    const int nr = hi - 1 == REF_BS_CARPENTER ? 4 : 2, np = hi - 1 == REF_BS_CARPENTER ? 6 : 5;
    const int r = N - 1 - c;
    if (r >= 0 && r < nr) {
      const double* w = hi - 1 == REF_BS_CARPENTER ? CARP1[r] : RA1[r];
      double acc = 0.0;
      for (int k = 0; k < np; ++k) acc += w[k] * f[x + (r - k) * st];
      return -acc;
    }
  }
  return W1[0] * f[x - 2 * st] + W1[1] * f[x - st] + W1[3] * f[x + st] + W1[4] * f[x + 2 * st];
}
/* Homogeneous second derivative D_xixi f (without 1/Delta^2) */
static double d2_gen(const double* f, ptrdiff_t x, ptrdiff_t st, int N, int lo, int hi, int c) {
  if (lo && c >= 0 && c < 2) {
    double acc = 0.0;
    for (int k = 0; k < 5; ++k) acc += BC2[c][k] * f[x + (k - c) * st];
    return acc;
  }
  if (hi && N - 1 - c >= 0 && N - 1 - c < 2) {
    const int r = N - 1 - c;
    double acc = 0.0;
    for (int k = 0; k < 5; ++k) acc += BC2[r][k] * f[x + (r - k) * st];
    return acc;
  }
  return W2[0] * f[x - 2 * st] + W2[1] * f[x - st] + W2[2] * f[x] + W2[3] * f[x + st] + W2[4] * f[x + 2 * st];
}
static inline int coord(int d, int i, int j, int k) { return d == 0 ? i : (d == 1 ? j : k); }

/* out = D_d f (first derivative in the computational coordinate) over range R */
static void deriv1(ref_state* s, const double* f, double* out, int d, range3 R) {
  const ptrdiff_t st = s->st[d];
  const double inv = 1.0 / s->c.delta[d];
  const int N = s->n[d], lo = s->mod[d][0], hi = s->mod[d][1];
  LOOP(R) {
    ptrdiff_t x = IX(s, i, j, k);
    out[x] = inv * d1_gen(f, x, st, N, lo, hi, coord(d, i, j, k));
  }
}
/* out = D_dd f (homogeneous second derivative) over range R */
static void deriv2(ref_state* s, const double* f, double* out, int d, range3 R) {
  const ptrdiff_t st = s->st[d];
  const double inv = 1.0 / s->c.delta[d];
  const double inv2 = inv * inv;
  const int N = s->n[d], lo = s->mod[d][0], hi = s->mod[d][1];
  LOOP(R) {
    ptrdiff_t x = IX(s, i, j, k);
    out[x] = inv2 * d2_gen(f, x, st, N, lo, hi, coord(d, i, j, k));
  }
}

/* ------------------------------------------------------------------------- */
/* Constituent relations (taylor_green_vortex.py:29-31, Sod_shock_tube.py:26-28) */
/* Evaluated over the whole allocated range minus the outermost points; the   */
/* generated code evaluates them over interior + halotype, which is a subset, */
/* and values outside that subset are never read.                             */
/* ------------------------------------------------------------------------- */
static void ducros_sensor(ref_state* s);

static void constituent_relations(ref_state* s) {
  const int nd = s->nd;
  const double g = s->c.gama, M2 = s->c.Minf * s->c.Minf;
  range3 R = rng_all(s, H);
  double* rho = s->q[0];
  double* rhoE = s->q[nd + 1];
  LOOP(R) {
    ptrdiff_t x = IX(s, i, j, k);
    double ke = 0.0;
    for (int m = 0; m < nd; ++m) {
      double um = s->q[1 + m][x] / rho[x];
      s->u[m][x] = um;
      ke += um * um;
    }
    double pp = (g - 1.0) * (rhoE[x] - 0.5 * rho[x] * ke);
    s->p[x] = pp;
    s->T[x] = pp * g * M2 / rho[x];
    s->a[x] = sqrt(g * pp / rho[x]);
  }
  if (s->c.curvilinear) { /* Eq(U_i, D_i_j*u_j) (apps/euler_wave_curvilinear/euler_wave.py:46) */
    LOOP(R) {
      ptrdiff_t x = IX(s, i, j, k);
      for (int a = 0; a < nd; ++a) {
        double v = 0.0;
        for (int b = 0; b < nd; ++b) v += s->met[a * nd + b][x] * s->u[b][x];
        s->U[a][x] = v;
      }
    }
  }
  if (s->c.sensor) ducros_sensor(s);
  if (s->c.viscous == REF_VISC_SUTHERLAND) { /* Eq(mu, T**1.5*(1+SuthT/RefT)/(T+SuthT/RefT)), katzer_SBLI.py:26 */
    const double sr = s->c.suthT / s->c.refT;
    LOOP(R) {
      ptrdiff_t x = IX(s, i, j, k);
      s->mu[x] = pow(s->T[x], 1.5) * (1.0 + sr) / (s->T[x] + sr);
    }
  }
}

/* metric factor D_dd at grid index c of a stretched direction (1 otherwise);
   halo points take the mirrored values of the metric boundary kernel
   (metric.py:359-379: D[-t] = D[t], D[N-1+t] = D[N-1-t]) */
static inline double met1(const ref_state* s, int d, int c) {
  if (!s->dm[d]) return 1.0;
  const int N = s->n[d];
  if (c < 0) c = -c;
  if (c > N - 1) c = 2 * (N - 1) - c;
  return s->dm[d][c];
}

/* Modified Ducros sensor (shock_sensors.py:12-49) at one point:
   1/2 (1 - tanh(2.5 (1 + 100 div u))) div^2 / (div^2 + |curl u|^2 + epsilon),
   with the block's central derivatives (closures and metric included). */
static double ducros_at(const ref_state* s, ptrdiff_t x, int i, int j, int k) {
  const int nd = s->nd;
  const int cc[3] = {i, j, k};
  double du[3][3] = {{0.0}};
  for (int m = 0; m < nd; ++m)
    for (int d = 0; d < nd; ++d)
      du[m][d] = met1(s, d, cc[d]) * ((1.0 / s->c.delta[d]) *
                 d1_gen(s->u[m], x, s->st[d], s->n[d], s->mod[d][0], s->mod[d][1], cc[d]));
  double div = 0.0, w2;
  for (int d = 0; d < nd; ++d) div += du[d][d];
  if (nd == 2) {
    const double w = du[1][0] - du[0][1];
    w2 = w * w;
  } else {
    const double wa = du[2][1] - du[1][2], wb = du[0][2] - du[2][0], wc = du[1][0] - du[0][1];
    w2 = wa * wa + wb * wb + wc * wc;
  }
  const double filt = 0.5 * (1.0 - tanh(2.5 * (1.0 + 100.0 * div)));
  return filt * (div * div) / (div * div + w2 + s->c.sensor_eps);
}

/* The sensor as a constituent relation of the adaptive TENO scheme, over the
   interior and three halo points. */
static void ducros_sensor(ref_state* s) {
  range3 R = rng_all(s, 3);
  LOOP(R) {
    ptrdiff_t x = IX(s, i, j, k);
    s->theta[x] = ducros_at(s, x, i, j, k);
  }
}

/* adaptive TENO cut-off at an interface from the sensor at its left point
   (teno.py:430-443): C_T = 1/10^floor(a1 - a2 (1 - (1-theta)^4 (1 + 4 theta))) */
static double adaptive_ct(const ref_state* s, double m) {
  const double t = 1.0 - m;
  const double t4 = (t * t) * (t * t);
  const double beta = s->c.teno_a1 - s->c.teno_a2 * (1.0 - t4 * (1.0 + 4.0 * m));
  const int n = (int)floor(beta);
  double r = 1.0; /* 10^|n| by products, exact for |n| <= 22 */
  for (int i = 0; i < (n < 0 ? -n : n); ++i) r *= 10.0;
  return n < 0 ? r : 1.0 / r;
}

/* ------------------------------------------------------------------------- */
/* Central convective terms                                                  */
/* ------------------------------------------------------------------------- */
/* Skew (Blaisdell) form, parsing.py:75-111 and scheme.py:187-252:            */
/*   R_rho  = -1/2 sum_j [D_j(rho u_j) + u_j D_j rho + rho D_j u_j]            */
/*   R_rhoui= -1/2 sum_j [D_j(rhou_i u_j) + u_j D_j rhou_i + rhou_i D_j u_j]   */
/*            - D_i p                                                          */
/*   R_rhoE = -1/2 sum_j [D_j(rhoE u_j) + u_j D_j rhoE + rhoE D_j u_j]         */
/*            - sum_j D_j(p u_j)                                               */
/* Product fields are evaluated into work arrays over interior +-2 in j       */
/* (scheme.py:218-225) and every derivative into its own work array.          */
static void central_skew(ref_state* s) {
  const int nd = s->nd, nv = s->nvar;
  range3 I = rng_all(s, 0);
  for (int m = 0; m < nv; ++m) {
    double* r = s->res[m];
    LOOP(I) r[IX(s, i, j, k)] = 0.0;
  }
  for (int d = 0; d < nd; ++d) {
    int e[3] = {0, 0, 0};
    e[d] = 2;
    range3 P = rng(s, e[0], e[1], e[2]);
    /* products: wk0 = rho*u_d, wk1+m = rhou_m*u_d, wk(1+nd) = rhoE*u_d, wk(2+nd) = p*u_d */
    double** w = s->wk;
    const double* ud = s->u[d];
    LOOP(P) {
      ptrdiff_t x = IX(s, i, j, k);
      w[0][x] = s->q[0][x] * ud[x];
      for (int m = 0; m < nd; ++m) w[1 + m][x] = s->q[1 + m][x] * ud[x];
      w[1 + nd][x] = s->q[nd + 1][x] * ud[x];
      w[2 + nd][x] = s->p[x] * ud[x];
    }
    /* derivative work arrays: 16.. */
    const int D0 = 16;
    double* dprod[8];
    double* dq[MAXV];
    for (int t = 0; t < nd + 3; ++t) {
      dprod[t] = w[D0 + t];
      deriv1(s, w[t], dprod[t], d, I);
    }
    for (int m = 0; m < nv; ++m) {
      dq[m] = w[D0 + 8 + m];
      deriv1(s, s->q[m], dq[m], d, I);
    }
    double* dud = w[D0 + 14];
    double* dp = w[D0 + 15];
    deriv1(s, ud, dud, d, I);
    deriv1(s, s->p, dp, d, I);
    /* a stretched direction scales every derivative by D_dd (metric.py:72-126) */
    if (s->dm[d]) {
      double* all[16];
      int na = 0;
      for (int t = 0; t < nd + 3; ++t) all[na++] = dprod[t];
      for (int m = 0; m < nv; ++m) all[na++] = dq[m];
      all[na++] = dud;
      all[na++] = dp;
      for (int t = 0; t < na; ++t) {
        double* f = all[t];
        LOOP(I) { ptrdiff_t x = IX(s, i, j, k); f[x] = met1(s, d, coord(d, i, j, k)) * f[x]; }
      }
    }
    /* residual contributions of direction d (the generated residual kernel sums over j) */
    LOOP(I) {
      ptrdiff_t x = IX(s, i, j, k);
      double u_ = ud[x], du_ = dud[x];
      s->res[0][x] += -0.5 * (dprod[0][x] + u_ * dq[0][x] + s->q[0][x] * du_);
      for (int m = 0; m < nd; ++m) {
        double t = -0.5 * (dprod[1 + m][x] + u_ * dq[1 + m][x] + s->q[1 + m][x] * du_);
        if (m == d) t -= dp[x];
        s->res[1 + m][x] += t;
      }
      s->res[nd + 1][x] += -0.5 * (dprod[1 + nd][x] + u_ * dq[nd + 1][x] + s->q[nd + 1][x] * du_) - dprod[2 + nd][x];
    }
  }
}

/* Conservative central form: R = -sum_j D_j F_j with F_j the Euler flux       */
/* (Conservative(rhou_j,x_j), Conservative(rhou_i*u_j + KD(_i,_j)*p, x_j),     */
/* Conservative((p+rhoE)*u_j, x_j); parsing.py:189-216, scheme.py:218-225).   */
static void central_conservative(ref_state* s) {
  const int nd = s->nd, nv = s->nvar;
  range3 I = rng_all(s, 0);
  for (int m = 0; m < nv; ++m) {
    double* r = s->res[m];
    LOOP(I) r[IX(s, i, j, k)] = 0.0;
  }
  for (int d = 0; d < nd; ++d) {
    int e[3] = {0, 0, 0};
    e[d] = 2;
    range3 P = rng(s, e[0], e[1], e[2]);
    double** w = s->wk;
    const double* ud = s->u[d];
    LOOP(P) {
      ptrdiff_t x = IX(s, i, j, k);
      for (int m = 0; m < nd; ++m) w[1 + m][x] = s->q[1 + m][x] * ud[x] + (m == d ? s->p[x] : 0.0);
      w[1 + nd][x] = (s->p[x] + s->q[nd + 1][x]) * ud[x];
    }
    deriv1(s, s->q[1 + d], w[20], d, I);
    for (int m = 0; m < nd; ++m) deriv1(s, w[1 + m], w[21 + m], d, I);
    deriv1(s, w[1 + nd], w[21 + nd], d, I);
    LOOP(I) {
      ptrdiff_t x = IX(s, i, j, k);
      const double g1 = met1(s, d, coord(d, i, j, k));
      if (s->dm[d]) {
        s->res[0][x] -= g1 * w[20][x];
        for (int m = 0; m < nd + 1; ++m) s->res[1 + m][x] -= g1 * w[21 + m][x];
      } else {
        s->res[0][x] -= w[20][x];
        for (int m = 0; m < nd + 1; ++m) s->res[1 + m][x] -= w[21 + m][x];
      }
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Viscous central terms (taylor_green_vortex.py:14-19 with constant          */
/* viscosity; katzer_SBLI.py:12-26 with mu(T) a constituent relation),        */
/* expanded by the product rule of Der.differentiate (parsing.py:141-186):    */
/*   D_j tau_mj = (mu/Re)(D_jj u_m + D_j D_m u_j - 2/3 d_mj D_m D_k u_k)       */
/*              + (1/Re) D_j mu (D_j u_m + D_m u_j - 2/3 d_mj D_k u_k)          */
/*   -D_j q_j   = cq (mu D_jj T + D_j mu D_j T),  cq = 1/((g-1) M^2 Pr Re)       */
/*   D_j(u_m tau_mj) = D_j u_m tau_mj + u_m D_j tau_mj                          */
/* Every term containing Re is accumulated into the residual                  */
/* (scheme.py:263-267,291-294).  Mixed derivatives are D_b(D_a u) with a < b   */
/* (opensblifunctions.py:341-353), the inner one evaluated over +-2 halo       */
/* points in b (scheme.py:360-387).  With a stretched direction every          */
/* derivative is transformed (metric.py:72-212): D_j -> D_jj CD_xi_j,          */
/* D_jj -> D_jj^2 CD2_xi_j + D_jj SD_jjj CD_xi_j, D_a D_b -> D_aa D_bb CD_ab.   */
/* ------------------------------------------------------------------------- */
static void viscous(ref_state* s) {
  const int nd = s->nd;
  const double Re = s->c.Re, g = s->c.gama, M = s->c.Minf, Pr = s->c.Pr;
  const double inv_re = 1.0 / Re;
  const double cq = 1.0 / ((g - 1.0) * M * M * Pr * Re);
  const int suth = s->c.viscous == REF_VISC_SUTHERLAND;
  double** w = s->wk;
  /* work layout: du[m][j] at 0+3m+j, d2u[m][j] at 9+3m+j, d2T[j] at 18+j, mixed at 21..29,
     dT[j] at 30+j, dmu[j] at 33+j */
  double* du[3][3];
  double* d2u[3][3];
  double* d2T[3];
  double* dT[3];
  double* dmu[3];
  double* mix[3][3][3]; /* mix[m][a][b], a<b: D_b(D_a u_m) */
  for (int m = 0; m < 3; ++m)
    for (int jj = 0; jj < 3; ++jj) { du[m][jj] = w[3 * m + jj]; d2u[m][jj] = w[9 + 3 * m + jj]; }
  for (int jj = 0; jj < 3; ++jj) { d2T[jj] = w[18 + jj]; dT[jj] = w[30 + jj]; dmu[jj] = w[33 + jj]; }
  int nm = 21;
  for (int m = 0; m < 3; ++m)
    for (int a_ = 0; a_ < 3; ++a_)
      for (int b = 0; b < 3; ++b) mix[m][a_][b] = (a_ < b && a_ < nd && b < nd) ? w[nm++] : NULL;
  int metric = 0;
  for (int d = 0; d < nd; ++d) metric |= s->dm[d] != NULL;

  range3 I = rng_all(s, 0);
  for (int m = 0; m < nd; ++m)
    for (int a_ = 0; a_ < nd; ++a_) {
      int e[3] = {2, 2, 2};
      e[a_] = 0;
      deriv1(s, s->u[m], du[m][a_], a_, rng(s, e[0], e[1], e[2]));
      deriv2(s, s->u[m], d2u[m][a_], a_, I);
    }
  for (int a_ = 0; a_ < nd; ++a_) deriv2(s, s->T, d2T[a_], a_, I);
  if (suth || metric)
    for (int a_ = 0; a_ < nd; ++a_) deriv1(s, s->T, dT[a_], a_, I);
  if (suth)
    for (int a_ = 0; a_ < nd; ++a_) deriv1(s, s->mu, dmu[a_], a_, I);
  for (int m = 0; m < nd; ++m)
    for (int a_ = 0; a_ < nd; ++a_)
      for (int b = a_ + 1; b < nd; ++b) deriv1(s, du[m][a_], mix[m][a_][b], b, I);

  LOOP(I) {
    ptrdiff_t x = IX(s, i, j, k);
    const int cc[3] = {i, j, k};
    double G[3], S2[3];
    for (int d = 0; d < 3; ++d) {
      G[d] = d < nd ? met1(s, d, cc[d]) : 1.0;
      S2[d] = (d < nd && s->sd[d]) ? s->sd[d][cc[d]] : 0.0;
    }
    /* transformed first and second derivatives */
    double DU[3][3], DD[3][3][3], D2T[3], DT[3], DMU[3];
    for (int m = 0; m < nd; ++m)
      for (int a_ = 0; a_ < nd; ++a_) {
        DU[m][a_] = metric ? G[a_] * du[m][a_][x] : du[m][a_][x];
        for (int b = 0; b < nd; ++b) {
          double v;
          if (a_ == b) v = metric ? G[a_] * G[a_] * d2u[m][a_][x] + G[a_] * S2[a_] * du[m][a_][x] : d2u[m][a_][x];
          else {
            const double mx = a_ < b ? mix[m][a_][b][x] : mix[m][b][a_][x];
            v = metric ? G[a_] * G[b] * mx : mx;
          }
          DD[m][a_][b] = v;
        }
      }
    for (int a_ = 0; a_ < nd; ++a_) {
      D2T[a_] = metric ? G[a_] * G[a_] * d2T[a_][x] + G[a_] * S2[a_] * dT[a_][x] : d2T[a_][x];
      DT[a_] = (suth || metric) ? (metric ? G[a_] * dT[a_][x] : dT[a_][x]) : 0.0;
      DMU[a_] = suth ? (metric ? G[a_] * dmu[a_][x] : dmu[a_][x]) : 0.0;
    }
    const double mu_ = suth ? s->mu[x] : 1.0;
    double divu = 0.0;
    for (int kk = 0; kk < nd; ++kk) divu += DU[kk][kk];
    double energy = 0.0;
    for (int jj = 0; jj < nd; ++jj) {
      if (suth) energy += cq * (mu_ * D2T[jj] + DMU[jj] * DT[jj]);
      else energy += cq * D2T[jj];
    }
    for (int m = 0; m < nd; ++m) {
      double visc = 0.0;
      for (int jj = 0; jj < nd; ++jj) {
        double t = DD[m][jj][jj] + DD[jj][m][jj];
        if (jj == m) {
          double sk = 0.0;
          for (int kk = 0; kk < nd; ++kk) sk += DD[kk][m][kk];
          t -= (2.0 / 3.0) * sk;
        }
        const double sr = DU[m][jj] + DU[jj][m] - (jj == m ? (2.0 / 3.0) * divu : 0.0);
        double dtau, tau;
        if (suth) {
          dtau = inv_re * mu_ * t + inv_re * DMU[jj] * sr;
          tau = inv_re * mu_ * sr;
        } else {
          dtau = inv_re * t;
          tau = inv_re * sr;
        }
        visc += dtau;
        energy += DU[m][jj] * tau + s->u[m][x] * dtau;
      }
      s->res[1 + m][x] += visc;
    }
    s->res[nd + 1][x] += energy;
  }
}

/* ------------------------------------------------------------------------- */
/* WENO / TENO reconstruction (weno.py, teno.py)                              */
/* f points at f[0] of a window f[-2..3].                                     */
/* ------------------------------------------------------------------------- */
/* Jiang-Shu smoothness indicators of WENO5 (weno.py:138-205): beta_r is the
 * sum over l = 1, 2 of h^(2l-1) * integral over the cell of (d^l p_r/dx^l)^2
 * for the quadratic p_r through the stencil's nodes, applied with shift 0
 * (right) or 1 (left).  The reference prints these as the expanded quadratic
 * form in Horner order; that form cancels to beta from terms of size f^2 and
 * carries rounding noise of order 1e-16*f^2, which WENO-Z's eps = 1e-14
 * amplifies in near-uniform regions.  The same quadratic form is evaluated
 * here as 13/12 (a - 2b + c)^2 + 1/4 (first-difference combination)^2, which
 * is algebraically identical and carries rounding relative to beta itself.
 * Agreement with the reference's own floating-point noise at that level is
 * unpinned. */
static void js_beta(const double* f, int shift, double* beta) {
  /* beta_r on nodes f[-r+shift], f[-r+1+shift], f[-r+2+shift] */
  for (int r = 0; r < 3; ++r) {
    const double* g = f - r + shift;
    const double a = g[0], b = g[1], c = g[2];
    const double t = a - 2.0 * b + c;
    double u;
    if (r == 0) u = 3.0 * a - 4.0 * b + c;
    else if (r == 1) u = a - c;
    else u = a - 4.0 * b + 3.0 * c;
    beta[r] = 13.0 / 12.0 * t * t + 0.25 * u * u;
  }
}

/* ENO coefficients c_rj and optimal weights d_r of WENO5 (weno.py:70-120),   */
/* func_points = -r+j+shift (weno.py:122-135).                                */
static const double WENO_C_R[3][3] = {{1.0 / 3.0, 5.0 / 6.0, -1.0 / 6.0}, {-1.0 / 6.0, 5.0 / 6.0, 1.0 / 3.0},
                                      {1.0 / 3.0, -7.0 / 6.0, 11.0 / 6.0}};
static const double WENO_D_R[3] = {3.0 / 10.0, 3.0 / 5.0, 1.0 / 10.0};
static const double WENO_C_L[3][3] = {{11.0 / 6.0, -7.0 / 6.0, 1.0 / 3.0}, {1.0 / 3.0, 5.0 / 6.0, -1.0 / 6.0},
                                      {-1.0 / 6.0, 5.0 / 6.0, 1.0 / 3.0}};
static const double WENO_D_L[3] = {1.0 / 10.0, 3.0 / 5.0, 3.0 / 10.0};

static double weno5(int conv, int side, const double* f, double* omega_out) {
  int shift = side > 0 ? 0 : 1;
  double beta[3], alpha[3];
  js_beta(f, shift, beta);
  const double(*c)[3] = side > 0 ? WENO_C_R : WENO_C_L;
  const double* dr = side > 0 ? WENO_D_R : WENO_D_L;
  if (conv == REF_CONV_WENO_Z) {
    const double eps = 1.0e-14;          /* weno.py:286 */
    double tau = fabs(beta[0] - beta[2]); /* weno.py:300-301 */
    for (int r = 0; r < 3; ++r) {
      double t = tau / (eps + beta[r]);
      alpha[r] = dr[r] * (1.0 + t * t);  /* weno.py:320-323, p = 2 */
    }
  } else {
    const double eps = 1e-6;             /* weno.py:344 */
    for (int r = 0; r < 3; ++r) alpha[r] = dr[r] / ((eps + beta[r]) * (eps + beta[r]));
  }
  double inv_alpha_sum = 1.0 / (alpha[0] + alpha[1] + alpha[2]);
  double recon = 0.0;
  for (int r = 0; r < 3; ++r) {
    double omega = alpha[r] * inv_alpha_sum;
    if (omega_out) omega_out[r] = omega;
    double eno = 0.0;
    for (int jj = 0; jj < 3; ++jj) eno += c[r][jj] * f[-r + jj + shift];
    recon += omega * eno;
  }
  return recon;
}

/* TENO candidate stencils, coefficients and optimal weights, literal in      */
/* teno.py:95-141; smoothness indicators literal in teno.py:159-177 including */
/* the unsquared 781/720 term of the right-side fourth stencil.               */
static const int TENO_PTS_R[4][4] = {{-1, 0, 1, 0}, {0, 1, 2, 0}, {-2, -1, 0, 0}, {0, 1, 2, 3}};
static const int TENO_PTS_L[4][4] = {{0, 1, 2, 0}, {-1, 0, 1, 0}, {1, 2, 3, 0}, {-2, -1, 0, 1}};
static const double TENO_C_R[4][4] = {{-1.0 / 6.0, 5.0 / 6.0, 2.0 / 6.0, 0.0},
                                      {2.0 / 6.0, 5.0 / 6.0, -1.0 / 6.0, 0.0},
                                      {2.0 / 6.0, -7.0 / 6.0, 11.0 / 6.0, 0.0},
                                      {3.0 / 12.0, 13.0 / 12.0, -5.0 / 12.0, 1.0 / 12.0}};
static const double TENO_C_L[4][4] = {{2.0 / 6.0, 5.0 / 6.0, -1.0 / 6.0, 0.0},
                                      {-1.0 / 6.0, 5.0 / 6.0, 2.0 / 6.0, 0.0},
                                      {11.0 / 6.0, -7.0 / 6.0, 1.0 / 3.0, 0.0},
                                      {1.0 / 12.0, -5.0 / 12.0, 13.0 / 12.0, 3.0 / 12.0}};
static const double TENO6_D[4] = {231.0 / 500.0, 3.0 / 10.0, 27.0 / 500.0, 23.0 / 125.0};
static const double TENO5_D[3] = {11.0 / 20.0, 4.0 / 10.0, 1.0 / 20.0};

static void teno_beta(int side, const double* f, double* b) {
  double q;
  if (side > 0) {
    q = f[-1] - f[1];
    b[0] = 0.25 * q * q;
    q = f[-1] - 2.0 * f[0] + f[1];
    b[0] += 13.0 / 12.0 * q * q;
    q = 3.0 * f[0] - 4.0 * f[1] + f[2];
    b[1] = 0.25 * q * q;
    q = f[0] - 2.0 * f[1] + f[2];
    b[1] += 13.0 / 12.0 * q * q;
    q = f[-2] - 4.0 * f[-1] + 3.0 * f[0];
    b[2] = 0.25 * q * q;
    q = f[-2] - 2.0 * f[-1] + f[0];
    b[2] += 13.0 / 12.0 * q * q;
    q = -11.0 * f[0] + 18.0 * f[1] - 9.0 * f[2] + 2.0 * f[3];
    b[3] = 1.0 / 36.0 * q * q;
    q = 2.0 * f[0] - 5.0 * f[1] + 4.0 * f[2] - f[3];
    b[3] += 13.0 / 12.0 * q * q;
    q = -f[0] + 3.0 * f[1] - 3.0 * f[2] + f[3];
    b[3] += 781.0 / 720.0 * q; /* teno.py:166-167: not squared */
  } else {
    q = f[0] - f[2];
    b[0] = 0.25 * q * q;
    q = f[0] - 2.0 * f[1] + f[2];
    b[0] += 13.0 / 12.0 * q * q;
    q = f[-1] - 4.0 * f[0] + 3.0 * f[1];
    b[1] = 0.25 * q * q;
    q = f[-1] - 2.0 * f[0] + f[1];
    b[1] += 13.0 / 12.0 * q * q;
    q = 3.0 * f[1] - 4.0 * f[2] + f[3];
    b[2] = 0.25 * q * q;
    q = f[1] - 2.0 * f[2] + f[3];
    b[2] += 13.0 / 12.0 * q * q;
    q = -2.0 * f[-2] + 9.0 * f[-1] - 18.0 * f[0] + 11.0 * f[1];
    b[3] = 1.0 / 36.0 * q * q;
    q = -f[-2] + 4.0 * f[-1] - 5.0 * f[0] + 2.0 * f[1];
    b[3] += 13.0 / 12.0 * q * q;
    q = -f[-2] + 3.0 * f[-1] - 3.0 * f[0] + f[1];
    b[3] += 781.0 / 720.0 * q * q;
  }
}

static double teno(int conv, int side, const double* f, double eps, double ct) {
  const int ns = conv == REF_CONV_TENO6 ? 4 : 3;
  const double* d = conv == REF_CONV_TENO6 ? TENO6_D : TENO5_D;
  const int(*pts)[4] = side > 0 ? TENO_PTS_R : TENO_PTS_L;
  const double(*c)[4] = side > 0 ? TENO_C_R : TENO_C_L;
  double b[4], alpha[4], delta[4];
  teno_beta(side, f, b);
  double tau = ns == 4 ? fabs(b[3] - (b[0] + b[2] - 4.0 * b[1]) / 6.0) /* teno.py:230 */
                       : fabs(b[0] - b[2]);                           /* teno.py:209 */
  double asum = 0.0;
  for (int r = 0; r < ns; ++r) {
    alpha[r] = pow(1.0 + tau / (eps + b[r]), 6.0); /* C = 1, q = 6 */
    asum += alpha[r];
  }
  double inv_alpha_sum = 1.0 / asum;
  double dsum = 0.0;
  for (int r = 0; r < ns; ++r) {
    delta[r] = (alpha[r] * inv_alpha_sum < ct) ? 0.0 : 1.0; /* teno.py:445-465 */
    dsum += d[r] * delta[r];
  }
  double inv_omega_sum = 1.0 / dsum;
  double recon = 0.0;
  for (int r = 0; r < ns; ++r) {
    int w = (r == 3) ? 4 : 3;
    double eno = 0.0;
    for (int jj = 0; jj < w; ++jj) eno += f[pts[r][jj]] * c[r][jj];
    recon += d[r] * delta[r] * inv_omega_sum * eno;
  }
  return recon;
}

double ref_recon_side(int conv, int side, const double* f6, double eps, double ct) {
  const double* f = f6 + 2;
  if (conv == REF_CONV_WENO_JS || conv == REF_CONV_WENO_Z) return weno5(conv, side, f, NULL);
  return teno(conv, side, f, eps, ct);
}

int ref_smoothness(int conv, int side, const double* f6, double* beta) {
  const double* f = f6 + 2;
  if (conv == REF_CONV_WENO_JS || conv == REF_CONV_WENO_Z) {
    js_beta(f, side > 0 ? 0 : 1, beta);
    return 3;
  }
  teno_beta(side, f, beta);
  return conv == REF_CONV_TENO6 ? 4 : 3;
}

/* Shock sensor of one WENO-JS reconstruction from its non-linear weights
   (weno.py:415-431, normalised, sensor_theta = 4):
   sum_r |omega_r/d_r - 1|^4 / (|1/min(d) - 1|^4 + k - 1). */
static double weno_sensor(const double* omega, const double* d) {
  double s = 0.0;
  for (int r = 0; r < 3; ++r) s += pow(fabs(omega[r] / d[r] - 1.0), 4.0);
  return s / (pow(fabs(1.0 / 0.1 - 1.0), 4.0) + 2.0);
}

/* WENO filter flux of one characteristic field (shock_capturing.py:82-96,
   401-416, 438-443): the WENO-JS reconstructions of f+ (right) and f- (left)
   summed, minus the sixth-order central flux of the characteristic fluxes on
   points -2..3, weighted by rj = max(sensor(right), sensor(left)). */
double ref_weno_filter_flux(const double* fp6, const double* fm6, const double* cf6, double* rj_out) {
  double om[3];
  const double right = weno5(REF_CONV_WENO_JS, 1, fp6 + 2, om);
  double rj = weno_sensor(om, WENO_D_R);
  const double left = weno5(REF_CONV_WENO_JS, -1, fm6 + 2, om);
  rj = fmax(rj, weno_sensor(om, WENO_D_L));
  const double central = (cf6[0] - 8.0 * cf6[1] + 37.0 * cf6[2] + 37.0 * cf6[3] - 8.0 * cf6[4] + cf6[5]) / 60.0;
  if (rj_out) *rj_out = rj;
  return rj * ((right + left) - central);
}

/* ------------------------------------------------------------------------- */
/* Euler eigensystems (euler_eigensystem.py:42-135), unit normal k = e_d for  */
/* Cartesian grids (metrics diagonal, euler_eigensystem.py:49-54).           */
/* L and R row-major nvar x nvar; 1/a and 1/rho used as inv_AVG_a/inv_AVG_rho */
/* (shock_capturing.py:319-331,428-437).                                     */
/* ------------------------------------------------------------------------- */
void ref_eigensystem(int nd, int d, double g, double rho, const double* u, double a, double* L, double* R,
                     double* lam) {
  if (nd == 1) {
    const double ia = 1.0 / a, gm1 = g - 1.0;
    double u0 = u[0];
    double Hh = a * a / gm1 + u0 * u0 / 2.0;
    double den = 2.0 * Hh - u0 * u0;
    lam[0] = u0 - a; lam[1] = u0; lam[2] = u0 + a;
    double Rm[9] = {1, 1, 1, u0 - a, u0, u0 + a, Hh - u0 * a, u0 * u0 / 2.0, Hh + u0 * a};
    double Lm[9] = {u0 * (2.0 * Hh + a * u0 - u0 * u0) * ia / (2.0 * den),
                    (-Hh - a * u0 + u0 * u0 / 2.0) * ia / den,
                    1.0 / den,
                    2.0 * (Hh - u0 * u0) / den,
                    2.0 * u0 / den,
                    2.0 / (-2.0 * Hh + u0 * u0),
                    u0 * (-2.0 * Hh + a * u0 + u0 * u0) * ia / (2.0 * den),
                    (Hh - a * u0 - u0 * u0 / 2.0) * ia / den,
                    1.0 / den};
    memcpy(R, Rm, sizeof(Rm));
    memcpy(L, Lm, sizeof(Lm));
    return;
  }
  double kv[3] = {0, 0, 0};
  kv[d] = 1.0;
  ref_eigensystem_k(nd, g, rho, u, a, kv, L, R, lam);
}

/* About the metric direction k = (D_d0, .., D_d(nd-1)) of a curvilinear grid
   (euler_eigensystem.py:18-40): k_t = k/|k| (|k| replaced by its inverse
   inv_AVG_met_fact, shock_capturing.py:319-331), theta_t = k_t.u in L and R,
   eigenvalues U, U(, U), U + a|k|, U - a|k| with U = k.u.  k = e_d gives the
   Cartesian system exactly. */
void ref_eigensystem_k(int nd, double g, double rho, const double* u, double a, const double* k, double* L,
                       double* R, double* lam) {
  const double ia = 1.0 / a, ir = 1.0 / rho, gm1 = g - 1.0;
  const double s2 = sqrt(2.0);
  const double alpha = rho / (a * s2), bta = ir * ia / s2;
  double kk = 0.0, U = 0.0;
  for (int m = 0; m < nd; ++m) { kk += k[m] * k[m]; U += k[m] * u[m]; }
  const double kmag = sqrt(kk), ikm = 1.0 / kmag;
  double kt[3] = {0, 0, 0};
  for (int m = 0; m < nd; ++m) kt[m] = k[m] * ikm;
  if (nd == 2) {
    double u0 = u[0], u1 = u[1];
    double th = kt[0] * u0 + kt[1] * u1;
    double phi = gm1 * ((u0 * u0 + u1 * u1) / 2.0);
    lam[0] = U; lam[1] = U; lam[2] = U + a * kmag; lam[3] = U - a * kmag;
    double Rm[16] = {1, 0, alpha, alpha,
                     u0, kt[1] * rho, alpha * (u0 + kt[0] * a), alpha * (u0 - kt[0] * a),
                     u1, -kt[0] * rho, alpha * (u1 + kt[1] * a), alpha * (u1 - kt[1] * a),
                     phi / gm1, rho * (kt[1] * u0 - kt[0] * u1), alpha * ((phi + a * a) / gm1 + a * th),
                     alpha * ((phi + a * a) / gm1 - a * th)};
    double Lm[16] = {1 - phi * ia * ia, gm1 * u0 * ia * ia, gm1 * u1 * ia * ia, -gm1 * ia * ia,
                     -(kt[1] * u0 - kt[0] * u1) * ir, kt[1] * ir, -kt[0] * ir, 0,
                     bta * (phi - a * th), bta * (kt[0] * a - gm1 * u0), bta * (kt[1] * a - gm1 * u1), bta * gm1,
                     bta * (phi + a * th), -bta * (kt[0] * a + gm1 * u0), -bta * (kt[1] * a + gm1 * u1), bta * gm1};
    memcpy(R, Rm, sizeof(Rm));
    memcpy(L, Lm, sizeof(Lm));
    return;
  }
  double u0 = u[0], u1 = u[1], u2 = u[2];
  double th = kt[0] * u0 + kt[1] * u1 + kt[2] * u2;
  double phi = gm1 * ((u0 * u0 + u1 * u1 + u2 * u2) / 2.0);
  double ia2 = ia * ia;
  lam[0] = U; lam[1] = U; lam[2] = U; lam[3] = U + a * kmag; lam[4] = U - a * kmag;
  double Rm[25] = {
      kt[0], kt[1], kt[2], alpha, alpha,
      kt[0] * u0, kt[1] * u0 - kt[2] * rho, kt[2] * u0 + kt[1] * rho, alpha * (u0 + kt[0] * a), alpha * (u0 - kt[0] * a),
      kt[0] * u1 + kt[2] * rho, kt[1] * u1, kt[2] * u1 - kt[0] * rho, alpha * (u1 + kt[1] * a), alpha * (u1 - kt[1] * a),
      kt[0] * u2 - kt[1] * rho, kt[1] * u2 + kt[0] * rho, kt[2] * u2, alpha * (u2 + kt[2] * a), alpha * (u2 - kt[2] * a),
      kt[0] * phi / gm1 + rho * (kt[2] * u1 - kt[1] * u2), kt[1] * phi / gm1 + rho * (kt[0] * u2 - kt[2] * u0),
      kt[2] * phi / gm1 + rho * (kt[1] * u0 - kt[0] * u1), alpha * ((phi + a * a) / gm1 + th * a),
      alpha * ((phi + a * a) / gm1 - th * a)};
  double Lm[25] = {
      kt[0] * (1 - phi * ia2) - (kt[2] * u1 - kt[1] * u2) * ir, kt[0] * gm1 * u0 * ia2, kt[0] * gm1 * u1 * ia2 + kt[2] * ir,
      kt[0] * gm1 * u2 * ia2 - kt[1] * ir, -kt[0] * gm1 * ia2,
      kt[1] * (1 - phi * ia2) - (kt[0] * u2 - kt[2] * u0) * ir, kt[1] * gm1 * u0 * ia2 - kt[2] * ir, kt[1] * gm1 * u1 * ia2,
      kt[1] * gm1 * u2 * ia2 + kt[0] * ir, -kt[1] * gm1 * ia2,
      kt[2] * (1 - phi * ia2) - (kt[1] * u0 - kt[0] * u1) * ir, kt[2] * gm1 * u0 * ia2 + kt[1] * ir,
      kt[2] * gm1 * u1 * ia2 - kt[0] * ir, kt[2] * gm1 * u2 * ia2, -kt[2] * gm1 * ia2,
      bta * (phi - th * a), -bta * (gm1 * u0 - kt[0] * a), -bta * (gm1 * u1 - kt[1] * a), -bta * (gm1 * u2 - kt[2] * a),
      bta * gm1,
      bta * (phi + th * a), -bta * (gm1 * u0 + kt[0] * a), -bta * (gm1 * u1 + kt[1] * a), -bta * (gm1 * u2 + kt[2] * a),
      bta * gm1};
  memcpy(R, Rm, sizeof(Rm));
  memcpy(L, Lm, sizeof(Lm));
}

/* Roe average (averaging.py:70-114); q = [rho, rhou.., rhoE, p] */
void ref_roe_average(int nd, double g, const double* qL, const double* qR, double* out) {
  double rL = qL[0], rR = qR[0];
  double sL = sqrt(rL), sR = sqrt(rR);
  double inv = 1.0 / (sR + sL);
  out[0] = sqrt(rL * rR);
  double ke = 0.0;
  for (int m = 0; m < nd; ++m) {
    double um = (sL * (qL[1 + m] / rL) + sR * (qR[1 + m] / rR)) * inv;
    out[1 + m] = um;
    ke += um * um;
  }
  ke = ke / 2.0;
  double HL = (qL[nd + 1] + qL[nd + 2]) / rL, HR = (qR[nd + 1] + qR[nd + 2]) / rR;
  double Hh = (sL * HL + sR * HR) * inv;
  out[nd + 1] = sqrt((g - 1.0) * (Hh - ke));
}

/* Characteristic LLF reconstruction kernel of one direction                  */
/* (shock_capturing.py:357-448, weno.py:486-521, teno.py:489-530), evaluated  */
/* at i+1/2 for i in [-1, N) along d; the flux is written into wk[5d+m].      */
/* filter = 1: the WENO filter's reconstruction (WENO-JS, SimpleAverage, the
   filter flux of ref_weno_filter_flux; WENO_filter.py:324-326) */
static void reconstruction_direction(ref_state* s, int d, int filter) {
  const int nd = s->nd, nv = s->nvar;
  const ptrdiff_t st = s->st[d];
  const double g = s->c.gama;
  const int conv = s->c.conv;
  const int averaging = filter ? REF_AVG_SIMPLE : s->c.averaging;
  range3 R = rng_all(s, 0);
  R.lo[d] = -1;
  double** out = s->wk + 5 * d;
  LOOP(R) {
    ptrdiff_t x = IX(s, i, j, k);
    double avg[5], ua[3];
    if (averaging == REF_AVG_ROE) {
      double qL[7], qR[7];
      for (int m = 0; m < nv; ++m) { qL[m] = s->q[m][x]; qR[m] = s->q[m][x + st]; }
      /* the Roe average uses the velocity datasets u_m = rhou_m/rho */
      qL[nv] = s->p[x];
      qR[nv] = s->p[x + st];
      double rL = qL[0], rR = qR[0], sL = sqrt(rL), sR = sqrt(rR);
      double inv = 1.0 / (sR + sL);
      avg[0] = sqrt(rL * rR);
      double ke = 0.0;
      for (int m = 0; m < nd; ++m) {
        ua[m] = (sL * s->u[m][x] + sR * s->u[m][x + st]) * inv;
        ke += ua[m] * ua[m];
      }
      ke = ke / 2.0;
      double HL = (qL[nd + 1] + qL[nv]) / rL, HR = (qR[nd + 1] + qR[nv]) / rR;
      double Hh = (sL * HL + sR * HR) * inv;
      avg[1] = sqrt((g - 1.0) * (Hh - ke));
    } else {
      /* SimpleAverage([0,1]) of the symbols the eigensystem uses (averaging.py:38-59) */
      avg[0] = (s->q[0][x] + s->q[0][x + st]) / 2.0;
      for (int m = 0; m < nd; ++m) ua[m] = (s->u[m][x] + s->u[m][x + st]) / 2.0;
      avg[1] = (s->a[x] + s->a[x + st]) / 2.0;
    }
    double L[25], Rv[25], lam[5];
    if (s->c.curvilinear) {
      /* metric terms of the direction, simple-averaged (averaging.py:49-57,104-113) */
      double kv[3];
      for (int m = 0; m < nd; ++m) kv[m] = (s->met[d * nd + m][x] + s->met[d * nd + m][x + st]) / 2.0;
      ref_eigensystem_k(nd, g, avg[0], ua, avg[1], kv, L, Rv, lam);
    } else {
      ref_eigensystem(nd, d, g, avg[0], ua, avg[1], L, Rv, lam);
    }
    /* window p = -2..3 */
    double CF[5][6], CS[5][6];
    double lmax[5] = {0, 0, 0, 0, 0};
    for (int pp = 0; pp < 6; ++pp) {
      ptrdiff_t y = x + (pp - 2) * st;
      double Fv[5], Qv[5];
      double udp = s->u[d][y], pr = s->p[y], kmag = 1.0;
      if (s->c.curvilinear) {
        /* fluxes detJ rho U_d, detJ (rhou_m U_d + p D_dm), detJ (p + rhoE) U_d
           (apps/euler_wave_curvilinear/euler_wave.py:20-25) */
        const double dj = s->met[nd * nd][y], Ud = s->U[d][y];
        Fv[0] = dj * s->q[0][y] * Ud;
        for (int m = 0; m < nd; ++m) Fv[1 + m] = dj * (s->q[1 + m][y] * Ud + pr * s->met[d * nd + m][y]);
        Fv[nd + 1] = dj * (pr + s->q[nd + 1][y]) * Ud;
        /* the point's eigenvalues use its own metric: U = k.u, a|k| */
        double kk = 0.0;
        udp = 0.0;
        for (int m = 0; m < nd; ++m) {
          const double km = s->met[d * nd + m][y];
          udp += km * s->u[m][y];
          kk += km * km;
        }
        kmag = sqrt(kk);
      } else {
        Fv[0] = s->q[1 + d][y];
        for (int m = 0; m < nd; ++m) Fv[1 + m] = s->q[1 + m][y] * udp + (m == d ? pr : 0.0);
        Fv[nd + 1] = (pr + s->q[nd + 1][y]) * udp;
      }
      for (int m = 0; m < nv; ++m) Qv[m] = s->q[m][y];
      for (int kk = 0; kk < nv; ++kk) {
        double cf = 0.0, cs = 0.0;
        for (int m = 0; m < nv; ++m) { cf += L[kk * nv + m] * Fv[m]; cs += L[kk * nv + m] * Qv[m]; }
        CF[kk][pp] = cf;
        CS[kk][pp] = cs;
      }
      /* pointwise eigenvalues with the stencil point's u_d and a (shock_capturing.py:512-536) */
      double ap = s->c.curvilinear ? s->a[y] * kmag : s->a[y];
      double lp[5];
      if (nd == 1) { lp[0] = udp - ap; lp[1] = udp; lp[2] = udp + ap; }
      else if (nd == 2) { lp[0] = udp; lp[1] = udp; lp[2] = udp + ap; lp[3] = udp - ap; }
      else { lp[0] = udp; lp[1] = udp; lp[2] = udp; lp[3] = udp + ap; lp[4] = udp - ap; }
      for (int kk = 0; kk < nv; ++kk) lmax[kk] = fmax(lmax[kk], fabs(lp[kk]));
    }
    double recon[5];
    for (int kk = 0; kk < nv; ++kk) {
      double fp[6], fm[6];
      for (int pp = 0; pp < 6; ++pp) {
        fp[pp] = 0.5 * (CF[kk][pp] + lmax[kk] * CS[kk][pp]);
        fm[pp] = 0.5 * (CF[kk][pp] - lmax[kk] * CS[kk][pp]);
      }
      if (filter) {
        recon[kk] = ref_weno_filter_flux(fp, fm, CF[kk], NULL);
        continue;
      }
      const double ct = s->c.sensor ? adaptive_ct(s, s->theta[x]) : s->c.teno_ct;
      recon[kk] = ref_recon_side(conv, 1, fp, s->c.eps, ct) + ref_recon_side(conv, -1, fm, s->c.eps, ct);
    }
    for (int m = 0; m < nv; ++m) {
      double fl = 0.0;
      for (int kk = 0; kk < nv; ++kk) fl += Rv[m * nv + kk] * recon[kk];
      out[m][x] = fl;
    }
  }
}

/* Shock-capturing residual (shock_capturing.py:21-34, opensblifunctions.py:575-593): */
/* R_m = -sum_d (wk_dm[i] - wk_dm[i-1]) / Delta_d, assigned.                   */
static void shock_capturing(ref_state* s, int filter) {
  const int nd = s->nd, nv = s->nvar;
  for (int d = 0; d < nd; ++d) reconstruction_direction(s, d, filter);
  range3 I = rng_all(s, 0);
  LOOP(I) {
    ptrdiff_t x = IX(s, i, j, k);
    for (int m = 0; m < nv; ++m) {
      double r = 0.0;
      if (s->c.curvilinear) {
        /* -Conservative(detJ F_j, xi_j)/detJ (apps/euler_wave_curvilinear/euler_wave.py:20-25) */
        for (int d = 0; d < nd; ++d) {
          const double* w = s->wk[5 * d + m];
          r += (w[x] - w[x - s->st[d]]) / s->c.delta[d];
        }
        s->res[m][x] = -r / s->met[nd * nd][x];
        continue;
      }
      for (int d = 0; d < nd; ++d) {
        const double* w = s->wk[5 * d + m];
        /* WD(F, x_d) -> D_dd WD(F, xi_d) on a stretched direction (opensblifunctions.py:607-616) */
        if (s->dm[d]) r += met1(s, d, coord(d, i, j, k)) * ((w[x] - w[x - s->st[d]]) / s->c.delta[d]);
        else r += (w[x] - w[x - s->st[d]]) / s->c.delta[d];
      }
      s->res[m][x] = -r;
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Boundary conditions                                                       */
/* ------------------------------------------------------------------------- */
/* Periodic exchange (periodic.py:42-56): side 0 copies interior planes        */
/* [0,H) into [N,N+H); side 1 copies [N-H,N) into [-H,0); extents in the other */
/* directions include their halos so edges fill through later directions.    */
static void periodic_exchange(ref_state* s, double** arrs, int narr, int d, int side) {
  int N = s->n[d];
  range3 R;
  for (int e = 0; e < 3; ++e) {
    R.lo[e] = -s->h[e];
    R.hi[e] = s->n[e] + s->h[e];
  }
  if (side == 0) { R.lo[d] = 0; R.hi[d] = H; }
  else { R.lo[d] = N - H; R.hi[d] = N; }
  ptrdiff_t shift = (side == 0 ? N : -N) * s->st[d];
  for (int v = 0; v < narr; ++v) {
    double* f = arrs[v];
    LOOP(R) {
      ptrdiff_t x = IX(s, i, j, k);
      f[x + shift] = f[x];
    }
  }
}

/* Zeroth-order extrapolation (extrapolation.py:29-58, bc_core.py:144-197):   */
/* side 0: q[0..-nh] = q[1]; side 1: q[N-1..N-1+nh] = q[N-2]; the kernel runs   */
/* on the boundary plane with the other directions' halos included.           */
static void extrapolation0(ref_state* s, double** arrs, int narr, int d, int side) {
  int N = s->n[d];
  range3 R;
  for (int e = 0; e < 3; ++e) {
    R.lo[e] = -s->h[e];
    R.hi[e] = s->n[e] + s->h[e];
  }
  int b = side == 0 ? 0 : N - 1;
  R.lo[d] = b;
  R.hi[d] = b + 1;
  const ptrdiff_t st = s->st[d];
  const int sgn = side == 0 ? -1 : 1;      /* from_side_factor */
  const int to = side == 0 ? 1 : -1;       /* to_side_factor */
  /* n_halos is |halo| of the boundary (3 below, 4 above for WENO/TENO halos,
     weno.py:17-29); the deeper halo points are never read by any kernel, so
     filling all H of them changes no result and keeps them finite */
  const int nh = H;
  for (int v = 0; v < narr; ++v) {
    double* f = arrs[v];
    LOOP(R) {
      ptrdiff_t x = IX(s, i, j, k);
      for (int hh = 0; hh <= nh; ++hh) f[x + sgn * hh * st] = f[x + to * st];
    }
  }
}

/* Halo depth a boundary kernel fills (bc_core.py:136-147): the largest halo
   of the schemes, -3/+4 for WENO5/TENO (weno.py:17-29), -2/+2 for Central(4). */
static int bc_halos(const ref_state* s, int side) {
  const int central = s->c.conv == REF_CONV_CENTRAL_SKEW || s->c.conv == REF_CONV_CENTRAL_CONS;
  /* the WENO filter's reconstruction halos join the block's boundary halos */
  if (central && !(s->c.filters & REF_FILTER_WENO)) return 2;
  return side == 0 ? 3 : 4;
}

/* temperature from the conservative variables at x (ns_physics.py:105-106) */
static double t_cons(const ref_state* s, ptrdiff_t x) {
  const int nd = s->nd;
  const double g = s->c.gama, M = s->c.Minf;
  double m2 = 0.0;
  for (int m = 0; m < nd; ++m) m2 += s->q[1 + m][x] * s->q[1 + m][x];
  return (s->q[nd + 1][x] - 0.5 * m2 / s->q[0][x]) * (g * (g - 1.0) * M * M) / s->q[0][x];
}

/* value of the per-face plane for variable v at plane indices (ia, ib),
   clamped onto the interior extents (halo corners are never read) */
static double plane_value(const ref_state* s, int d, int side, int v, int ia, int ib) {
  const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
  const int n1 = s->n[e1], n2 = s->n[e2];
  ia = ia < 0 ? 0 : (ia >= n1 ? n1 - 1 : ia);
  ib = ib < 0 ? 0 : (ib >= n2 ? n2 - 1 : ib);
  return s->plane[d][side][((size_t)v * n2 + ib) * n1 + ia];
}

/* One point (x) of the boundary plane of (d, side): st points inward
   (to_side_factor), (ia, ib) are the plane indices of the other directions. */
static void bc_point(ref_state* s, int d, int side, ptrdiff_t x, ptrdiff_t st, int ia, int ib) {
  const int nv = s->nvar, nd = s->nd;
  const int nh = bc_halos(s, side);
  const double g = s->c.gama, M = s->c.Minf;
  double** q = s->q;
  switch (s->c.bc[d][side]) {
    case REF_BC_EXTRAP1: /* halo i = 2 q[i-1] - q[i-2] going outward (extrapolation.py:42-54) */
      for (int v = 0; v < nv; ++v)
        for (int hh = 1; hh <= nh; ++hh) q[v][x - hh * st] = 2.0 * q[v][x - (hh - 1) * st] - q[v][x - (hh - 2) * st];
      break;
    case REF_BC_SYMMETRY: /* symmetry.py:24-48: q[-i] = q[i] with the normal momentum v - 2 (v.n) n; the
                             boundary point is left as it is */
      for (int hh = 1; hh <= nh; ++hh)
        for (int v = 0; v < nv; ++v) q[v][x - hh * st] = (v == 1 + d) ? -q[v][x + hh * st] : q[v][x + hh * st];
      break;
    case REF_BC_ZERO_GRADIENT: /* q[0] = q[1]; q[-i] = q[i] (zero_gradient_outlet.py:16-23) */
      for (int v = 0; v < nv; ++v) {
        q[v][x] = q[v][x + st];
        for (int hh = 1; hh <= nh; ++hh) q[v][x - hh * st] = q[v][x + hh * st];
      }
      break;
    case REF_BC_ADIABATIC_WALL: { /* adiabatic_wall.py:28-69 */
      /* wall: zero momentum; T_wall from Carpenter's one-sided dT/dn = 0 */
      for (int m = 0; m < nd; ++m) q[1 + m][x] = 0.0;
      const double T1 = t_cons(s, x + st), T2 = t_cons(s, x + 2 * st), T3 = t_cons(s, x + 3 * st);
      const double Tw = (6.0 / 11.0) * (3.0 * T1 + (1.0 / 3.0) * T3 - (3.0 / 2.0) * T2);
      const double pw = q[0][x] * Tw / (g * M * M);
      q[nd + 1][x] = pw / (g - 1.0);
      for (int hh = 1; hh <= nh; ++hh) {
        q[0][x - hh * st] = q[0][x + hh * st];
        for (int m = 0; m < nd; ++m) q[1 + m][x - hh * st] = -q[1 + m][x + hh * st];
        q[nd + 1][x - hh * st] = q[nd + 1][x + hh * st];
      }
      break;
    }
    case REF_BC_ISOTHERMAL_WALL: { /* isothermal_wall.py:33-88 with Eq(rhoE, rho*Twall/(gama*(gama-1)*Minf**2)) */
      const double Tw = s->c.twall;
      for (int m = 0; m < nd; ++m) q[1 + m][x] = 0.0;
      q[nd + 1][x] = q[0][x] * Tw / (g * (g - 1.0) * (M * M));
      double m2 = 0.0;
      for (int m = 0; m < nd; ++m) m2 += q[1 + m][x] * q[1 + m][x];
      const double pw = (g - 1.0) * (q[nd + 1][x] - 0.5 * m2 / q[0][x]);
      const double Ta = t_cons(s, x + st);
      double uh[5][3];
      for (int hh = 1; hh <= nh; ++hh)
        for (int m = 0; m < nd; ++m) uh[hh - 1][m] = q[1 + m][x + hh * st] / q[0][x + hh * st];
      for (int hh = 1; hh <= nh; ++hh) {
        const double Th = (hh + 1) * Tw - hh * Ta;
        const double rh = pw * g * M * M / Th;
        double ke = 0.0;
        for (int m = 0; m < nd; ++m) {
          q[1 + m][x - hh * st] = -rh * uh[hh - 1][m];
          ke += uh[hh - 1][m] * uh[hh - 1][m];
        }
        q[nd + 1][x - hh * st] = pw / (g - 1.0) + 0.5 * rh * ke;
        q[0][x - hh * st] = rh;
      }
      break;
    }
    case REF_BC_DIRICHLET: /* boundary point and halos take the plane values (dirichlet.py:28-41) */
      for (int v = 0; v < nv; ++v) {
        const double val = plane_value(s, d, side, v, ia, ib);
        for (int hh = 0; hh <= nh; ++hh) q[v][x - hh * st] = val;
      }
      break;
    case REF_BC_INLET_PEXTRAP: { /* inlet_pressure_extrapolate.py:32-66, side 0 */
      const double rb = q[0][x];
      double ke = 0.0, un = 0.0;
      for (int m = 0; m < nd; ++m) {
        const double ub = fabs(q[1 + m][x] / rb);
        ke += ub * ub;
        if (m == d) un = ub;
      }
      const double pb = (g - 1.0) * (q[nd + 1][x] - 0.5 * rb * ke);
      const double ab = sqrt(g * pb / rb);
      const int super = un >= ab;
      for (int v = 0; v < nv; ++v) q[v][x] = super ? q[v][x - st] : q[v][x];
      for (int hh = 1; hh <= nh; ++hh) q[nd + 1][x - hh * st] = super ? q[nd + 1][x - hh * st] : q[nd + 1][x];
      break;
    }
    default: break;
  }
}

/* Boundary kernel of (d, side) over its plane, the other directions' halos
   included (bc_core.py:163-171). */
static void bc_face(ref_state* s, int d, int side) {
  if (s->c.bc[d][side] == REF_BC_EXTRAP0) {
    extrapolation0(s, s->q, s->nvar, d, side);
    return;
  }
  const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
  const int b = side == 0 ? 0 : s->n[d] - 1;
  const ptrdiff_t st = (side == 0 ? 1 : -1) * s->st[d];
  for (int ib = -s->h[e2]; ib < s->n[e2] + s->h[e2]; ++ib)
    for (int ia = -s->h[e1]; ia < s->n[e1] + s->h[e1]; ++ia) {
      int c[3];
      c[d] = b;
      c[e1] = ia;
      c[e2] = ib;
      bc_point(s, d, side, IX(s, c[0], c[1], c[2]), st, ia, ib);
    }
}

/* Boundary kernels in the order of SimulationBlock.apply_boundary_conditions
   (block.py:199-210): direction 0 sides 0, 1, then direction 1, ... */
static void apply_bcs(ref_state* s) {
  for (int d = 0; d < s->nd; ++d)
    for (int side = 0; side < 2; ++side) {
      if (s->c.bc[d][side] == REF_BC_PERIODIC) periodic_exchange(s, s->q, s->nvar, d, side);
      else bc_face(s, d, side);
    }
}

/* ------------------------------------------------------------------------- */
static void evaluate_rhs(ref_state* s) {
  constituent_relations(s);
  switch (s->c.conv) {
    case REF_CONV_CENTRAL_SKEW: central_skew(s); break;
    case REF_CONV_CENTRAL_CONS: central_conservative(s); break;
    default: shock_capturing(s, 0); break;
  }
  if (s->c.viscous != REF_VISC_NONE) viscous(s);
  const double* c = s->c.body_force;
  if (c[0] != 0.0 || c[1] != 0.0 || c[2] != 0.0) {
    /* - KD(_i,_j)*c_j and - Dot(c_j, u_j) (apps/channel_flow/compressible_TCF_TENO/turbulent_channel.py:15-16) */
    const int nd = s->nd;
    range3 I = rng_all(s, 0);
    LOOP(I) {
      ptrdiff_t x = IX(s, i, j, k);
      double cu = 0.0;
      for (int m = 0; m < nd; ++m) {
        s->res[1 + m][x] -= c[m];
        cu += c[m] * s->u[m][x];
      }
      s->res[nd + 1][x] -= cu;
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Filters after the full time step (InTheSimulation, algorithm.py:412-470)   */
/* ------------------------------------------------------------------------- */

/* WENO non-linear filter (WENO_filter.py:145-339): constituent relations u, p,
   a (and the Yee Mach sensor) on the post-step state; the dissipation sensor
   kappa over the interior (the sensor array's halos are never written, so they
   hold zeros); the filter residual -sum_d (F_d[i] - F_d[i-1])/Delta_d of the
   filter fluxes; then Q += Wall * shock_filter_control * kappa_max * dt * R,
   with kappa_max the maximum of kappa over the points +-2 along each direction
   and Wall = 0 within 6 points of a wall face. */
static void weno_filter(ref_state* s) {
  const int nd = s->nd, nv = s->nvar;
  const double g = s->c.gama;
  range3 R = rng_all(s, H);
  LOOP(R) {
    ptrdiff_t x = IX(s, i, j, k);
    const double inv_rho = 1.0 / s->q[0][x];
    double m2 = 0.0, uu = 0.0;
    for (int m = 0; m < nd; ++m) {
      const double um = s->q[1 + m][x] * inv_rho;
      s->u[m][x] = um;
      m2 += s->q[1 + m][x] * s->q[1 + m][x];
      uu += um * um;
    }
    const double pp = (g - 1.0) * (s->q[nd + 1][x] - 0.5 * m2 * inv_rho);
    s->p[x] = pp;
    s->a[x] = sqrt(g * pp * inv_rho);
    if (s->c.wf_mach) { /* WENO_filter.py:178-196 */
      const double inv_a = 1.0 / sqrt(g * pp * inv_rho);
      const double M = sqrt(uu) * inv_a, M2 = M * M;
      s->msens[x] = fmin(0.5 * M2 * sqrt(4.0 + (1.0 - M2) * (1.0 - M2)) / (1.0 + M2), 1.0);
    }
  }
  range3 I = rng_all(s, 0);
  if (s->c.wf_sensor == REF_WF_DUCROS) {
    LOOP(I) {
      ptrdiff_t x = IX(s, i, j, k);
      const double th = ducros_at(s, x, i, j, k);
      s->kappa[x] = s->c.wf_mach ? s->msens[x] * th : th;
    }
  }
  shock_capturing(s, 1);
  int wall[3][2] = {{0}};
  for (int d = 0; d < nd; ++d)
    for (int side = 0; side < 2; ++side)
      wall[d][side] = s->c.bc[d][side] == REF_BC_ADIABATIC_WALL || s->c.bc[d][side] == REF_BC_ISOTHERMAL_WALL;
  LOOP(I) {
    ptrdiff_t x = IX(s, i, j, k);
    const int cc[3] = {i, j, k};
    double kmax;
    if (s->c.wf_sensor == REF_WF_DUCROS) {
      kmax = s->kappa[x];
      for (int d = 0; d < nd; ++d)
        for (int o = -2; o <= 2; ++o) kmax = fmax(kmax, s->kappa[x + o * s->st[d]]);
    } else {
      kmax = s->c.wf_mach ? s->msens[x] : 1.0;
    }
    double w = 1.0;
    for (int d = 0; d < nd; ++d) {
      if (wall[d][0] && cc[d] <= 5) w = 0.0;
      if (wall[d][1] && cc[d] >= s->n[d] - 6) w = 0.0;
    }
    for (int m = 0; m < nv; ++m)
      s->q[m][x] = s->q[m][x] + w * s->c.shock_filter_control * kmax * s->c.dt * s->res[m][x];
  }
}

/* Selective frequency damping (SFD.py:64-86), pointwise over the interior:
   vfilt' = (v (1 - E) + vfilt (c + E)) / (c + 1),
   v'     = (v (E c + 1) + vfilt c (1 - E)) / (c + 1),
   c = chi/omega, E = exp(-(chi + omega) dt). */
static void sfd_filter(ref_state* s) {
  const int nv = s->nvar;
  const double chi = s->c.chi_filt, om = s->c.omega_filt;
  range3 I = rng_all(s, 0);
  for (int m = 0; m < nv; ++m) {
    double *q = s->q[m], *f = s->qf[m];
    LOOP(I) {
      ptrdiff_t x = IX(s, i, j, k);
      const double v = q[x], old = f[x];
      const double E = exp(-(chi + om) * s->c.dt);
      f[x] = (v * (1.0 - E) + old * (chi / om + E)) / (chi / om + 1.0);
      q[x] = (v * (E * chi / om + 1.0) + old * chi / om * (1.0 - E)) / (chi / om + 1.0);
    }
  }
}

/* Binomial filter of even order N (binomial_filter.py:36-110): per direction
   the weighted sum with weights binomial(N, i)/2^N on points -N/2..N/2, the
   directional values averaged, then q = (1 - sigma) q + sigma q_f where the
   grid condition holds.  Every point reads the pre-filter state (the
   reference writes q in place while reading its neighbours, which makes its
   result depend on the loop order; this is the order-free reading). */
static void binomial_filter(ref_state* s) {
  const int nd = s->nd, nv = s->nvar, N = s->c.binomial_order;
  double w[11];
  double c = 1.0;
  for (int i = 0; i <= N; ++i) {
    w[i] = c / pow(2.0, N);
    c = c * (N - i) / (i + 1);
  }
  const double sg = s->c.sigma_filt;
  range3 I = rng_all(s, 0);
  for (int m = 0; m < nv; ++m) {
    const double* q = s->q[m];
    double* o = s->tmp[m];
    LOOP(I) {
      ptrdiff_t x = IX(s, i, j, k);
      double sum = 0.0;
      for (int d = 0; d < nd; ++d) {
        double fd = 0.0;
        for (int t = 0; t <= N; ++t) fd += w[t] * q[x + (t - N / 2) * s->st[d]];
        sum += fd;
      }
      const double qf = sum / (double)nd;
      int on = 1;
      if (s->fmask) on = s->fmask[(size_t)i + (size_t)s->n[0] * ((size_t)j + (size_t)s->n[1] * (size_t)k)] != 0;
      o[x] = on ? (1.0 - sg) * q[x] + sg * qf : q[x];
    }
  }
  for (int m = 0; m < nv; ++m) {
    double *q = s->q[m], *o = s->tmp[m];
    LOOP(I) { ptrdiff_t x = IX(s, i, j, k); q[x] = o[x]; }
  }
}

static void filters(ref_state* s) {
  if (s->c.filters & REF_FILTER_WENO) weno_filter(s);
  if (s->c.filters & REF_FILTER_SFD) sfd_filter(s);
  if (s->c.filters & REF_FILTER_BINOMIAL) binomial_filter(s);
}

int ref_rk_coefficients(int rk, double* a, double* b) {
  if (rk == REF_RK3_SBLI) { /* rknew (stage) and rkold (solution), rk_sbli.py:58-60 */
    a[0] = 2.0 / 3.0; a[1] = 5.0 / 12.0; a[2] = 3.0 / 5.0;
    b[0] = 1.0 / 4.0; b[1] = 3.0 / 20.0; b[2] = 3.0 / 5.0;
    return 3;
  }
  if (rk == REF_RK3_LS) { /* rk_LS.py:91-94 */
    a[0] = 0.0; a[1] = -5.0 / 9.0; a[2] = -153.0 / 128.0;
    b[0] = 1.0 / 3.0; b[1] = 15.0 / 16.0; b[2] = 8.0 / 15.0;
    return 3;
  }
  if (rk == REF_RK3_LS_SSP) { /* rk_LS.py:74-89, evaluated in double */
    double c = 0.924574;
    double z1 = sqrt(36 * c * c * c * c + 36 * c * c * c - 135 * c * c + 84 * c - 12);
    double z2 = 2 * c * c + c - 2;
    double z3 = 12 * c * c * c * c - 18 * c * c * c + 18 * c * c - 11 * c + 2;
    double z4 = 36 * c * c * c * c - 36 * c * c * c + 13 * c * c - 8 * c + 4;
    double z5 = 69 * c * c * c - 62 * c * c + 28 * c - 8;
    double z6 = 34 * c * c * c * c - 46 * c * c * c + 34 * c * c - 13 * c + 2;
    b[0] = 0.924574;
    b[1] = (12 * c * (c - 1) * (3 * z2 - z1) - (3 * z2 - z1) * (3 * z2 - z1)) / (144 * c * (3 * c - 2) * (c - 1) * (c - 1));
    b[2] = (-24 * (3 * c - 2) * (c - 1) * (c - 1)) / ((3 * z2 - z1) * (3 * z2 - z1) - 12 * c * (c - 1) * (3 * z2 - z1));
    a[0] = 0.0;
    a[1] = (-z1 * (6 * c * c - 4 * c + 1) + 3 * z3) / ((2 * c + 1) * z1 - 3 * (c + 2) * (2 * c - 1) * (2 * c - 1));
    a[2] = (-z1 * z4 + 108 * (2 * c - 1) * c * c * c * c * c - 3 * (2 * c - 1) * z5) /
           (24 * z1 * c * (c - 1) * (c - 1) * (c - 1) * (c - 1) + 72 * c * z6 + 72 * c * c * c * c * c * c * (2 * c - 13));
    return 3;
  }
  /* rk_LS.py:96-97, five stages */
  const double A[5] = {0, -0.4178904745, -1.192151694643, -1.697784692471, -1.514183444257};
  const double B[5] = {0.1496590219993, 0.3792103129999, 0.8229550293869, 0.6994504559488, 0.1530572479681};
  for (int i = 0; i < 5; ++i) { a[i] = A[i]; b[i] = B[i]; }
  return 5;
}

static void time_step(ref_state* s) {
  const int nv = s->nvar;
  const double dt = s->c.dt;
  double A[5], B[5];
  int ns = ref_rk_coefficients(s->c.rk, A, B);
  range3 I = rng_all(s, 0);
  apply_bcs(s); /* temporal_start = bc_kernels + start kernels (algorithm.py:423-425) */
  if (s->c.rk == REF_RK3_SBLI) {
    for (int m = 0; m < nv; ++m) { /* "Save equations" Q_old = Q (rk_sbli.py:126-133) */
      double *o = s->qold[m], *q = s->q[m];
      LOOP(I) { ptrdiff_t x = IX(s, i, j, k); o[x] = q[x]; }
    }
  }
  for (int st = 0; st < ns; ++st) {
    evaluate_rhs(s);
    for (int m = 0; m < nv; ++m) {
      double *o = s->qold[m], *q = s->q[m], *r = s->res[m];
      if (s->c.rk == REF_RK3_SBLI) {
        double an = A[st], bo = B[st];
        LOOP(I) { ptrdiff_t x = IX(s, i, j, k); q[x] = o[x] + dt * an * r[x]; }   /* Sub stage advancement */
        LOOP(I) { ptrdiff_t x = IX(s, i, j, k); o[x] = o[x] + dt * bo * r[x]; }   /* Temporal solution advancement */
      } else {
        double a_ = A[st], b_ = B[st];
        LOOP(I) {
          ptrdiff_t x = IX(s, i, j, k);
          o[x] = a_ * o[x] + dt * r[x];   /* tempRK = rkA*tempRK + dt*R */
          q[x] = q[x] + b_ * o[x];        /* Q = Q + rkB*tempRK */
        }
      }
    }
    apply_bcs(s);
  }
  filters(s);
}

/* ------------------------------------------------------------------------- */
int ref_create(const ref_config* cfg, ref_state** out) {
  if (cfg->ndim < 1 || cfg->ndim > 3) return -1;
  if (cfg->sensor && cfg->ndim < 2) return -1; /* vorticity needs two directions */
  /* the filters define their fields for 2-D and 3-D blocks only */
  if (cfg->filters && cfg->ndim < 2) return -1;
  if ((cfg->filters & REF_FILTER_BINOMIAL) &&
      (cfg->binomial_order < 2 || cfg->binomial_order > 10 || cfg->binomial_order % 2)) return -1;
  if ((cfg->filters & REF_FILTER_SFD) && !(cfg->omega_filt != 0.0)) return -1;
  if (cfg->curvilinear) {
    /* periodic Euler blocks with a characteristic scheme (the curvilinear app) */
    if (cfg->ndim < 2 || cfg->viscous != REF_VISC_NONE || cfg->filters || cfg->sensor) return -1;
    if (cfg->conv == REF_CONV_CENTRAL_SKEW || cfg->conv == REF_CONV_CENTRAL_CONS) return -1;
    for (int d = 0; d < cfg->ndim; ++d)
      if (cfg->stretch[d] || cfg->bc[d][0] != REF_BC_PERIODIC || cfg->bc[d][1] != REF_BC_PERIODIC) return -1;
  }
  ref_state* s = (ref_state*)calloc(1, sizeof(ref_state));
  s->c = *cfg;
  s->nd = cfg->ndim;
  s->nvar = cfg->ndim + 2;
  for (int d = 0; d < 3; ++d) {
    s->n[d] = d < s->nd ? cfg->n[d] : 1;
    s->h[d] = d < s->nd ? H : 0;
  }
  s->sx = s->n[0] + 2 * s->h[0];
  s->sy = s->n[1] + 2 * s->h[1];
  s->sz = s->n[2] + 2 * s->h[2];
  s->ntot = s->sx * s->sy * s->sz;
  s->st[0] = 1; s->st[1] = s->sx; s->st[2] = s->sx * s->sy;
  const size_t bytes = (size_t)s->ntot * sizeof(double);
  for (int m = 0; m < s->nvar; ++m) {
    s->q[m] = (double*)calloc(1, bytes);
    s->qold[m] = (double*)calloc(1, bytes);
    s->res[m] = (double*)calloc(1, bytes);
  }
  for (int m = 0; m < s->nd; ++m) s->u[m] = (double*)calloc(1, bytes);
  s->p = (double*)calloc(1, bytes);
  s->T = (double*)calloc(1, bytes);
  s->a = (double*)calloc(1, bytes);
  s->mu = (double*)calloc(1, bytes);
  s->theta = (double*)calloc(1, bytes);
  if (cfg->filters & REF_FILTER_WENO) {
    s->kappa = (double*)calloc(1, bytes);
    s->msens = (double*)calloc(1, bytes);
  }
  for (int m = 0; m < s->nvar; ++m) {
    if (cfg->filters & REF_FILTER_SFD) s->qf[m] = (double*)calloc(1, bytes);
    if (cfg->filters & REF_FILTER_BINOMIAL) s->tmp[m] = (double*)calloc(1, bytes);
  }
  if (cfg->curvilinear) {
    for (int m = 0; m <= s->nd * s->nd; ++m) s->met[m] = (double*)calloc(1, bytes);
    for (int m = 0; m < s->nd; ++m) s->U[m] = (double*)calloc(1, bytes);
    /* until coordinates are set: the identity map x_d = i_d Delta_d */
    for (int a = 0; a < s->nd; ++a)
      for (ptrdiff_t x = 0; x < s->ntot; ++x) s->met[a * s->nd + a][x] = 1.0;
    for (ptrdiff_t x = 0; x < s->ntot; ++x) s->met[s->nd * s->nd][x] = 1.0;
  }
  carpenter_init();
  for (int d = 0; d < s->nd; ++d)
    for (int side = 0; side < 2; ++side) {
      const int bc = cfg->bc[d][side];
      /* every BC but periodic, ZeroGradientOutlet and Symmetry is a ModifyCentralDerivative */
      if (bc != REF_BC_PERIODIC && bc != REF_BC_ZERO_GRADIENT && bc != REF_BC_SYMMETRY)
        s->mod[d][side] = 1 + cfg->bscheme[d][side];
    }
  /* the residual kernels consume every convective work array before the
     viscous kernels run, so the two groups share one pool */
  int nwk = (cfg->conv == REF_CONV_CENTRAL_SKEW || cfg->conv == REF_CONV_CENTRAL_CONS) ? 32 : 15;
  if ((cfg->filters & REF_FILTER_WENO) && nwk < 15) nwk = 15;
  if (cfg->viscous != REF_VISC_NONE && nwk < 36) nwk = 36;
  for (int w = 0; w < nwk; ++w) s->wk[w] = (double*)calloc(1, bytes);
  *out = s;
  return 0;
}

void ref_destroy(ref_state* s) {
  if (!s) return;
  for (int m = 0; m < MAXV; ++m) { free(s->q[m]); free(s->qold[m]); free(s->res[m]); }
  for (int m = 0; m < 3; ++m) free(s->u[m]);
  free(s->p); free(s->T); free(s->a); free(s->mu); free(s->theta);
  free(s->kappa); free(s->msens); free(s->fmask);
  for (int m = 0; m < MAXV; ++m) { free(s->qf[m]); free(s->tmp[m]); }
  for (int w = 0; w < NWK; ++w) free(s->wk[w]);
  for (int m = 0; m < 10; ++m) free(s->met[m]);
  for (int m = 0; m < 3; ++m) free(s->U[m]);
  for (int d = 0; d < 3; ++d) {
    free(s->dm[d]); free(s->sd[d]);
    free(s->plane[d][0]); free(s->plane[d][1]);
  }
  free(s);
}

int ref_set_field(ref_state* s, int var, const double* v) {
  if (var < 0 || var >= s->nvar) return -1;
  double* f = s->q[var];
  ptrdiff_t c = 0;
  for (int k = 0; k < s->n[2]; ++k)
    for (int j = 0; j < s->n[1]; ++j)
      for (int i = 0; i < s->n[0]; ++i) f[IX(s, i, j, k)] = v[c++];
  return 0;
}

int ref_get_field(const ref_state* s, int var, double* v) {
  if (var < 0 || var >= s->nvar) return -1;
  const double* f = s->q[var];
  ptrdiff_t c = 0;
  for (int k = 0; k < s->n[2]; ++k)
    for (int j = 0; j < s->n[1]; ++j)
      for (int i = 0; i < s->n[0]; ++i) v[c++] = f[IX(s, i, j, k)];
  return 0;
}

/* Metric of a direction stretched only along itself (metric.py:72-126,241-273):
   detJ = CD(x_d, xi_d), D_dd = 1/detJ, then SD_ddd = CD(D_dd, xi_d), both with
   the faces' stencil closures; both faces must modify the central stencils
   (the one-sided rows never read the halos). */
int ref_set_coordinates(ref_state* s, int d, const double* xc) {
  if (d < 0 || d >= s->nd || !s->c.stretch[d]) return -1;
  if (!s->mod[d][0] || !s->mod[d][1]) return -2;
  const int N = s->n[d];
  const double inv = 1.0 / s->c.delta[d];
  free(s->dm[d]);
  free(s->sd[d]);
  s->dm[d] = (double*)calloc((size_t)N, sizeof(double));
  s->sd[d] = (double*)calloc((size_t)N, sizeof(double));
  for (int c = 0; c < N; ++c) s->dm[d][c] = 1.0 / (inv * d1_gen(xc, c, 1, N, s->mod[d][0], s->mod[d][1], c));
  for (int c = 0; c < N; ++c) s->sd[d][c] = inv * d1_gen(s->dm[d], c, 1, N, s->mod[d][0], s->mod[d][1], c);
  return 0;
}

/* Metric of a curvilinear block (metric.py:72-147): the Jacobian
   J_ab = CD(x_a, xi_b) of the grid coordinates by the 4th-order central first
   derivative over the interior, detJ = det J and D = adj(J)/detJ
   (generate_fd_metrics_equations, metric.py:136-147), then the periodic
   exchange of D and detJ into every halo layer (metric_boundary_condition,
   metric.py:315-341; curvilinear blocks are periodic in every direction). */
int ref_set_grid_coordinates(ref_state* s, int halo, const double* xc) {
  if (!s->c.curvilinear) return -1;
  if (halo < 2 || halo > H) return -2;
  const int nd = s->nd;
  int e[3], o[3];
  for (int d = 0; d < 3; ++d) {
    o[d] = d < nd ? halo : 0;
    e[d] = s->n[d] + 2 * o[d];
  }
  const ptrdiff_t per = (ptrdiff_t)e[0] * e[1] * e[2];
  double* x[3] = {NULL, NULL, NULL};
  for (int c = 0; c < nd; ++c) {
    x[c] = (double*)calloc((size_t)s->ntot, sizeof(double));
    for (int kk = 0; kk < e[2]; ++kk)
      for (int jj = 0; jj < e[1]; ++jj)
        for (int ii = 0; ii < e[0]; ++ii)
          x[c][IX(s, ii - o[0], jj - o[1], kk - o[2])] = xc[c * per + ii + (ptrdiff_t)e[0] * (jj + (ptrdiff_t)e[1] * kk)];
  }
  double inv[3];
  for (int d = 0; d < nd; ++d) inv[d] = 1.0 / s->c.delta[d];
  range3 I = rng_all(s, 0);
  LOOP(I) {
    ptrdiff_t p = IX(s, i, j, k);
    double J[3][3] = {{0}};
    for (int a = 0; a < nd; ++a)
      for (int b = 0; b < nd; ++b) J[a][b] = inv[b] * d1_gen(x[a], p, s->st[b], s->n[b], 0, 0, 0);
    double adj[3][3], det;
    if (nd == 2) {
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      adj[0][0] = J[1][1]; adj[0][1] = -J[0][1];
      adj[1][0] = -J[1][0]; adj[1][1] = J[0][0];
    } else {
      /* adj(J)_ab = cofactor C_ba */
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          const int r0 = (b + 1) % 3, r1 = (b + 2) % 3, c0 = (a + 1) % 3, c1 = (a + 2) % 3;
          adj[a][b] = J[r0][c0] * J[r1][c1] - J[r0][c1] * J[r1][c0];
        }
      det = J[0][0] * adj[0][0] + J[0][1] * adj[1][0] + J[0][2] * adj[2][0];
    }
    s->met[nd * nd][p] = det;
    for (int a = 0; a < nd; ++a)
      for (int b = 0; b < nd; ++b) s->met[a * nd + b][p] = adj[a][b] / det;
  }
  for (int c = 0; c < nd; ++c) free(x[c]);
  for (int d = 0; d < nd; ++d)
    for (int side = 0; side < 2; ++side) periodic_exchange(s, s->met, nd * nd + 1, d, side);
  return 0;
}

int ref_get_metric_field(const ref_state* s, int idx, double* v) {
  if (!s->c.curvilinear || idx < 0 || idx > s->nd * s->nd) return -1;
  ptrdiff_t c = 0;
  for (int k = 0; k < s->n[2]; ++k)
    for (int j = 0; j < s->n[1]; ++j)
      for (int i = 0; i < s->n[0]; ++i) v[c++] = s->met[idx][IX(s, i, j, k)];
  return 0;
}

int ref_get_sensor(const ref_state* s, double* v) {
  if (!s->c.sensor) return -1;
  ptrdiff_t c = 0;
  for (int k = 0; k < s->n[2]; ++k)
    for (int j = 0; j < s->n[1]; ++j)
      for (int i = 0; i < s->n[0]; ++i) v[c++] = s->theta[IX(s, i, j, k)];
  return 0;
}

int ref_get_metric(const ref_state* s, int d, double* dm, double* sd) {
  if (d < 0 || d >= s->nd || !s->dm[d]) return -1;
  for (int c = 0; c < s->n[d]; ++c) { dm[c] = s->dm[d][c]; sd[c] = s->sd[d][c]; }
  return 0;
}

/* The plane of a Dirichlet face, or the inflow state of an inlet face, whose
   halo points all take it now (the initial condition covers the halos,
   katzer_init.py:233-245, and the inlet kernel never writes rho, rhou there). */
int ref_set_bc_plane(ref_state* s, int d, int side, const double* plane) {
  if (d < 0 || d >= s->nd || side < 0 || side > 1) return -1;
  const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
  const size_t np = (size_t)s->nvar * s->n[e1] * s->n[e2];
  free(s->plane[d][side]);
  s->plane[d][side] = (double*)malloc(np * sizeof(double));
  memcpy(s->plane[d][side], plane, np * sizeof(double));
  if (s->c.bc[d][side] == REF_BC_INLET_PEXTRAP) {
    const int b = side == 0 ? 0 : s->n[d] - 1;
    const ptrdiff_t st = (side == 0 ? 1 : -1) * s->st[d];
    for (int ib = -s->h[e2]; ib < s->n[e2] + s->h[e2]; ++ib)
      for (int ia = -s->h[e1]; ia < s->n[e1] + s->h[e1]; ++ia) {
        int c[3];
        c[d] = b;
        c[e1] = ia;
        c[e2] = ib;
        const ptrdiff_t x = IX(s, c[0], c[1], c[2]);
        for (int v = 0; v < s->nvar; ++v)
          for (int hh = 1; hh <= H; ++hh) s->q[v][x - hh * st] = plane_value(s, d, side, v, ia, ib);
      }
  }
  return 0;
}

int ref_rhs(ref_state* s, double* residual) {
  apply_bcs(s);
  evaluate_rhs(s);
  ptrdiff_t c = 0;
  for (int m = 0; m < s->nvar; ++m)
    for (int k = 0; k < s->n[2]; ++k)
      for (int j = 0; j < s->n[1]; ++j)
        for (int i = 0; i < s->n[0]; ++i) residual[c++] = s->res[m][IX(s, i, j, k)];
  return 0;
}

int ref_run(ref_state* s, int64_t niter) {
  if ((s->c.filters & REF_FILTER_SFD) && niter > 0) {
    /* "Initialize the filter" before the time loop: vfilt = v (SFD.py:50-61) */
    range3 I = rng_all(s, 0);
    for (int m = 0; m < s->nvar; ++m) {
      if (s->sfd_set[m]) continue;
      double *f = s->qf[m], *q = s->q[m];
      LOOP(I) { ptrdiff_t x = IX(s, i, j, k); f[x] = q[x]; }
      s->sfd_set[m] = 1;
    }
  }
  for (int64_t it = 0; it < niter; ++it) time_step(s);
  return 0;
}

int ref_set_filter_mask(ref_state* s, const uint8_t* mask) {
  if (!(s->c.filters & REF_FILTER_BINOMIAL)) return -1;
  const size_t n = (size_t)s->n[0] * s->n[1] * s->n[2];
  free(s->fmask);
  s->fmask = NULL;
  if (!mask) return 0;
  s->fmask = (unsigned char*)malloc(n);
  memcpy(s->fmask, mask, n);
  return 0;
}

int ref_get_sfd_field(const ref_state* s, int var, double* v) {
  if (!(s->c.filters & REF_FILTER_SFD) || var < 0 || var >= s->nvar) return -1;
  const double* f = s->qf[var];
  ptrdiff_t c = 0;
  for (int k = 0; k < s->n[2]; ++k)
    for (int j = 0; j < s->n[1]; ++j)
      for (int i = 0; i < s->n[0]; ++i) v[c++] = f[IX(s, i, j, k)];
  return 0;
}

int ref_set_sfd_field(ref_state* s, int var, const double* v) {
  if (!(s->c.filters & REF_FILTER_SFD) || var < 0 || var >= s->nvar) return -1;
  double* f = s->qf[var];
  ptrdiff_t c = 0;
  for (int k = 0; k < s->n[2]; ++k)
    for (int j = 0; j < s->n[1]; ++j)
      for (int i = 0; i < s->n[0]; ++i) f[IX(s, i, j, k)] = v[c++];
  s->sfd_set[var] = 1;
  return 0;
}
