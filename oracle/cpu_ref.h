/*
 * cpu_ref.h -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * CPU restatement of the kernel sequence that OpenSBLI's OPSC target
 * (opensbli/code_generation/opsc.py) emits for the RHS and Runge-Kutta time
 * advance.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load this library, and only as the checker.  Nothing in the product
 * (opensbli_amd/) links, loads or calls it.
 */
#ifndef OSBL_CPU_REF_H
#define OSBL_CPU_REF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Convective discretisations (the scheme objects the problem script installs). */
enum {
  REF_CONV_CENTRAL_SKEW = 0,  /* Central(4) on the Skew/Blaisdell form, taylor_green_vortex.py:13-15 */
  REF_CONV_CENTRAL_CONS = 1,  /* Central(4) on Conservative(...) fluxes, parsing.py:189-216 */
  REF_CONV_WENO_JS = 2,       /* LLFWeno(5, formulation='JS'), weno.py:341-369 */
  REF_CONV_WENO_Z = 3,        /* LLFWeno(5, formulation='Z'),  weno.py:283-338 */
  REF_CONV_TENO5 = 4,         /* LLFTeno(5), teno.py:195-213 */
  REF_CONV_TENO6 = 5          /* LLFTeno(6), teno.py:216-234 */
};
enum { REF_AVG_SIMPLE = 0, REF_AVG_ROE = 1 };           /* averaging.py:31,62 */
enum { REF_VISC_NONE = 0, REF_VISC_CONST = 1,          /* taylor_green_vortex.py:18-19 */
       REF_VISC_SUTHERLAND = 2 };                       /* mu(T) CR, katzer_SBLI.py:26 */
enum { REF_RK3_SBLI = 0, REF_RK3_LS = 1, REF_RK3_LS_SSP = 2, REF_RK4_LS = 3 };  /* rk_sbli.py, rk_LS.py */
enum {
  REF_BC_PERIODIC = 0,        /* periodic.py */
  REF_BC_EXTRAP0 = 1,         /* ExtrapolationBC(order=0), extrapolation.py:37-41 */
  REF_BC_EXTRAP1 = 2,         /* ExtrapolationBC(order=1), extrapolation.py:42-54 */
  REF_BC_ZERO_GRADIENT = 3,   /* ZeroGradientOutletBC, zero_gradient_outlet.py:11-23 */
  REF_BC_ADIABATIC_WALL = 4,  /* AdiabaticWallBC, adiabatic_wall.py:28-69 */
  REF_BC_ISOTHERMAL_WALL = 5, /* IsothermalWallBC, isothermal_wall.py:33-88 */
  REF_BC_DIRICHLET = 6,       /* DirichletBC, dirichlet.py:28-41 */
  REF_BC_INLET_PEXTRAP = 7,   /* InletPressureExtrapolateBC (side 0), inlet_pressure_extrapolate.py:32-66 */
  REF_BC_SYMMETRY = 8         /* SymmetryBC, symmetry.py:24-48 */
};
/* near-boundary replacement of the central stencils (bc_core.py:15-16,
   opensblifunctions.py:523-534) at faces whose BC modifies them */
enum { REF_BS_CARPENTER = 0, REF_BS_REDUCED_ACCESS = 1 };
/* filters applied after every full time step (InTheSimulation kernels,
   algorithm.py:412-470), in this order: WENO filter, SFD, binomial filter */
enum { REF_FILTER_WENO = 1,       /* WENOFilter, WENO_filter.py:15-339 */
       REF_FILTER_SFD = 2,        /* SFD, SFD.py:27-86 */
       REF_FILTER_BINOMIAL = 4 }; /* BinomialFilter, binomial_filter.py:15-123 */
enum { REF_WF_DUCROS = 0, REF_WF_CONSTANT = 1 };  /* dissipation_sensor, WENO_filter.py:265-275 */

typedef struct ref_config {
  int32_t ndim;
  int32_t n[3];          /* block0np0..2 (1 for absent dimensions) */
  double delta[3];       /* Delta{d}block0 */
  int32_t conv, averaging, viscous, rk;
  int32_t bc[3][2];
  double gama, Minf, Re, Pr, dt, eps, teno_ct;
  int32_t bscheme[3][2]; /* REF_BS_* of each face (used when its BC modifies the stencils) */
  int32_t stretch[3];    /* 1: the coordinate of direction d is given (metric.py:37-54) */
  int32_t filters;       /* REF_FILTER_* bitmask */
  double suthT, refT;    /* Sutherland constants SuthT, RefT */
  double twall;          /* Twall of the isothermal wall */
  int32_t sensor;        /* 1: Ducros sensor theta with the adaptive TENO cut-off */
  int32_t curvilinear;   /* 1: general metric D_ij = d xi_i / d x_j and detJ from grid coordinates
                            (MetricsEquation with curvilinear directions, metric.py:72-147) */
  double teno_a1, teno_a2, sensor_eps;  /* teno_a1, teno_a2, epsilon of the sensor */
  int32_t binomial_order;  /* even order N of the binomial filter, 2..10 */
  int32_t wf_sensor;       /* REF_WF_*: dissipation sensor of the WENO filter */
  int32_t wf_mach;         /* 1: Yee low-Mach correction of the WENO filter */
  int32_t pad3_;
  double sigma_filt;       /* binomial blending weight sigma_filt */
  double chi_filt, omega_filt;   /* SFD damping coefficient and cut frequency */
  double shock_filter_control;   /* global scale of the WENO filter */
  double body_force[3];          /* c_j: -c_i in the momentum, -c_j u_j in the energy (apps/channel_flow) */
} ref_config;

typedef struct ref_state ref_state;

int ref_create(const ref_config* cfg, ref_state** out);
void ref_destroy(ref_state* s);
/* interior values, x fastest, n0*n1*n2 doubles */
int ref_set_field(ref_state* s, int var, const double* interior);
int ref_get_field(const ref_state* s, int var, double* interior);
/* coordinates x_d (n[d] values) of a stretched direction; computes the
   metric D_dd = 1/CD(x_d, xi_d) and SD_ddd = CD(D_dd, xi_d) */
int ref_set_coordinates(ref_state* s, int d, const double* x);
/* grid coordinates of a curvilinear block: ndim arrays x_c (c = 0..ndim-1),
   each over the interior extended by `halo` (2..5) points on both sides of
   every active direction, x fastest; computes J_ab = CD(x_a, xi_b), detJ =
   det J and D = adj(J)/detJ over the interior, then the periodic exchange of
   D and detJ (metric.py:72-147, 315-341) */
int ref_set_grid_coordinates(ref_state* s, int halo, const double* coords);
/* interior values of a curvilinear metric array: idx = a*ndim + b for D_ab,
   ndim*ndim for detJ */
int ref_get_metric_field(const ref_state* s, int idx, double* interior);
/* per-face plane of conservative values, [nvar][n_e2][n_e1] over the interior
   extents of the other two directions (e1 < e2): the Dirichlet values, or the
   inflow state that an inlet face keeps in its halos */
int ref_set_bc_plane(ref_state* s, int d, int side, const double* plane);
/* apply the boundary conditions, then evaluate one RHS into residual[nvar][cells] */
int ref_rhs(ref_state* s, double* residual);
/* niter full iterations of the generated time loop (algorithm.py:466-470) */
int ref_run(ref_state* s, int64_t niter);

/* ---- per-function entry points used by the pinning tests ---- */
/* f points at the 6-point window f[-2..3] (array of 6); side +1 or -1 */
double ref_recon_side(int conv, int side, const double* f6, double eps, double teno_ct);
/* smoothness indicators (WENO: 3, TENO5: 3, TENO6: 4) written to beta */
int ref_smoothness(int conv, int side, const double* f6, double* beta);
/* eigensystem at an averaged state (rho, u[ndim], a) for direction d:
   L (nvar*nvar row-major), R (nvar*nvar), lambda (nvar) */
void ref_eigensystem(int ndim, int d, double gama, double rho, const double* u, double a,
                     double* L, double* R, double* lam);
/* the same about a metric direction k (ndim >= 2, euler_eigensystem.py:18-40,
   88-135): unit normal k/|k| in L and R, lambda = (U.., U + a|k|, U - a|k|)
   with U = k.u */
void ref_eigensystem_k(int ndim, double gama, double rho, const double* u, double a, const double* k,
                       double* L, double* R, double* lam);
/* Roe average between two states; out = [rho, u0..u(ndim-1), a] */
void ref_roe_average(int ndim, double gama, const double* qL, const double* qR, double* out);
/* central weights on points -2..2 of the first and second derivative */
void ref_central_weights(double* w1, double* w2);
/* boundary closure weights of the first (order 1: rows x 6) or second
   (order 2: 2 x 6) derivative at side 0 for scheme REF_BS_*; returns rows */
int ref_closure_weights(int scheme, int order, double* w);
/* the Ducros sensor of the last RHS evaluation over the interior (x fastest) */
int ref_get_sensor(const ref_state* s, double* theta);
/* metric arrays of direction d after ref_set_coordinates (n[d] values each) */
int ref_get_metric(const ref_state* s, int d, double* dm, double* sd);
/* binomial filter grid condition: 1 where the filter applies (interior cells,
   x fastest); without a mask it applies everywhere */
int ref_set_filter_mask(ref_state* s, const uint8_t* mask);
/* SFD filtered solution vfilt (interior, x fastest); it starts as the state at
   the first ref_run unless set */
int ref_get_sfd_field(const ref_state* s, int var, double* interior);
int ref_set_sfd_field(ref_state* s, int var, const double* interior);
/* WENO filter flux of one characteristic field at an interface: fp, fm the
   split fluxes f+ and f- and cf the characteristic flux on points -2..3; returns
   rj (recon(f+) + recon(f-) - central6(cf)) and writes rj */
double ref_weno_filter_flux(const double* fp6, const double* fm6, const double* cf6, double* rj);
/* RK coefficients of a temporal scheme: returns the number of stages */
int ref_rk_coefficients(int rk, double* a, double* b);

#ifdef __cplusplus
}
#endif
#endif
