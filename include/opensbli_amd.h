/*
 * opensbli_amd.h -- C-ABI of libosbl_hip.so, the MI355X (gfx950) implementation
 * of the right-hand-side and Runge-Kutta time-advance kernels that OpenSBLI
 * otherwise emits as OPS C.
 *
 * The reference has no FFI for this path: its generated program calls the OPS
 * runtime directly.  Each entry point below names the OPS calls / generated
 * constructs it replaces (paths relative to the OpenSBLI source tree).
 *
 * Conventions
 *  - Plain C types only; no ownership crosses the boundary.  Host buffers are
 *    borrowed for the duration of a call.  Device buffers passed to the *_dev
 *    entry points must live on the context's device.
 *  - Every int-returning call returns 0 on success and a negative OSBL_E_* code
 *    on failure; osbl_last_error() then returns a thread-local message.
 *  - A context is used by one host thread at a time.  All device work of a
 *    context is ordered on its own HIP stream.
 *  - Field layout across the boundary: interior points only, x fastest,
 *    n[0]*n[1]*n[2] doubles per conservative variable, variable order
 *    rho, rhou0..rhou(ndim-1), rhoE (the time_advance_arrays order,
 *    opensbliequations.py:426-438).
 */
#ifndef OPENSBLI_AMD_H
#define OPENSBLI_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OSBL_ABI_VERSION 1

/* error codes */
#define OSBL_OK 0
#define OSBL_E_ARG (-1)      /* invalid argument / unsupported configuration */
#define OSBL_E_HIP (-2)      /* HIP runtime failure */
#define OSBL_E_NOMEM (-3)    /* device allocation failed */
#define OSBL_E_STATE (-4)    /* call not valid in the context's state */

/* convective discretisation: the spatial scheme objects of the problem script */
#define OSBL_CONV_CENTRAL_SKEW 0 /* Central(4) on Skew(...) terms (parsing.py:75-111, scheme.py:187-252) */
#define OSBL_CONV_CENTRAL_CONS 1 /* Central(4) on Conservative(...) fluxes (parsing.py:189-216) */
#define OSBL_CONV_WENO_JS 2      /* LLFWeno(5, formulation='JS') (weno.py:341-369,468-521) */
#define OSBL_CONV_WENO_Z 3       /* LLFWeno(5, formulation='Z') (weno.py:283-338) */
#define OSBL_CONV_TENO5 4        /* LLFTeno(5) (teno.py:195-213,468-530) */
#define OSBL_CONV_TENO6 5        /* LLFTeno(6) (teno.py:216-234) */

#define OSBL_AVG_SIMPLE 0 /* SimpleAverage([0,1]) (averaging.py:31-59) */
#define OSBL_AVG_ROE 1    /* RoeAverage([0,1])    (averaging.py:62-114) */

#define OSBL_VISC_NONE 0
#define OSBL_VISC_CONST 1 /* tau_ij=(1/Re)(...), q_j = D_j T/((gama-1)Minf^2 Pr Re) (taylor_green_vortex.py:18-19) */
#define OSBL_VISC_SUTHERLAND 2 /* mu = T^1.5 (1+SuthT/RefT)/(T+SuthT/RefT) as a CR (katzer_SBLI.py:12-26) */

#define OSBL_RK3 0        /* RungeKutta(3), 3-register (rk_sbli.py:16-146) */
#define OSBL_RK3_LS 1     /* RungeKuttaLS(3) (rk_LS.py:91-94) */
#define OSBL_RK3_LS_SSP 2 /* RungeKuttaLS(3, formulation='SSP') (rk_LS.py:74-89) */
#define OSBL_RK4_LS 3     /* RungeKuttaLS(4), five stages (rk_LS.py:96-97) */

#define OSBL_BC_PERIODIC 0        /* PeriodicBC (periodic.py:7-56, exchange.py:9-57) */
#define OSBL_BC_EXTRAP0 1         /* ExtrapolationBC(order=0) (extrapolation.py:37-41) */
#define OSBL_BC_EXTRAP1 2         /* ExtrapolationBC(order=1) (extrapolation.py:42-54) */
#define OSBL_BC_ZERO_GRADIENT 3   /* ZeroGradientOutletBC (zero_gradient_outlet.py:11-23) */
#define OSBL_BC_ADIABATIC_WALL 4  /* AdiabaticWallBC (adiabatic_wall.py:28-69) */
#define OSBL_BC_ISOTHERMAL_WALL 5 /* IsothermalWallBC with rhoE = rho Twall/(gama(gama-1)Minf^2) (isothermal_wall.py:33-88) */
#define OSBL_BC_DIRICHLET 6       /* DirichletBC, values from osbl_set_bc_plane (dirichlet.py:28-41) */
#define OSBL_BC_INLET_PEXTRAP 7   /* InletPressureExtrapolateBC, side 0 (inlet_pressure_extrapolate.py:32-66) */
#define OSBL_BC_SYMMETRY 8        /* SymmetryBC: mirrored halos with the normal momentum negated; the central
                                     stencils are kept (symmetry.py:24-48) */

/* stencil closure at faces whose BC is a ModifyCentralDerivative (all but
   periodic and zero-gradient; bc_core.py:15-16, opensblifunctions.py:523-534) */
#define OSBL_BS_CARPENTER 0      /* Carpenter() (Carpenter_scheme.py:6-102), the BCs' default */
#define OSBL_BS_REDUCED_ACCESS 1 /* ReducedAccess() (reduced_access_scheme.py:6-83) */

/* filters run after every full time step, in this order (UserDefinedEquations
   placed InTheSimulation, algorithm.py:412-470); 2-D and 3-D blocks */
#define OSBL_FILTER_WENO 1     /* WENOFilter(block, 5, ...) (WENO_filter.py:15-339) */
#define OSBL_FILTER_SFD 2      /* SFD(block, chifilt, omegafilt) (SFD.py:27-86) */
#define OSBL_FILTER_BINOMIAL 4 /* BinomialFilter(block, order, grid_condition, sigma) (binomial_filter.py:15-123) */

/* dissipation sensor of the WENO filter (WENO_filter.py:265-275) */
#define OSBL_WF_DUCROS 0   /* modified Ducros sensor, maximum over +-2 points per direction */
#define OSBL_WF_CONSTANT 1 /* no flow sensor: shock_filter_control alone */

/*
 * The problem definition, a flat mirror of the constants OpenSBLI declares
 * (ConstantsToDeclare, kernel.py:20-35; substitute_simulation_parameters,
 * helperfunctions.py:130-149) and of the scheme/BC objects on the block.
 */
typedef struct osbl_config {
  int32_t ndim;        /* 1, 2 or 3 */
  int32_t n[3];        /* block0np0..2; 1 for absent dimensions */
  double delta[3];     /* Delta0block0..Delta2block0 */
  int32_t conv;        /* OSBL_CONV_* */
  int32_t averaging;   /* OSBL_AVG_* (characteristic schemes only) */
  int32_t viscous;     /* OSBL_VISC_* */
  int32_t rk;          /* OSBL_RK* */
  int32_t bc[3][2];    /* OSBL_BC_* per direction and side */
  double gama, Minf, Re, Pr, dt, eps, teno_ct;
  int32_t bscheme[3][2]; /* OSBL_BS_* per face (faces that modify the central stencils) */
  int32_t stretch[3];    /* 1: direction d is stretched along itself; coordinates via osbl_set_coordinates
                            (MetricsEquation.generate_transformations, metric.py:37-54) */
  int32_t filters;       /* OSBL_FILTER_* bitmask */
  double suthT, refT;    /* SuthT, RefT of the Sutherland law */
  double twall;          /* Twall of the isothermal wall */
  int32_t sensor;        /* 1: modified Ducros sensor with the adaptive TENO cut-off
                            (shock_sensors.py:12-49, teno.py:430-443, LLFTeno(formulation='adaptive')) */
  int32_t curvilinear;   /* 1: general curvilinear metric D_ij = d xi_i/d x_j and detJ from the grid
                            coordinates (osbl_set_grid_coordinates): the strong conservation form
                            -Conservative(detJ F_j, xi_j)/detJ with the characteristic system about the
                            metric normal (MetricsEquation with curvilinear directions, metric.py:72-147;
                            euler_eigensystem.py:18-40; apps/euler_wave_curvilinear/euler_wave.py);
                            periodic Euler blocks (2-D, 3-D) with a WENO/TENO scheme */
  double teno_a1, teno_a2, sensor_eps;  /* teno_a1, teno_a2, epsilon of the sensor */
  int32_t binomial_order;  /* even order of the binomial filter, 2..8 (binomial_filter.py:16-24) */
  int32_t wf_sensor;       /* OSBL_WF_*: dissipation_sensor of the WENO filter */
  int32_t wf_mach;         /* 1: Mach_correction of the WENO filter (Yee sensor, WENO_filter.py:178-196) */
  int32_t reserved3;
  double sigma_filt;       /* sigma_filt: blending weight of the binomial filter */
  double chi_filt, omega_filt;  /* chi_filt, omega_filt of SFD */
  double shock_filter_control;  /* shock_filter_control: global scale of the WENO filter */
  double body_force[3];    /* c_j of a constant body force: -KD(_i,_j)*c_j in the momentum and
                              -Dot(c_j, u_j) in the energy equation (apps/channel_flow); blocks with
                              closures, a metric or Sutherland viscosity (the general update kernels) */
} osbl_config;

typedef struct osbl_ctx osbl_ctx;

/* ABI version of the loaded library (OSBL_ABI_VERSION). */
int osbl_abi_version(void);

/* Thread-local message describing the last failure. */
const char* osbl_last_error(void);

/*
 * Create a context: device, stream and every dataset of the block.
 * Replaces ops_init / ops_decl_block / ops_decl_dat / ops_decl_const /
 * ops_partition of the generated main() (opsc.py:540-553, 617-662, 693-722).
 */
int osbl_create(const osbl_config* cfg, osbl_ctx** out);

/* Release every device resource of the context (replaces ops_exit, opsc.py:435-443). */
void osbl_destroy(osbl_ctx* ctx);

/*
 * Upload / download one conservative variable (interior points, x fastest).
 * set_field replaces the initial-condition kernel of GridBasedInitialisation
 * (gridbasedinit.py:46-57) or ops_decl_dat_hdf5 (opsc.py:702-705);
 * get_field replaces ops_fetch_dat_hdf5_file (io_hdf5.py:99-127).
 */
int osbl_set_field(osbl_ctx* ctx, int var, const double* host);
int osbl_get_field(osbl_ctx* ctx, int var, double* host);

/*
 * One conservative variable WITH the halo layers the boundary conditions
 * fill, after applying them to the current state: what the dataset of an
 * OPS dat holds when ops_fetch_dat_hdf5_file writes it at the end of a stage
 * (io_hdf5.py:121-123; the boundary kernels close every stage,
 * algorithm.py:440-442).  d_m[d] <= 0 and d_p[d] >= 0 receive the halo depths
 * per direction (every layer of a periodic direction, the schemes' boundary
 * halos on other faces; 0 beyond ndim); host receives
 * prod(n[d] - d_m[d] + d_p[d]) doubles, x fastest.  host NULL only reports
 * the depths.
 */
int osbl_get_field_halo(osbl_ctx* ctx, int var, int32_t* d_m, int32_t* d_p, double* host);

/* Same, with a device pointer (e.g. a torch tensor's data_ptr()); enqueued on the context's stream. */
int osbl_set_field_dev(osbl_ctx* ctx, int var, const double* dev);
int osbl_get_field_dev(osbl_ctx* ctx, int var, double* dev);

/*
 * Run niter iterations of the time loop: boundary conditions, save of the
 * RK register, then for every stage the constituent relations, the spatial
 * scheme kernels, the RK update and the boundary conditions
 * (the Timers/DoLoop tree of algorithm.py:423-477).  Stream-ordered; returns
 * once the work is enqueued.
 */
int osbl_run(osbl_ctx* ctx, int64_t niter);

/*
 * Apply the boundary conditions to the current state and evaluate one RHS;
 * copies the residual (nvar blocks of interior points) to host memory.
 * Corresponds to one pass of the spatial kernels of a stage (Residual arrays,
 * opensbliequations.py:238-246).
 */
int osbl_rhs(osbl_ctx* ctx, double* residual_host);

/* Replay the time loop from hipGraphs of two iterations (default on; the
   launches are issued directly while profiling is enabled). */
int osbl_use_graphs(osbl_ctx* ctx, int enable);

/* Coordinates x_d[0..n[d]) of a stretched direction.  Replaces the metric
   kernels of MetricsEquation (metric.py:241-273): D_dd = 1/CD(x_d, xi_d) and
   SD_ddd = CD(D_dd, xi_d) with the faces' stencil closures.  Both faces of d
   must modify the central stencils. */
int osbl_set_coordinates(osbl_ctx* ctx, int d, const double* x_host);

/* Grid coordinates of a curvilinear block: ndim arrays x_c (c = 0..ndim-1),
   each over the interior extended by `halo` >= 2 points on both sides of every
   direction, x fastest ((n[0]+2 halo)(n[1]+2 halo)(n[2]+2 halo) doubles, 1 for
   absent directions).  Replaces the metric kernels of MetricsEquation
   (metric.py:72-147, 241-273): J_ab = CD(x_a, xi_b) with the 4th-order central
   derivative, detJ = det J and D = adj(J)/detJ over the interior, then the
   periodic exchange of D and detJ (metric.py:315-341).  The coordinate arrays
   are what the reference's initial condition writes into x0..x2 over the
   interior and its halos (gridbasedinit.py:46-57). */
int osbl_set_grid_coordinates(osbl_ctx* ctx, int halo, const double* coords);

/* Interior values (x fastest) of a metric array: on a curvilinear block
   idx = a*ndim + b for D_ab and ndim*ndim for detJ; on a block with stretched
   directions idx = d*ndim + d for D_dd of a stretched direction d (a function
   of its own index).  The datasets D00.., detJ the reference writes with
   iohdf5 (apps/euler_wave_curvilinear/euler_wave.py:135,
   apps/katzer_SBLI/katzer_SBLI.py:135). */
int osbl_get_metric_field(osbl_ctx* ctx, int idx, double* host);

/* Interior values (x fastest) of the Ducros shock sensor theta of the last
   stage evaluated (shock_sensors.py:12-49), on blocks with the adaptive TENO
   or the WENO filter's Ducros sensor.  The adaptive cut-off the reference
   stores as the TENO dataset (teno.py:430-443, store_sensor) is
   10^-floor(teno_a1 - teno_a2 (1 - (1 - theta)^4 (1 + 4 theta))) of it. */
int osbl_get_sensor(osbl_ctx* ctx, double* host);

/* Number of NaN values among the interior values of conservative variable
   `var` of the current state, or of every variable when var < 0: the check
   OPS's ops_NaNcheck(dat) makes for a simulation monitor
   (core/diagnostics/simulation_monitors.py:112-113) or for
   print_iteration_ops(NaN_check=...) (utilities/helperfunctions.py:172-190).
   The caller decides what a non-zero count stops. */
int osbl_nan_count(osbl_ctx* ctx, int var, int64_t* count);

/* Value of conservative variable `var` at interior index `index` (ndim
   entries, 0-based): the one-point reduction of a simulation monitor's probe
   (core/diagnostics/simulation_monitors.py:73-106). */
int osbl_get_point(osbl_ctx* ctx, int var, const int32_t* index, double* value);

/* Per-face plane of conservative values, [nvar][n_e2][n_e1] over the interior
   extents of the two other directions (e1 < e2): the values a DirichletBC face
   sets (dirichlet.py:28-41), or the inflow state an InletPressureExtrapolateBC
   face keeps in its halos (written now; the reference's initial-condition
   kernel covers the halos, katzer_init.py:233-245). */
int osbl_set_bc_plane(osbl_ctx* ctx, int d, int side, const double* plane_host);

/* Grid condition of the binomial filter: one byte per interior cell (x
   fastest), nonzero where the filter applies -- the boolean expression the
   reference evaluates from the coordinates (binomial_filter.py:67-75),
   evaluated by the caller.  NULL applies it everywhere (the default). */
int osbl_set_filter_mask(osbl_ctx* ctx, const uint8_t* mask_host);

/* SFD filtered solution vfilt of one conservative variable (interior points,
   x fastest): get for a restart file, set to restart from one ("write out
   vfilt for restart with sfd", SFD.py:14-16).  Variables not set take the
   state at the first osbl_run (the "Initialize the filter" kernel,
   SFD.py:50-61). */
int osbl_get_sfd_field(osbl_ctx* ctx, int var, double* host);
int osbl_set_sfd_field(osbl_ctx* ctx, int var, const double* host);

/* Wait for all work enqueued on the context's stream. */
int osbl_sync(osbl_ctx* ctx);

/* The context's hipStream_t, as an opaque pointer (for event timing by the caller). */
void* osbl_stream(osbl_ctx* ctx);

/*
 * Per-kernel timing: when enabled, every launch is bracketed with HIP events
 * on the context's stream and accumulated per kernel family.
 * osbl_kernel_stats fills name (NUL-terminated, up to name_len bytes),
 * total milliseconds and launch count of family `index`; returns the number
 * of families, or a negative error.  Replaces OPS's per-kernel timing output
 * (ops_timing_output, opsc.py:435-443).
 */
int osbl_profile(osbl_ctx* ctx, int enable);
int osbl_kernel_stats(osbl_ctx* ctx, int index, char* name, int name_len, double* total_ms, int64_t* launches);

/*
 * Algorithmic work of one RK stage of this configuration: HBM bytes
 * 8*(4*nvar) per interior cell (read Q and the RK register, write both) and
 * FP64 operations per interior cell counted from the kernel formulas.
 */
int osbl_work_per_stage(osbl_ctx* ctx, double* bytes, double* flops);

/* Algorithmic HBM bytes and FP64 operations of one launch of kernel family
   `index` (the families of osbl_kernel_stats). */
int osbl_kernel_work(osbl_ctx* ctx, int index, double* bytes, double* flops);

/* Number of RK stages of the configured temporal scheme. */
int osbl_num_stages(osbl_ctx* ctx);

/*
 * Discrete adjoint of the central skew-symmetric scheme and of the LLF
 * characteristic WENO5-Z scheme, each with or without constant-viscosity
 * terms, on periodic blocks without metric or filters (no counterpart in the
 * reference, which generates forward code only; TENO's cut-off and
 * Sutherland's law are excluded).  All arrays are device pointers to packed interior fields
 * (nvar blocks of n[0]*n[1]*n[2] doubles), enqueued on the context's stream.
 *  osbl_rhs_vjp_dev:      out += (dR/dQ at q)^T mu
 *  osbl_step_stages_dev:  advance one RK step from the current state and store
 *                         the input state of every stage in stages
 *                         (nstages x nvar blocks)
 *  osbl_adjoint_step_dev: lambda <- (dQ^{n+1}/dQ^n)^T lambda for one
 *                         RungeKutta(3) step whose stage inputs are in stages
 */
int osbl_rhs_vjp_dev(osbl_ctx* ctx, const double* q, const double* mu, double* out);
int osbl_step_stages_dev(osbl_ctx* ctx, double* stages);
int osbl_adjoint_step_dev(osbl_ctx* ctx, const double* stages, double* lambda);

/*
 * osbl_adjoint_step_dev_noalloc: osbl_adjoint_step_dev without any device
 * allocation (no counterpart in the reference).  For callers that cannot
 * allocate: code under HIP graph capture, tight backward loops, harnesses
 * that forbid hipMalloc.  Its work buffers (the three stage adjoints, the
 * halo-padded state and weights, the viscous VJP's fields) come from a
 * fixed device array of OSBL_NOALLOC_SCRATCH_MIB MiB that each device running
 * the library holds until the process ends; a block needing more returns
 * OSBL_E_ARG.  Every call on one device shares that array (its contents are
 * the previous call's; what must start at zero is cleared by a kernel on the
 * context's stream): calls that overlap on one device race.  Same arguments,
 * kernels and results as osbl_adjoint_step_dev.
 */
#define OSBL_NOALLOC_SCRATCH_MIB 128
int osbl_adjoint_step_dev_noalloc(osbl_ctx* ctx, const double* stages, double* lambda);

/*
 * Ensembles of independent 1-D problems of uneven size (characteristic
 * WENO/TENO fluxes, zeroth-order extrapolation ends, RungeKutta(3) or
 * RungeKuttaLS(3)).  No counterpart in the reference, which would build and
 * run one generated executable per instance (one OPSC program each,
 * opsc.py:250-282); these entry points replace that loop of executables.
 * Instance i has n[i] points, spacing delta[i], its own constant dt[i]
 * (rk_LS.py:62-67 allows only a constant dt) and nsteps[i] iterations.
 * base supplies the scheme, averaging, RK scheme, gama, eps and teno_ct.
 */
typedef struct osbl_batch osbl_batch;
int osbl_batch_create(const osbl_config* base, int ninst, const int32_t* n, const double* delta, const double* dt,
                      const int64_t* nsteps, osbl_batch** out);
void osbl_batch_destroy(osbl_batch* b);
/* state of one instance: 3 blocks of n[inst] doubles (rho, rhou0, rhoE) */
int osbl_batch_set_state(osbl_batch* b, int inst, const double* host);
int osbl_batch_get_state(osbl_batch* b, int inst, double* host);
/* integrate every instance to its nsteps: mode 0 = one persistent launch,
   mode 1 = one launch per RK stage over all running instances, replayed from
   hipGraphs; returns the device time in *elapsed_ms */
int osbl_batch_run(osbl_batch* b, int mode, double* elapsed_ms);
/* persistent mode: threads per workgroup (64..512, multiple of 64) for instances with n <= max_n[k] */
int osbl_batch_set_threads(osbl_batch* b, int nbuckets, const int32_t* max_n, const int32_t* threads);
/* staged mode: iterations captured per hipGraph */
int osbl_batch_set_graph_steps(osbl_batch* b, int steps);
/* snapshot / restore the ensemble state on the device (for repeated timing) */
int osbl_batch_save(osbl_batch* b);
int osbl_batch_restore(osbl_batch* b);

#ifdef __cplusplus
}
#endif
#endif /* OPENSBLI_AMD_H */
