/*
 * gatt_dns.h -- C ABI of the MI355X (gfx950) time-advance library.
 *
 * The library replaces the Fortran module procedures of the reference time
 * loop.  Every entry point below names the reference routine it stands in for
 * (paths relative to the reference checkout).  All arrays use the Fortran
 * layout of the reference: column-major, x fastest, 1-based index boxes.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative GATT_E_* code on
 *     failure; gatt_last_error() gives a message.  Nothing is thrown across
 *     the ABI and nothing calls exit().
 *   - A context owns all device memory: the state q = (U,V,W,R,E) of `batch`
 *     members, stage work arrays, coefficient tables and the boundary program.
 *     State stays resident on the device between calls; only scalars return.
 *   - Calls on one context are ordered on the context's HIP stream and are not
 *     re-entrant.  Host pointers are copied before the call returns.  That
 *     stream is non-blocking: device pointers passed in must hold their data
 *     when the call is made (synchronise the stream that produced them first);
 *     the entry points that take device pointers return after their own work
 *     on them has completed.
 *   - 2-D problems (nz == 1) carry U,V,R,E; W is identically zero and the
 *     reported W change is 0 (RK4_2D passes W through untouched,
 *     source/Fortran/timeIntegration.F90:8-156).
 */
#ifndef GATT_DNS_H
#define GATT_DNS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GATT_OK = 0,
  GATT_E_INVALID = -1,   /* bad argument or descriptor */
  GATT_E_HIP = -2,       /* HIP runtime error */
  GATT_E_NOMEM = -3,     /* device allocation failed */
  GATT_E_UNSUPPORTED = -4
};

/* Fields of the state vector, in the reference's order U,V,W,R,E. */
enum { GATT_U = 0, GATT_V = 1, GATT_W = 2, GATT_R = 3, GATT_E = 4 };

/* Operator kinds of a line solve. */
enum { GATT_DERIV = 0, GATT_FILTER = 1 };

/* Pre-factored compact operator along one direction: the arrays the reference
 * generates in matrices.F90 (source/writeFortranMatrices.m:70-172) and passes
 * to solveSystem{X,Y,Z}[per] (source/Fortran/derivs.F90:556-948).
 *   A : (n-1) x ntypes      sub-diagonal factors (periodic: A(1) = LHS(1,n))
 *   B : (n-periodic) x nt   inverted pivots
 *   C : (n-1) x ntypes      scaled super-diagonal (periodic: C(1) = LHS(n,1))
 *   D : n x ntypes          first row of inv(LHS), periodic only (else NULL)
 *   R : n x (2nR-1) x nt    RHS band, R(i,nR+l,t) multiplies F(i+l)          */
typedef struct gatt_lineop {
  int32_t n;
  int32_t ntypes;
  int32_t nR;
  int32_t periodic;
  const double* A;
  const double* B;
  const double* C;
  const double* D;
  const double* R;
} gatt_lineop;

/* A list of boundary boxes (iUd/vUd, iUn/dUn ... of boundaryInfo.F90).
 *   box   : count x 6 (column-major), [i1 i2 j1 j2 k1 k2], 1-based inclusive
 *   value : Dirichlet values (count) or NULL
 *   dir   : Neumann / 2nd-derivative directions 1..6 = xi,xf,yi,yf,zi,zf or NULL */
typedef struct gatt_box_list {
  int32_t count;
  const int32_t* box;
  const double* value;
  const int32_t* dir;
} gatt_box_list;

/* Boundary-list variable slots: U,V,W,P,E (P acts on R through P = R E). */
enum { GATT_BC_U = 0, GATT_BC_V = 1, GATT_BC_W = 2, GATT_BC_P = 3, GATT_BC_E = 4 };

/* Disturbance hooks called at the end of applyBoundCond (runDisturbances.F90). */
enum {
  GATT_DIST_LID_DRIVEN = 1,    /* source/disturbances/lidDrivenMovement.F90: par[0] = U0 (on U) */
  GATT_DIST_HOLD_INLET = 2,    /* source/Fortran/holdInlet.F90: par[0] = holdDensity flag */
  GATT_DIST_PERIODIC_SINE = 3, /* source/disturbances/periodicSine.F90: par = {omega, amplitude}, on var */
  GATT_DIST_PACKET_2D = 4      /* source/disturbances/packet_2d.F90: par = {omega, nmodes, amplitude}, on var */
};

typedef struct gatt_disturbance {
  int32_t kind;
  int32_t box[6];
  const double* x;      /* X(box(1):box(2)) as written in runDisturbances.F90 */
  double par[4];
  int32_t var;          /* field the routine receives as u (GATT_U..GATT_E); kinds 3 and 4 */
} gatt_disturbance;

/* Forcing hooks called at the end of nseq (runForcings{2,3}D.F90). */
enum {
  GATT_FORCING_TURBULENCE = 1,        /* source/disturbances/turbulenceGenerator.F90: par = {epsilon, e0} */
  GATT_FORCING_PRESSURE_GRADIENT = 2  /* source/disturbances/pressureGradient.F90: par[0] = dpdx;
                                         var = {field passed as r, field passed as u} ('RU') */
};

typedef struct gatt_forcing {
  int32_t kind;
  int32_t box[6];
  double par[4];
  int32_t var[2];
} gatt_forcing;

/* Everything applyBoundCond receives (source/Fortran/nseq.F90:327-350). */
typedef struct gatt_bc_desc {
  gatt_box_list dir[5];     /* Dirichlet, slots U,V,W,P,E */
  gatt_box_list neu[5];     /* Neumann (zero gradient), slots U,V,W,P,E */
  gatt_box_list sec[5];     /* null second derivative, slots U,V,W,P,E */
  int32_t nCorners;
  const int32_t* cL;        /* nCorners x 6 */
  const int32_t* cD;        /* nCorners x 3, entries -1/0/1 */
  const int32_t* cAdiabatic;/* nCorners */
  int32_t neumannLength;
  int32_t neumann2Length;
  const double* neumannCoeffs;
  const double* neumann2Coeffs;
  int32_t nDisturb;
  const gatt_disturbance* disturb;
} gatt_bc_desc;

/* One problem = the content of parameters.F90, matrices.F90, boundaryInfo.F90,
 * runDisturbances.F90, runForcings*.F90 and SFD.mat of the reference. */
typedef struct gatt_problem_desc {
  int32_t nx, ny, nz;
  double Re, Ma, Pr, gamma, T0;
  gatt_lineop deriv[3];     /* x, y, z (z ignored when nz == 1) */
  gatt_lineop filter[3];
  const int32_t* types[3];  /* typeMapX (ny*nz), typeMapY (nx*nz), typeMapZ (nx*ny), 1-based */
  int32_t filterDir[3];     /* filterX/Y/Z */
  double filterCharTime;    /* FilterCharTime; <= 0 means full replacement */
  gatt_bc_desc bc;
  int32_t nForcing;
  const gatt_forcing* forcing;
  int32_t sfdType;          /* 0 off, 1 scalar chi, 2 chi field */
  double sfdDelta;          /* SFD_Delta; <= 0 freezes the mean (Delta = inf) */
  double sfdXval;           /* SFD_X_val (type 1) */
  const double* sfdX;       /* chi field nx*ny*nz (type 2) or NULL */
  double dxmin[3];
  int32_t timeStepping;     /* timeStepping of parameters.F90: 1 RK4 (also 0), 2 Euler, 3 SSPRK3
                               (writeFortranParameters.m:50-58, main.F90:255-350) */
} gatt_problem_desc;

typedef struct gatt_ctx gatt_ctx;

/* Last error message of the calling thread ("" if none). */
const char* gatt_last_error(void);

/* Library and device information: writes a short description into buf. */
int gatt_device_info(char* buf, int32_t len);

/* Create a context for `batch` independent members of one problem
 * (replaces main.F90:75-80 set-up: decomp_2d_init + include matrices/boundaryInfo). */
int gatt_create(const gatt_problem_desc* desc, int32_t batch, gatt_ctx** out);
int gatt_destroy(gatt_ctx* ctx);

/* Host <-> device state copies of member m (importFlow.F90 / exportFlow.F90
 * without the MAT file).  q is an array of 5 pointers U,V,W,R,E to nx*ny*nz
 * doubles each (W may be NULL in 2-D). */
int gatt_set_state(gatt_ctx* ctx, int32_t member, const double* const* q);
int gatt_get_state(gatt_ctx* ctx, int32_t member, double* const* q);
/* SFD mean of member m (main.F90:148-164; defaults to the state at the first step). */
/* The five values of the current state at npts points (tracked points of the
 * log file, main.F90:478-486): out[5*i + f] = field f (U,V,W,R,E) of `member` at
 * the 0-based linear point idx[i] = i + nx*(j + ny*k); W = 0 in 2-D.  At most
 * 2^20 points per call; every index must lie in the grid. */
int gatt_get_points(gatt_ctx* ctx, int32_t member, int32_t npts, const int64_t* idx, double* out);
int gatt_set_mean(gatt_ctx* ctx, int32_t member, const double* const* q);
int gatt_get_mean(gatt_ctx* ctx, int32_t member, double* const* q);

/* One RK4 step q -> q_new with boundary conditions after every stage
 * (RK4_3D / RK4_2D, source/Fortran/timeIntegration.F90:338-498, 8-156).
 * The result replaces the state. */
int gatt_rk4(gatt_ctx* ctx, double t, double dt);

/* One explicit Euler step (EULER_3D / EULER_2D, timeIntegration.F90:500-566,
 * 158-222) and one three-stage SSP Runge-Kutta step (SSPRK3_3D / SSPRK3_2D,
 * timeIntegration.F90:569-685, 225-335), boundary conditions after every stage.
 * The result replaces the state.  gatt_step and gatt_advance use the integrator
 * named by the descriptor's timeStepping. */
int gatt_euler(gatt_ctx* ctx, double t, double dt);
int gatt_ssprk3(gatt_ctx* ctx, double t, double dt);

/* Compact filter in x, y, z with blend factor theta (dt/FilterCharTime;
 * theta <= 0 replaces the field) -- applyFilters3D / applyFilters2D,
 * source/Fortran/derivs.F90:399-552, 303-397.  Acts on the state in place. */
int gatt_apply_filters(gatt_ctx* ctx, double theta);

/* applyBoundCond on the state at time t (source/Fortran/nseq.F90:327-518). */
int gatt_apply_bound_cond(gatt_ctx* ctx, double t);

/* SFD blend on the state (main.F90:381-402). */
int gatt_sfd(gatt_ctx* ctx, double dt);

/* Right-hand side of the Navier-Stokes equations, nseq3d / nseq2d
 * (source/Fortran/nseq.F90:142-323, 10-139), of member 0's state.
 * dq: 5 host pointers receiving dU,dV,dW,dR,dE. */
int gatt_rhs(gatt_ctx* ctx, double* const* dq);

/* One full time step of the main loop (main.F90:196-462) for every member:
 * CFL scalars on the current state, RK4, filter, SFD, boundary conditions,
 * max changes, then q = q_new.  Outputs (for member 0, may be NULL):
 *   umax[3]      max |U|,|V|,|W| before the step (main.F90:197-205)
 *   maxchange[5] max |q_new - q| per field (main.F90:414-418)
 *   diverged     1 if any change is NaN or all are zero (main.F90:434) */
int gatt_step(gatt_ctx* ctx, double t, double dt, double* umax, double* maxchange, int32_t* diverged);

/* nsteps fixed-dt steps starting at time t0 (timeControl = 1), replayed from
 * a captured HIP graph; per-step scalars of the last step are returned.  The
 * time advances on the device as t = t + dt after each step (main.F90:424), so
 * time-dependent disturbances see the same times as the reference loop.  The
 * state is used as given, as after main.F90:112; only a problem that holds its
 * inflow (holdInlet) and has had no boundary application yet gets one first,
 * which captures the inflow. */
int gatt_advance(gatt_ctx* ctx, int32_t nsteps, double t0, double dt, double* maxchange, int32_t* diverged);

/* The state before the last step (as gatt_get_state): after a gatt_step whose
 * flow diverged it is the U that main.F90:433-441 writes to flow_diverged
 * (exportFlow runs before U = Unew). */
int gatt_get_previous_state(gatt_ctx* ctx, int32_t member, double* const* q);

/* Device copy of the state (and, with SFD, the mean) of every member inside the
 * context: op 0 saves, op 1 restores the saved copy (allocated on the first
 * save).  No reference counterpart: the time loop saves before a replayed
 * stretch, and when gatt_advance reports a divergence it restores and repeats
 * the stretch one gatt_step at a time, so it stops at the diverged step as
 * main.F90:433-448 does. */
int gatt_checkpoint(gatt_ctx* ctx, int32_t op);

/* Unit entry point: one operator along one direction on nfields device
 * arrays of nx*ny*nz doubles (calcDerivs{X,Y,Z} / one direction of
 * applyFilters: solveSystem{X,Y,Z}[per], source/Fortran/derivs.F90:556-948).
 * in/out are arrays of nfields device pointers; out may not alias in. */
int gatt_line_solve(gatt_ctx* ctx, int32_t dir, int32_t kind, int32_t nfields,
                    const double* const* in, double* const* out);

/* Synchronise the context's stream. */
int gatt_sync(gatt_ctx* ctx);

/* Launch shape of the line solves along `dir` for operator `kind`.  No
 * reference counterpart (launch tuning); chosen by a built-in heuristic at
 * create time and overridden by the autotuner (gatt_dns_amd/autotune.py).
 *   variant 0, lane groups:   param = G lanes of a wavefront per line (power of
 *                             two, 2..64), waves_per_block in 1..4;
 *   variant 1, wave segments: param = M points per lane (16 or 32); each
 *                             lane owns a line, ceil(n/M) wavefronts per
 *                             workgroup (waves_per_block ignored); only for
 *                             directions whose tiles of 64 consecutive lines
 *                             each have one operator type;
 *   variant 2, fused wave segments: as variant 1, with per-point coefficient
 *                             records folded into single fused multiply-adds
 *                             and no branches in the sweeps (results agree
 *                             with variants 0 and 1 to rounding);
 *   variant 3 (x only): the same sweeps with the lines of a workgroup
 *                             staged through an LDS tile: 32 lines
 *                             (waves_per_block 1) or 16 (waves_per_block 2);
 *   variant 4 (x only), one wavefront per line: param = M points per lane
 *                             (4, 8 or 16; 64*M >= n); any operator type
 *                             per line.
 * Shapes that do not fit the line or the device's LDS are rejected. */
int gatt_set_line_shape(gatt_ctx* ctx, int32_t dir, int32_t kind, int32_t variant, int32_t param,
                        int32_t waves_per_block);
int gatt_get_line_shape(gatt_ctx* ctx, int32_t dir, int32_t kind, int32_t* variant, int32_t* G, int32_t* M,
                        int32_t* waves_per_block);

/* Kernel classes of one time step, for timing.  DERIV: first derivatives of the
 * state fields along one direction; FILTER: the filter in place; VISC and
 * ASSEMBLE: the stresses and the assembly of the unfused stage; FLUX_X / FLUX_Z:
 * derivatives of the viscous fluxes along x / z with the fluxes computed as they
 * are loaded; RHS_Y: the y flux derivatives fused with the assembly and the
 * stage update (the last two kinds only where the context's step uses them).
 * STEP: one whole main-loop step, as gatt_advance replays it (a graph of two
 * steps launched on the context's stream; gatt_advance must have run at least
 * two steps first, and the timing advances the state). */
enum {
  GATT_K_DERIV_X = 0, GATT_K_DERIV_Y = 1, GATT_K_DERIV_Z = 2,
  GATT_K_FILTER_X = 3, GATT_K_FILTER_Y = 4, GATT_K_FILTER_Z = 5,
  GATT_K_VISC = 6, GATT_K_ASSEMBLE = 7,
  GATT_K_FLUX_X = 8, GATT_K_FLUX_Z = 9, GATT_K_RHS_Y = 10,
  GATT_K_STEP = 11
};

/* Time one kernel class on the context's stream with HIP events: one launch of
 * that class is replayed `reps` times.  Returns the mean duration of one launch
 * in ms and the algorithmic bytes of that launch (8 B per point for every
 * distinct array read or written once).  For STEP the duration is that of one
 * step and the bytes are those of the minimum-pass schedule of an RK4 step
 * with the filter and the step scalars: 8 B x (377 + s) per point and member in
 * 3-D, 8 B x (196 + s) in 2-D, s the SFD arrays (3-D / 2-D: 11 / 9 for type 2
 * with an evolving mean, 10 / 8 for type 1, 6 / 5 for a frozen mean, 0 off).
 * Filter timing overwrites the state.  No reference counterpart (measurement
 * only). */
int gatt_time_kernel(gatt_ctx* ctx, int32_t kernel, int32_t reps, double* avg_ms, double* bytes);

/* ------------------------------------------------------------------------
 * Jacobian-free stability analysis on the members of a context
 * (calcInstability.m).  The reference runs one DNS process per Krylov vector;
 * here a block of vectors becomes a block of members, advanced together by
 * gatt_advance.  Krylov vectors are device arrays in the flow2vec layout
 * (calcInstability.m:480-515): the flow-region points of U,V,R,E (2-D) or
 * U,V,W,R,E (3-D), one field after the other, region points in increasing
 * column-major order.  V, W, X, Y below are device pointers to column-major
 * blocks of such vectors; H and M are host arrays.
 * ------------------------------------------------------------------------ */

/* Flow region (calcInstability.m:276-281, flowRegion = ~isnan(baseflow.U)
 * minus optional buffer zones): `count` strictly increasing 0-based
 * column-major point indices.  idx == NULL selects every point (the default). */
int gatt_set_flow_region(gatt_ctx* ctx, int64_t count, const int64_t* idx);

/* singleBeta packing (calcInstability.m:14, 493-509, 557-579) on a 3-D grid
 * (on = 1) or the full 3-D packing (on = 0, the default).  A singleBeta vector
 * holds U,V,W,R,E at the flow-region points of the z = 0 plane: perturbations
 * are v cos(z_k) on plane k (v sin(z_k) for W), z_k = 2 pi k / nz, and
 * collected states are projected as sum_k (2/nz) cos(z_k) q_k (sin for W).
 * Switching resets the flow region; gatt_set_flow_region then takes indices
 * of the z = 0 plane (0 <= idx < nx*ny). */
int gatt_set_single_beta(gatt_ctx* ctx, int32_t on);

/* Krylov vector length N (calcInstability.m:284-290): 4 or 5 times the
 * number of flow-region points. */
int gatt_vector_length(gatt_ctx* ctx, int64_t* n);

/* Base flow q0 (calcInstability.m:313-330, baseflow.U ... baseflow.E), host
 * arrays as gatt_set_state.  It is also the frozen SFD mean of perturbed
 * members (buffer-zone damping with Delta = inf). */
int gatt_set_base(gatt_ctx* ctx, const double* const* q);

/* Members first..first+count-1 become base + scale * vec2flow(V[:, b])
 * (calcInstability.m:345-350 with scale = epsilon*sqrt(N); vec2flow
 * 517-582).  V == NULL sets the members to the base flow.  Points outside the
 * flow region take the base value.  With SFD on, each member's mean is reset
 * to the base flow. */
int gatt_perturb(gatt_ctx* ctx, int32_t first, int32_t count, const double* V, int64_t ldv, double scale);

/* W[:, b] = scale * (flow2vec(member first+b) - S_b) for b < count, with
 * S_b = flow2vec(member minus_first+b) if minus_first >= 0 (second-order
 * differences, calcInstability.m:370-385), else the vector `ref` (device,
 * first order: UM = flow2vec(Phi(q0)), 385-388), else 0. */
int gatt_collect(gatt_ctx* ctx, int32_t first, int32_t count, int32_t minus_first, const double* ref, double scale,
                 double* W, int64_t ldw);

/* H (k x p, column-major, host) = X^T Y for X: n x k and Y: n x p device
 * blocks -- the projections H(j,k) = zeta(:,j)'*B of the Gram-Schmidt loop
 * (calcInstability.m:389-393), a whole block at once.  Sums are formed in a
 * fixed order, so results are reproducible. */
int gatt_block_dot(gatt_ctx* ctx, int64_t n, int32_t k, int32_t p, const double* X, int64_t ldx, const double* Y,
                   int64_t ldy, double* H);

/* gatt_block_dot without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same arguments, kernels
 * and summation order as gatt_block_dot, so H is bitwise the same; the partial
 * sums live in a fixed 8 MiB array of library device memory per device, set up
 * at gatt_create, so k*p*(chunks+1) doubles must fit in 2^20, with
 * chunks = min(512, ceil(n/8192)) (32 x 32 blocks fit at any n).  Larger blocks
 * return GATT_E_INVALID.  Every call on one device shares that array: calls
 * that overlap on one device (from several host threads or contexts) race. */
int gatt_block_dot_noalloc(gatt_ctx* ctx, int64_t n, int32_t k, int32_t p, const double* X, int64_t ldx,
                           const double* Y, int64_t ldy, double* H);

/* Y = alpha*Y + beta*X*M for X: n x k, M: k x p (host), Y: n x p -- the
 * update uPrime = uPrime - H(j,k)*zeta(:,j) (calcInstability.m:391-392), the
 * normalisation zeta(:,k+1) = uPrime/H(k+1,k) (395-397) and the modes
 * psiA = zeta*psiH (437-440), a whole block at once.  X and Y must not
 * overlap; alpha == 0 does not read Y. */
int gatt_block_update(gatt_ctx* ctx, int64_t n, int32_t k, int32_t p, const double* X, int64_t ldx, const double* M,
                      double alpha, double beta, double* Y, int64_t ldy);

/* ------------------------------------------------------------------------
 * Tangent-linear and adjoint steps (no reference counterpart: the
 * differentiation of the step of main.F90:255-420 as implemented here).
 * Each member carries a perturbation q' (5 fields, same layout as the state).
 *   gatt_tlm_step : q <- Phi(q) and q' <- J(q) q'   (J = dPhi/dq at the old q)
 *   gatt_adj_step : q' <- J(q)^T q'                 (q is not advanced)
 * Where the SFD mean evolves (SFD on, Delta > 0) the mean m is part of the
 * state: Phi maps (q, m) to (q, m) after the step, the member also carries the
 * mean's perturbation m' (slot GATT_SLOT_MEAN_PERTURBATION), both steps act on
 * the pair (q', m'), and gatt_adj_step leaves the base mean as given.
 * The adjoint is the exact transpose of the tangent: RK4 stages, the
 * Navier-Stokes RHS (nseq.F90), the compact solves as transposed Thomas
 * sweeps with the transposed periodic correction and the transposed RHS band,
 * the filter blend, the SFD damping with a frozen mean and the boundary
 * program (Dirichlet, Neumann and null-second-derivative boxes, pressure
 * conditions through R = P/E, corners, lid and held inflow) and the
 * turbulenceGenerator forcing (its global sums differentiated: the energy sum's
 * and the means' tangents are global sums of the perturbation against the base
 * state, their adjoints global sums of the RHS adjoint).
 * ------------------------------------------------------------------------ */
enum { GATT_SLOT_STATE = 0, GATT_SLOT_PERTURBATION = 1, GATT_SLOT_MEAN = 2, GATT_SLOT_MEAN_PERTURBATION = 3 };

/* Host copies of member m's perturbation, as gatt_set_state / gatt_get_state. */
int gatt_set_perturbation(gatt_ctx* ctx, int32_t member, const double* const* q);
int gatt_get_perturbation(gatt_ctx* ctx, int32_t member, double* const* q);

/* Device-to-device copy of member m's state, perturbation, SFD mean or the
 * mean's perturbation (slot) to or from a caller buffer of 5*nx*ny*nz doubles
 * (U,V,W,R,E one after the other): to_ctx != 0 copies dev into the context.
 * For checkpoints kept on the GPU.  The mean slot needs SFD on; before it is
 * set or first used, the mean reads as the state. */
int gatt_copy_fields(gatt_ctx* ctx, int32_t member, int32_t slot, int32_t to_ctx, double* dev);

/* Unit entry point: the transpose of gatt_line_solve along `dir` (the
 * adjoint of solveSystem{X,Y,Z}[per], source/Fortran/derivs.F90:556-948) on
 * nfields device arrays.  The in arrays are used as workspace and are
 * overwritten; out may not alias in. */
int gatt_line_solve_t(gatt_ctx* ctx, int32_t dir, int32_t kind, int32_t nfields, double* const* in,
                      double* const* out);

/* One step of the state with its tangent (all members); scalars as gatt_step. */
int gatt_tlm_step(gatt_ctx* ctx, double t, double dt, double* maxchange, int32_t* diverged);

/* Adjoint of one step about the current state (all members). */
int gatt_adj_step(gatt_ctx* ctx, double t, double dt);

/* ------------------------------------------------------------------------
 * Ragged groups: independent problems of different grid sizes and
 * parameters (a parameter sweep, runBatch.m / one DNS process per case in the
 * reference) advanced together.  Each member is its own context.
 * ------------------------------------------------------------------------ */
typedef struct gatt_group gatt_group;

/* Group n distinct contexts of one device.  mode 0: one captured graph per
 * double step in which every member's step is a parallel branch; mode 1:
 * per-member graphs launched round-robin on the members' own streams; mode 2:
 * one captured graph of ragged launches, each kernel launch of the step
 * serving every member (a per-member table of argument blocks and the first
 * block of each member, walked inside the kernels).  Mode 2 needs members whose
 * steps make the same launches (same dimensionality, scheme bands, integrator,
 * SFD and forcing settings; boundary programs may differ); it gives the
 * members one common launch shape per direction and operator for the life of
 * the group and restores their own on gatt_group_destroy, and
 * gatt_group_advance returns GATT_E_UNSUPPORTED when the steps differ.  The
 * contexts stay owned by the caller and must outlive the group. */
int gatt_group_create(gatt_ctx* const* ctxs, int32_t n, int32_t mode, gatt_group** out);
int gatt_group_destroy(gatt_group* group);

/* nsteps fixed-dt steps of every member (member m from time t0[m] with step
 * dt[m]), each exactly as gatt_advance on that context alone.  maxchange
 * (n x 5) and diverged (n) receive each member's last-step scalars (may be
 * NULL). */
int gatt_group_advance(gatt_group* group, int32_t nsteps, const double* t0, const double* dt, double* maxchange,
                       int32_t* diverged);

#ifdef __cplusplus
}
#endif

#endif /* GATT_DNS_H */
