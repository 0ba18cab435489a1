/*
 * cs3d.h -- C ABI of the MI355X compact-RHS library (lib3dcompact_hip.so).
 *
 * Drop-in boundary for the per-stage right-hand-side pipeline of
 * emathew1/3DCompactSolver.  Each entry point names the reference interface
 * it replaces (<file>:<line>).  All functions return CS3D_OK (0) or an error
 * code; cs3d_last_error() gives the message of the calling thread's last
 * failure.  Plain pointers and sizes only.
 *
 * Layout: fields are Nx*Ny*Nz arrays in the reference's XYZ order,
 * index = k*Nx*Ny + j*Nx + i (Macros.hpp:8), unless a call says otherwise.
 * Threading: one solver handle per thread at a time (one HIP stream each);
 * calls on one handle are not re-entrant.
 */
#ifndef CS3D_H
#define CS3D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CS3D_OK = 0,
    CS3D_ERR_ARG = 1,          /* invalid argument or unsupported size */
    CS3D_ERR_HIP = 2,          /* HIP runtime failure (allocation, launch) */
    CS3D_ERR_UNSUPPORTED = 3,  /* configuration outside the GPU path */
    CS3D_ERR_STATE = 4,        /* call out of order (e.g. step before ICs) */
    CS3D_ERR_NONFINITE = 5     /* NaN/Inf met in a reduction (dt, diagnostics) */
};

/* BC.hpp:10 BCType */
enum { CS3D_PERIODIC_SOLVE = 0, CS3D_DIRICHLET_SOLVE = 1 };
/* BC.hpp:11 BCKind.  Walls (CSolver_AWS.cpp:553-826, 1506-1605): no-slip
 * or moving, adiabatic (the wall temperature makes the one-sided 7-point
 * normal derivative vanish), at least 7 points along the wall's direction.
 * INLET is declared by the reference but handled nowhere in it; cs3d_create
 * refuses it. */
enum {
    CS3D_BC_PERIODIC = 0,
    CS3D_BC_SPONGE = 1,
    CS3D_BC_ADIABATIC_WALL = 2,
    CS3D_BC_MOVING_ADIABATIC_WALL = 3,
    CS3D_BC_INLET = 4
};
/* Derivatives.hpp:48 Direct */
enum { CS3D_DIRX = 0, CS3D_DIRY = 1, CS3D_DIRZ = 2 };
/* operator kinds; the _T kinds are the transposed operators F^T, D1^T,
 * D2^T (the adjoint's; no reference counterpart) */
enum {
    CS3D_FILTER = 0, CS3D_DERIV1 = 1, CS3D_DERIV2 = 2,
    CS3D_FILTER_T = 3, CS3D_DERIV1_T = 4, CS3D_DERIV2_T = 5
};
/* arithmetic of the bulk path */
enum { CS3D_FP64 = 0, CS3D_FP32 = 1 };
/* time integrator (RK4.hpp, TVDRK3.hpp) */
enum { CS3D_TVDRK3 = 3, CS3D_RK4 = 4 };
/* pointer kinds */
enum { CS3D_HOST = 0, CS3D_DEVICE = 1 };
/* field layouts for the standalone operators */
enum { CS3D_LAYOUT_XYZ = 0, CS3D_LAYOUT_REFERENCE = 1 };

/*
 * Solver configuration: the values the reference takes from Domain
 * (Domain.hpp:16), BC (BC.hpp:18), TimeStepping (TimeStepping.hpp:21),
 * the CSolver_AWS constructor (CSolver_AWS.hpp:119) and IdealGas
 * (IdealGas.hpp:35), plus the GPU path's own choices.
 */
typedef struct {
    int Nx, Ny, Nz;
    double Lx, Ly, Lz;        /* x_i = i*L/(N-1) (Domain.hpp:30-45) */
    int bc_type[3];           /* CS3D_PERIODIC_SOLVE / CS3D_DIRICHLET_SOLVE per direction */
    int bc_kind[6];           /* X0, X1, Y0, Y1, Z0, Z1 */
    double cfl;               /* TimeStepping::CFL */
    int filter_step;          /* TimeStepping::filterStep */
    int max_steps;            /* TimeStepping::maxTimeStep */
    double max_time;          /* TimeStepping::maxTime */
    double alpha_f;           /* filter parameter */
    double mu_ref;            /* IdealGas::mu_ref */
    double gamma;             /* IdealGas::gamma (1.4) */
    double Pr;                /* IdealGas::Pr (reference default 0.7) */
    int rk;                   /* CS3D_RK4 or CS3D_TVDRK3 */
    int precision;            /* CS3D_FP64 or CS3D_FP32 */
    int compensated;          /* FP32 only: state and RK sums carried as FP32 hi + lo pairs */
    double sponge_frac[3];    /* sponge length / domain length (SpongeBC.hpp:57-59) */
    double sponge_strength;   /* 6.0 */
    double sponge_avgT;       /* 15.0 */
    double sponge_epsP;       /* 0.005 */
    double dt_fixed;          /* > 0: constant dt instead of the CFL dt */
    /* moving-wall velocity (U, V, W) per face X0, X1, Y0, Y1, Z0, Z1, used
     * on CS3D_BC_MOVING_ADIABATIC_WALL faces: the tangential components are
     * CSolver_AWS's X0WallV, X0WallW, ..., Z1WallV (CSolver_AWS.hpp:105-107,
     * default 0); the normal component is ignored (always 0 on a wall) */
    double wall_vel[6][3];
} cs3d_config;

typedef struct cs3d_solver cs3d_solver;
typedef struct cs3d_ops cs3d_ops;

/* message of this thread's last failed call ("" if none) */
const char *cs3d_last_error(void);
/* library version string */
const char *cs3d_version(void);

/* Reference defaults (3DCompact.cpp:45-102 / IdealGas.hpp:39-40) for a grid. */
int cs3d_config_default(cs3d_config *cfg, int Nx, int Ny, int Nz);

/* new CSolver_AWS(dom, bc, ts, alphaF, mu_ref, blocksize, useTiming)
 * (CSolver_AWS.hpp:119-185) + initializeSolverData (CSolver_AWS.cpp:3-196):
 * allocates the device state on the current HIP device. */
int cs3d_create(const cs3d_config *cfg, cs3d_solver **out);
void cs3d_destroy(cs3d_solver *s);

/* setInitialConditions (CSolver_AWS.cpp:199-504): host arrays rho0, U0, V0,
 * W0, p0 in XYZ order. */
int cs3d_set_initial_conditions(cs3d_solver *s, const double *rho0, const double *U0,
                                const double *V0, const double *W0, const double *p0);

/* preStep -> calcDtFromCFL (CSolver_AWS.cpp:507-551, 2897-2904) */
int cs3d_pre_step(cs3d_solver *s);
/* one RK stage: preSubStep + solveEqnSet + postSubStep (CSolver_AWS.cpp:2906-2921),
 * AbstractRK::updateConservedData (RK4.hpp:35-118, TVDRK3.hpp:38-96),
 * updateData (CSolver_AWS.cpp:2923-2928). */
int cs3d_sub_step(cs3d_solver *s, int rkStep, int rkLast);
/* postStep -> updateSponge (CSolver_AWS.cpp:1937-2031, 2930-2946) */
int cs3d_post_step(cs3d_solver *s);
/* nsteps of the executeSolverLoop body (RK4.hpp:120-170, TVDRK3.hpp:98-139) */
int cs3d_step(cs3d_solver *s, int nsteps);
/* the stage right-hand side alone, dt*RHS into the *k2 set (solveEqnSet) */
int cs3d_rhs(cs3d_solver *s, int rkStep);

/* Host copies of a conserved set: set 1 = *1 (state), 2 = *k, 3 = *k2, 4 = *2;
 * with sponge faces also 5 = the sponge running averages (SpongeBC.hpp:
 * spongeRhoAvg ...) and 6 = the copy of the sponge planes taken before the
 * last sponge reset.  Any output pointer may be NULL. */
int cs3d_get_conserved(cs3d_solver *s, int set, double *rho, double *rhoU, double *rhoV,
                       double *rhoW, double *rhoE);
int cs3d_set_conserved(cs3d_solver *s, int set, const double *rho, const double *rhoU,
                       const double *rhoV, const double *rhoW, const double *rhoE);
/* U, V, W, p, T, mu, Amu, sos of the state (updateNonConservedData,
 * CSolver_AWS.cpp:1895-1935); any pointer may be NULL */
int cs3d_get_primitives(cs3d_solver *s, double *U, double *V, double *W, double *p,
                        double *T, double *mu, double *Amu, double *sos);
int cs3d_get_time(cs3d_solver *s, double *time, double *dt, int *timeStep);

/* The run state beyond the conserved sets, for restarts (no reference
 * counterpart: the reference dumps the state, CSolver_AWS.cpp:2493-2535,
 * but has no reader).  rho_ref, rhoE_ref: the reference state of the
 * perturbation storage; cfl_max: the CFL maximum the next dt comes from
 * (taken before the sponge update, so not recomputable from the dumped
 * state); stale: the next first stage takes sponge-plane primitives from
 * set 6.  Restart order: cs3d_set_initial_conditions (any field),
 * cs3d_set_run_state, cs3d_set_conserved for set 1 (and 5, 6 with
 * sponges), cs3d_set_run_state again (setting set 1 recomputes cfl_max). */
typedef struct {
    int timeStep, filterTimeStep;
    double time, dt, cfl_max;
    double rho_ref, rhoE_ref;
    int stale;
} cs3d_run_state;
int cs3d_get_run_state(cs3d_solver *s, cs3d_run_state *rs);
int cs3d_set_run_state(cs3d_solver *s, const cs3d_run_state *rs);

/* When enabled, the last stage of every step keeps the per-point inputs of
 * the step diagnostics (the reference computes them in postStep,
 * CSolver_AWS.cpp:2930-2934) and the post-step reduction forms them. */
int cs3d_set_diagnostics(cs3d_solver *s, int enabled);

/* calcTaylorGreenQuantities (CSolver_AWS.cpp:2175-2257), FP64 reductions:
 * out = time, kineticEngSum, turbDissSum, enstrophySum, enstrophySum2Mu,
 * meanMu, dilprime.  Valid after a completed step. */
int cs3d_diag_taylor_green(cs3d_solver *s, double out[7]);
/* calcTurbulenceQuantities (CSolver_AWS.cpp:2033-2173): out = time, tau,
 * taylorReyn3, turbMach, meanTurbDiss, uprime, kolNu, rhoprime, dilprime,
 * vortprime, meanKineticEng */
int cs3d_diag_turbulence(cs3d_solver *s, double out[11]);
/* shearLayerInfoCalc (CSolver_AWS.cpp:2259-2456) of the current state, FP64:
 * out = 11 rows of Ny values, each row the x-z plane statistic per y index,
 * in the order of its files: rhoprime2, uprime2, vprime2, wprime2, pprime2,
 * tprime2, uvprime, uwprime, vwprime, uAvg, rhoAvg.  fix_q12 = 0 keeps the
 * reference's accumulation of v'w' into uwprime (vwprime zero); 1 separates
 * them. */
int cs3d_diag_shear_layer(cs3d_solver *s, int fix_q12, double *out);
/* No counterpart in the reference: cs3d_diag_shear_layer for callers that
 * cannot allocate memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same statistics and output,
 * with no hipMalloc, hipFree or host buffer: the work arrays live in a
 * file-scope device array of 10.6 MiB that each GPU holds until the
 * process ends.  Grids with Ny * Nz above 131072 or Ny above 8192 are
 * refused with CS3D_ERR_ARG.  Calls that overlap on one GPU share that
 * array and race; calls on one solver handle never overlap. */
int cs3d_diag_shear_layer_noalloc(cs3d_solver *s, int fix_q12, double *out);
/* AeroOptics::computeAO (AeroOptics.cpp:45-93, the reference calls it
 * from postStep, CSolver_AWS.cpp:2941-2942, with num_angles 5 and max_angle
 * 70, CSolver_AWS.hpp:161-165): for each angle p/(num_angles-1) * max_angle
 * (degrees) and every (x_i, z_k), the integral of the density along the
 * ray from y = 0.2 Ly to 0.8 Ly leaning by that angle in x (x wrapped with
 * fmod(., Lx)), by trilinear interpolation (Utils.cpp:261-332) and the
 * trapezoid rule, FP64.  opl[p*Nx*Nz + k*Nx + i], num_angles*Nx*Nz doubles.
 * set 1 integrates the state's density; set 4 the *2 array (the reference
 * passes rho2, which after a filtered step holds the density filtered along
 * two of the three directions).  2 <= num_angles <= 64, 0 <= max_angle <
 * 90.  The reference also appends each angle's integrals to "<angle>.OPL"
 * (writeOPLFiles, AeroOptics.cpp:27-43); cs3d.hpp does that. */
int cs3d_aero_optics(cs3d_solver *s, int set, int num_angles, double max_angle, double *opl);
/* the same integrals of any FP64 density field rho (XYZ) on the grid
 * x_i = i/(Nx-1) Lx, ... (Domain.hpp:30-45), without a solver */
int cs3d_aero_optics_field(int Nx, int Ny, int Nz, double Lx, double Ly, double Lz,
                           const double *rho, int num_angles, double max_angle, double *opl);
/* volume-averaged kinetic energy of the current state, FP64 */
int cs3d_kinetic_energy(cs3d_solver *s, double *ke);

/* Discrete adjoint (no reference counterpart: the reference has no
 * adjoint; these serve gradient-based callers such as optimisation and
 * sensitivity studies).  Periodic or non-periodic lines (the boundary
 * closures transposed), sponge faces, walls or both; in the compensated
 * FP32 mode the replays retrace the compensated forward pass and the
 * products take the FP32 states.
 *
 * cs3d_rhs_vjp: the transpose of the linearised stage map Q -> dt*RHS(Q)
 * (the map cs3d_rhs evaluates, with the dt of the last pre-step) applied to
 * the cotangents lam_*, at the conserved set `set` (1 = *1, 2 = *k).  With
 * sponges the averages and, at the first stage after a sponge update, the
 * pre-reset values the x derivatives read on the sponge planes are held
 * fixed; with walls the map is Q -> dt*RHS(W(Q)), W the wall correction
 * cs3d_rhs applies first.  Host fields in XYZ order; a NULL out_* skips
 * that field. */
int cs3d_rhs_vjp(cs3d_solver *s, int set, const double *lam_rho, const double *lam_rhoU,
                 const double *lam_rhoV, const double *lam_rhoW, const double *lam_rhoE,
                 double *out_rho, double *out_rhoU, double *out_rhoV, double *out_rhoW,
                 double *out_rhoE);
/* Advance nsteps steps of the configured scheme (RK4 or TVDRK3) from the
 * current state (as cs3d_step does) and return J, the final volume-averaged
 * kinetic energy, and dJ/dU0, dJ/dV0, dJ/dW0: its gradient with respect to
 * the initial velocity fields at fixed initial density and pressure, by the
 * discrete adjoint of the steps taken (the dt sequence of the forward pass
 * held fixed, the filter included; with sponges also the running averages,
 * the plane reset and the stale primitives, the averages starting as the
 * initial state, so the call must follow cs3d_set_initial_conditions
 * directly: CS3D_ERR_ARG otherwise).  States are recomputed from a
 * checkpoint every `checkpoint` steps; the reverse pass holds
 * ceil(nsteps/checkpoint) + S*checkpoint + 1 states (S = 4 for RK4, 3 for
 * TVDRK3; 5 fields each; with sponges 2 more per checkpoint and 3 more per
 * step of a segment) and 77 work fields on periodic lines without
 * sponges, 108 on the general path (non-periodic lines), 122 with sponges
 * (with walls 116, and one state per wall face),
 * kept from call to call until cs3d_adjoint_release or cs3d_destroy.
 * The solver is left at the final state.  NULL outputs are skipped. */
int cs3d_adjoint_ke_gradient(cs3d_solver *s, int nsteps, int checkpoint, double *J, double *gU,
                             double *gV, double *gW);
/* frees the adjoint's device memory (the next adjoint call allocates again) */
int cs3d_adjoint_release(cs3d_solver *s);

int cs3d_synchronize(cs3d_solver *s);
/* HIP stream the solver launches on (hipStream_t) */
void *cs3d_stream(cs3d_solver *s);

/* Per-launch timing with HIP events recorded on the solver's stream around
 * every kernel launch (no reference counterpart: the reference's useTiming
 * tic/toc, Utils.cpp:377-386, times host phases).  Classes: 0 X sweep,
 * 1 Y sweep, 2 Z sweep, 3 filter sweeps, 4 reductions, 5 other (dt,
 * sponge).  The report sums milliseconds and counts launches since the
 * last enable or report, then restarts. */
int cs3d_profile_enable(cs3d_solver *s, int enabled);
int cs3d_profile_report(cs3d_solver *s, double ms[6], int counts[6]);

/* The HBM rate this GPU sustains on plain streams, from hand-written
 * 16-byte-vector kernels (no reference counterpart: a measured ceiling for
 * the sweeps' roofline fractions, and a copy of known size to calibrate
 * HBM traffic counters).  kind 0: copy b = a (2 arrays), 1: triad
 * a = b + s c (3 arrays), 2: the column sweeps' mix (5 arrays read and 9
 * written per element: 14 arrays) of `bytes` each, in six launch shapes (grid-
 * strided or one pass, 4 to 16 vectors in flight per lane); per shape 2
 * untimed then `reps` timed launches on the null stream, HIP events around
 * each.  Bytes read plus written per second (1e9): the best launch of any
 * shape, and the mean of that shape's launches.  Allocates and frees its
 * arrays. */
int cs3d_stream_probe(int kind, size_t bytes, int reps, double *best_gbps, double *mean_gbps);

/*
 * Standalone compact operators: Derivatives(dom, bcType, dir)
 * (Derivatives.hpp:54) + calc1stDerivField / calc2ndDerivField
 * (Derivatives.cpp:3-96), and Filter(alphaF, dom, bcType, dir)
 * (Filter.hpp:31) + filterField (Filter.cpp:129-171).
 * CS3D_LAYOUT_REFERENCE uses the reference's per-direction layout
 * (DIRX: XYZ, DIRY: YZX, DIRZ: ZXY, Derivatives.cpp:13,27,40).
 * precision CS3D_FP64: in/out are double*; CS3D_FP32: float*.
 * ptr_kind CS3D_HOST stages through device buffers; CS3D_DEVICE runs in
 * place on device memory, ordered on the ops stream.
 * Kinds CS3D_FILTER_T, CS3D_DERIV1_T, CS3D_DERIV2_T apply the transposed
 * operator (out = M^T in for the operator's matrix M along each line: the
 * product the discrete adjoint needs, e.g. <D f, g> = <f, D^T g>; no
 * reference counterpart).  They need in != out and, on non-periodic lines,
 * at least 18 points (24 for the filter).
 */
int cs3d_ops_create(int Nx, int Ny, int Nz, double Lx, double Ly, double Lz, int bc_type,
                    int dir, double alpha_f, int precision, cs3d_ops **out);
int cs3d_ops_apply(cs3d_ops *o, int kind, int layout, const void *in, void *out, int ptr_kind);
int cs3d_ops_synchronize(cs3d_ops *o);
void cs3d_ops_destroy(cs3d_ops *o);

/*
 * The host FP64 factorisation of one compact operator (no reference
 * counterpart: the reference refactorises every line inside solveTri /
 * cyclic, Utils.cpp:4-46).  Returns the tables the kernels run on, for
 * inspection and for CPU checks of the factorisation:
 * ints = {n, M, P, K, uniform, cyclic, NB, nreals}; reals = {scale, tap[5],
 * bcl[4][7], bcr[4][7], m[n], iw[n], mu[n], pf[n], pb[n], phi[P*K],
 * psi[P*K]}; segs = {sstart[P], slen[P]}.  With reals == NULL only ints is
 * filled.
 */
int cs3d_plan_info(int kind, int bc_type, int n, double dh, double alpha_f, int precision,
                   int ints[8], double *reals, size_t nreals, int *segs, size_t nsegs);

#ifdef __cplusplus
}
#endif
#endif
