/*
 * oracle.h -- FP64 CPU restatement of the 3DCompactSolver per-stage RHS
 * pipeline.  TEST INFRASTRUCTURE ONLY: this code is the checker that the HIP
 * path is compared against (and the timed CPU baseline in bench.py).  The
 * product library never links, loads or calls it.
 *
 * Every function follows a reference routine and cites it as
 * <file>:<line> of emathew1/3DCompactSolver.  Layout is the reference's XYZ
 * order, index = k*Nx*Ny + j*Nx + i (Macros.hpp:8).
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - solve_tri / cyclic / D1 / D2 / filter lines and fields: checked against
 *     the reference's own Utils.cpp, Derivatives.cpp, Filter.cpp compiled into
 *     oracle/_ref (when buildable) and against dense solves, modified
 *     wavenumber / transfer-function closed forms and polynomial exactness.
 *   - ideal gas, sponge sigma, RK4/TVDRK3 updates: against the reference
 *     headers compiled into oracle/_ref.
 *   - calcNeumann: against the reference's Utils.hpp compiled into
 *     oracle/_ref and polynomial exactness; wall handling: quiescent-box
 *     invariance, the no-slip/adiabatic values on the wall planes, and
 *     mass conservation in a closed box.
 *   - RHS assembly / full step: invariants (uniform state -> zero RHS, mass
 *     conservation, filter commutation), Taylor-Green closed forms at t=0, and
 *     the reference's recorded taylorgreen.out first row (64^3).
 */
#ifndef CS3D_ORACLE_H
#define CS3D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* BC.hpp:10-11 */
enum { OR_PERIODIC_SOLVE = 0, OR_DIRICHLET_SOLVE = 1 };
enum { OR_PERIODIC = 0, OR_SPONGE = 1, OR_ADIABATIC_WALL = 2,
       OR_MOVING_ADIABATIC_WALL = 3, OR_INLET = 4 };
/* Derivatives.hpp:48 */
enum { OR_DIRX = 0, OR_DIRY = 1, OR_DIRZ = 2 };

/* ---- line solvers (Utils.cpp:4-46) ---- */
void or_solve_tri(const double *a, const double *b, const double *c,
                  const double *d, double *x, double *work, int n);
void or_cyclic(const double *a, const double *b, const double *c,
               double alpha, double beta, const double *r, int n, double *x);

/* ---- compact operators on one contiguous line ----
 * kind: 1 = first derivative, 2 = second derivative, 0 = 8th-order filter.
 * bctype: OR_PERIODIC_SOLVE / OR_DIRICHLET_SOLVE.  dh is the grid spacing
 * (ignored by the filter), alpha_f the filter parameter (ignored by D1/D2). */
void or_line_op(int kind, int bctype, int n, double dh, double alpha_f,
                const double *phi, double *out);
/* the RHS stencil alone (B*phi) and the LHS tridiagonal (a, b, c) */
void or_line_rhs(int kind, int bctype, int n, double dh, double alpha_f,
                 const double *phi, double *rhs);
void or_line_lhs(int kind, int bctype, int n, double alpha_f,
                 double *a, double *b, double *c);

/* ---- compact operators on a full field ----
 * refLayout = 0: field in XYZ order, lines gathered with a stride.
 * refLayout = 1: field in the direction's own reference layout (DIRX: XYZ,
 *   DIRY: YZX, DIRZ: ZXY), i.e. the calc*DerivField contract
 *   (Derivatives.cpp:3-96, Filter.cpp:129-171). */
void or_field_op(int kind, int dir, int bctype, int Nx, int Ny, int Nz,
                 double dh, double alpha_f, int refLayout,
                 const double *in, double *out);

/* ---- solver ---- */
typedef struct {
    int Nx, Ny, Nz;
    double Lx, Ly, Lz;          /* Domain.hpp:16-52: x_i = i*L/(N-1) */
    int bc_type[3];             /* per direction */
    int bc_kind[6];             /* X0, X1, Y0, Y1, Z0, Z1 */
    double alpha_f;
    double mu_ref;
    double gamma, Pr;           /* IdealGas.hpp:39-40 defaults 1.4 / 0.7 */
    double cfl;
    int filter_step;
    int rk;                     /* 4 = RK4.hpp, 3 = TVDRK3.hpp */
    double sponge_frac[3];      /* SpongeBC.hpp: 0.25, 0.125, 0.25 */
    double sponge_strength;     /* 6.0 */
    double sponge_avgT;         /* 15.0 */
    double sponge_epsP;         /* 0.005 */
    /* moving-wall velocity (U, V, W) per face X0..Z1: the tangential
     * components are CSolver_AWS.hpp:105-107 (X0WallV, X0WallW, ...,
     * default 0); the normal component is ignored */
    double wall_vel[6][3];
} or_config;

void or_config_default(or_config *c, int Nx, int Ny, int Nz);

typedef struct or_solver or_solver;

or_solver *or_create(const or_config *cfg);
void or_destroy(or_solver *s);
/* setInitialConditions (CSolver_AWS.cpp:199-504); arrays are XYZ, length N */
void or_set_ic(or_solver *s, const double *rho0, const double *U0,
               const double *V0, const double *W0, const double *p0);
/* one full time step: preStep, 4 (or 3) stages, postStep minus IO.
 * Returns dt used. */
double or_step(or_solver *s);
/* pieces, for stage-level parity tests */
void or_pre_step(or_solver *s);                 /* calcDtFromCFL */
void or_sub_step(or_solver *s, int rkStep, int rkLast);
void or_post_step(or_solver *s);                /* updateSponge */
/* the stage RHS alone (preSubStep + solveEqnSet + postSubStep) */
void or_rhs(or_solver *s, int rkStep);
double or_dt(const or_solver *s);
double or_time(const or_solver *s);
int or_timestep(const or_solver *s);
void or_set_dt_override(or_solver *s, double dt); /* <0: CFL (default) */

/* RK4.hpp:35-118 / TVDRK3.hpp:38-96 on plain arrays:
 * a[0..4] = *1, a[5..9] = *k, a[10..14] = *k2, a[15..19] = *2 */
void or_rk_update_arrays(int rk, int rkStep, long N, double **a);

/* accessors: which = 0..4 -> rho, rhoU, rhoV, rhoW, rhoE;
 * set: 1 = *1 (state), 2 = *k, 3 = *k2 (dt*RHS), 4 = *2 */
double *or_field(or_solver *s, int set, int which);
/* primitives: 0 U, 1 V, 2 W, 3 T, 4 p, 5 mu, 6 Amu, 7 sos */
double *or_prim(or_solver *s, int which);
/* derivative arrays by name index (see oracle.c, OR_D_* enum) */
double *or_deriv(or_solver *s, int which);

/* calcTaylorGreenQuantities (CSolver_AWS.cpp:2175-2257):
 * out[0..6] = time, kineticEngSum, turbDissSum, enstrophySum,
 * enstrophySum2Mu, meanMu, dilprime */
void or_diag_tgv(or_solver *s, double out[7]);
/* calcTurbulenceQuantities (CSolver_AWS.cpp:2033-2173):
 * out[0..10] = time, tau, taylorReyn3, turbMach, meanTurbDiss, uprime,
 * kolNu, rhoprime, dilprime, vortprime, meanKineticEng */
void or_diag_turb(or_solver *s, double out[11]);

/* sponge strength field (SpongeBC.hpp:35-154), XYZ, length N (NULL if none) */
const double *or_sponge_sigma(or_solver *s);
/* sponge running average of conserved variable `which` (0..4: rho, rhoU,
 * rhoV, rhoW, rhoE; spongeRhoAvg .. spongeRhoEAvg of SpongeBC.hpp), XYZ,
 * length N, writable (NULL if no sponge) */
double *or_sponge_avg(or_solver *s, int which);

/* calcNeumann (Utils.hpp:52): the boundary value that makes the one-sided
 * 7-point first derivative vanish, from the 6 values next to the boundary */
double or_calc_neumann(double f1, double f2, double f3, double f4, double f5,
                       double f6);

/* AeroOptics::computeAO (AeroOptics.cpp:45-93): optical-path integrals of
 * the density rho (XYZ, FP64) along nA rays per (x_i, z_k), the angles
 * p/(nA-1) * maxAngle degrees; out[p*Nx*Nz + k*Nx + i] */
void or_aero_optics(int Nx, int Ny, int Nz, double Lx, double Ly, double Lz, const double *rho,
                    int nA, double maxAngle, double *out);

#ifdef __cplusplus
}
#endif
#endif
