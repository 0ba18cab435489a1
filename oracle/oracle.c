/*
 * oracle.c -- FP64 CPU restatement of the 3DCompactSolver RHS pipeline.
 * TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline); the HIP
 * product never links or calls this file.  See oracle.h for what pins it.
 *
 * The reference works on transposed copies so that every line it solves is
 * contiguous (Utils.cpp:82-240).  The restatement gathers each line into a
 * contiguous buffer instead; the per-line arithmetic is the reference's, in
 * the reference's order, so the results are the same numbers.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXLINE 2048

/* Mutation switch for the oracle's own tests: tests/test_oracle_sponge.py
 * compiles private copies of this file with OR_PLANT set to 1..7, each
 * carrying one plausible mistake in the sponge path, and shows that its
 * closed-form checks reject every one.  Every library the product is checked
 * against is built with OR_PLANT = 0. */
#ifndef OR_PLANT
#define OR_PLANT 0
#endif

/* ------------------------------------------------------------------ */
/* Line solvers                                                         */
/* ------------------------------------------------------------------ */

/* Utils.cpp:4-20: Thomas algorithm, re-factorising on every call. */
void or_solve_tri(const double *a, const double *b, const double *c,
                  const double *d, double *x, double *work, int n)
{
    memcpy(work, b, (size_t)n * sizeof(double));
    memcpy(x, d, (size_t)n * sizeof(double));
    for (int i = 1; i < n; i++) {
        double m = a[i] / work[i - 1];
        work[i] = work[i] - m * c[i - 1];
        x[i] = x[i] - m * x[i - 1];
    }
    x[n - 1] /= work[n - 1];
    for (int i = n - 2; i >= 0; i--)
        x[i] = (x[i] - c[i] * x[i + 1]) / work[i];
}

/* Utils.cpp:22-46: Sherman-Morrison periodic solve (Numerical Recipes
 * "cyclic"): alpha is the bottom-left corner, beta the top-right one. */
void or_cyclic(const double *a, const double *b, const double *c,
               double alpha, double beta, const double *r, int n, double *x)
{
    double bb[MAXLINE], u[MAXLINE], z[MAXLINE], work[MAXLINE];
    if (n <= 2 || n > MAXLINE) {
        fprintf(stderr, "or_cyclic: bad n=%d\n", n);
        abort();
    }
    double gamma = -b[0];
    for (int i = 0; i < n; i++) bb[i] = b[i];
    bb[0] = b[0] - gamma;
    bb[n - 1] = b[n - 1] - alpha * beta / gamma;
    or_solve_tri(a, bb, c, r, x, work, n);
    for (int i = 0; i < n; i++) u[i] = 0.0;
    u[0] = gamma;
    u[n - 1] = alpha;
    or_solve_tri(a, bb, c, u, z, work, n);
    double fact = (x[0] + beta * x[n - 1] / gamma) /
                  (1.0 + z[0] + beta * z[n - 1] / gamma);
    for (int i = 0; i < n; i++) x[i] -= fact * z[i];
}

/* ------------------------------------------------------------------ */
/* Compact coefficients                                                 */
/* ------------------------------------------------------------------ */

/* Derivatives.hpp:78-120 */
static const double D1_ALPHA = 1.0 / 3.0, D1_A = 14.0 / 9.0, D1_B = 1.0 / 9.0;
static const double D1_AL11 = 5.0, D1_A1 = -197.0 / 60.0, D1_B1 = -5.0 / 12.0,
                    D1_C1 = 5.0, D1_D1 = -5.0 / 3.0, D1_E1 = 5.0 / 12.0,
                    D1_F1 = -1.0 / 20.0;
static const double D1_AL21 = 1.0 / 8.0, D1_AL22 = 3.0 / 4.0,
                    D1_A2 = -43.0 / 96.0, D1_B2 = -5.0 / 6.0, D1_C2 = 9.0 / 8.0,
                    D1_D2 = 1.0 / 6.0, D1_E2 = -1.0 / 96.0;
static const double D2_ALPHA = 2.0 / 11.0, D2_A = 12.0 / 11.0,
                    D2_B = 3.0 / 11.0;
static const double D2_AL11 = 10.0, D2_A1 = 145.0 / 12.0, D2_B1 = -76.0 / 3.0,
                    D2_C1 = 29.0 / 2.0, D2_D1 = -4.0 / 3.0, D2_E1 = 1.0 / 12.0;
static const double D2_AL21 = 2.0 / 11.0, D2_AL22 = -131.0 / 22.0,
                    D2_A2 = 177.0 / 88.0, D2_B2 = -507.0 / 44.0,
                    D2_C2 = 783.0 / 44.0, D2_D2 = -201.0 / 22.0,
                    D2_E2 = 81.0 / 88.0, D2_F2 = -3.0 / 44.0;

/* Derivatives.hpp:133-161 and Filter.hpp:83-104 */
void or_line_lhs(int kind, int bctype, int n, double alpha_f, double *a,
                 double *b, double *c)
{
    double al = kind == 1 ? D1_ALPHA : kind == 2 ? D2_ALPHA : alpha_f;
    for (int i = 0; i < n; i++) {
        b[i] = 1.0;
        a[i] = al;
        c[i] = al;
    }
    if (bctype != OR_DIRICHLET_SOLVE) return;
    if (kind == 1 || kind == 2) {
        double al11 = kind == 1 ? D1_AL11 : D2_AL11;
        double al21 = kind == 1 ? D1_AL21 : D2_AL21;
        double al22 = kind == 1 ? D1_AL22 : D2_AL22;
        c[0] = al11;
        c[1] = al22;
        a[1] = al21;
        a[n - 1] = al11;
        a[n - 2] = al22;
        c[n - 2] = al21;
    } else {
        a[0] = a[1] = a[2] = 0.0;
        c[0] = c[1] = c[2] = 0.0;
        a[n - 1] = a[n - 2] = a[n - 3] = 0.0;
        c[n - 1] = c[n - 2] = c[n - 3] = 0.0;
    }
}

/* wrap an index into [0, n) for periodic stencils */
static inline int wrapi(int i, int n)
{
    i %= n;
    return i < 0 ? i + n : i;
}

/* Derivatives.cpp:119-150 */
static void rhs_d1_periodic(double dh, const double *f, int n, double *r)
{
    double cc2 = -D1_B / 4.0, cc3 = -D1_A / 2.0, cc4 = D1_A / 2.0,
           cc5 = D1_B / 4.0;
    for (int i = 0; i < n; i++)
        r[i] = cc2 * f[wrapi(i - 2, n)] + cc3 * f[wrapi(i - 1, n)] +
               cc4 * f[wrapi(i + 1, n)] + cc5 * f[wrapi(i + 2, n)];
    for (int i = 0; i < n; i++) r[i] /= dh;
}

/* Derivatives.cpp:153-181 */
static void rhs_d2_periodic(double dh, const double *f, int n, double *r)
{
    double cc1 = D2_B / 4.0, cc2 = D2_A, cc3 = -(D2_B / 2.0 + 2.0 * D2_A),
           cc4 = D2_A, cc5 = D2_B / 4.0;
    for (int i = 0; i < n; i++)
        r[i] = cc1 * f[wrapi(i - 2, n)] + cc2 * f[wrapi(i - 1, n)] +
               cc3 * f[i] + cc4 * f[wrapi(i + 1, n)] +
               cc5 * f[wrapi(i + 2, n)];
    for (int i = 0; i < n; i++) r[i] /= dh * dh;
}

/* Derivatives.cpp:183-212 */
static void rhs_d1_dirichlet(double dh, const double *f, int n, double *r)
{
    double cc2 = -D1_B / 4.0, cc3 = -D1_A / 2.0, cc4 = D1_A / 2.0,
           cc5 = D1_B / 4.0;
    r[0] = D1_A1 * f[0] + D1_B1 * f[1] + D1_C1 * f[2] + D1_D1 * f[3] +
           D1_E1 * f[4] + D1_F1 * f[5];
    r[1] = D1_A2 * f[0] + D1_B2 * f[1] + D1_C2 * f[2] + D1_D2 * f[3] +
           D1_E2 * f[4];
    for (int i = 2; i < n - 2; i++)
        r[i] = cc2 * f[i - 2] + cc3 * f[i - 1] + cc4 * f[i + 1] +
               cc5 * f[i + 2];
    r[n - 2] = -D1_E2 * f[n - 5] - D1_D2 * f[n - 4] - D1_C2 * f[n - 3] -
               D1_B2 * f[n - 2] - D1_A2 * f[n - 1];
    r[n - 1] = -D1_F1 * f[n - 6] - D1_E1 * f[n - 5] - D1_D1 * f[n - 4] -
               D1_C1 * f[n - 3] - D1_B1 * f[n - 2] - D1_A1 * f[n - 1];
    for (int i = 0; i < n; i++) r[i] /= dh;
}

/* Derivatives.cpp:214-247 */
static void rhs_d2_dirichlet(double dh, const double *f, int n, double *r)
{
    double cc1 = D2_B / 4.0, cc2 = D2_A, cc3 = -(D2_B / 2.0 + 2.0 * D2_A),
           cc4 = D2_A, cc5 = D2_B / 4.0;
    r[0] = D2_A1 * f[0] + D2_B1 * f[1] + D2_C1 * f[2] + D2_D1 * f[3] +
           D2_E1 * f[4];
    r[1] = D2_A2 * f[0] + D2_B2 * f[1] + D2_C2 * f[2] + D2_D2 * f[3] +
           D2_E2 * f[4] + D2_F2 * f[5];
    for (int i = 2; i < n - 2; i++)
        r[i] = cc1 * f[i - 2] + cc2 * f[i - 1] + cc3 * f[i] +
               cc4 * f[i + 1] + cc5 * f[i + 2];
    r[n - 2] = D2_F2 * f[n - 6] + D2_E2 * f[n - 5] + D2_D2 * f[n - 4] +
               D2_C2 * f[n - 3] + D2_B2 * f[n - 2] + D2_A2 * f[n - 1];
    r[n - 1] = D2_E1 * f[n - 5] + D2_D1 * f[n - 4] + D2_C1 * f[n - 3] +
               D2_B1 * f[n - 2] + D2_A1 * f[n - 1];
    for (int i = 0; i < n; i++) r[i] /= dh * dh;
}

/* Filter.hpp:54-81 */
typedef struct {
    double a0_8, a1_8, a2_8, a3_8, a4_8;
    double a0_6, a1_6, a2_6, a3_6;
} filt_coef;

static filt_coef filter_coefs(double af)
{
    filt_coef k;
    k.a0_8 = (93.0 + 70.0 * af) / 128.0;
    k.a1_8 = (7.0 + 18.0 * af) / 16.0;
    k.a2_8 = (-7.0 + 14.0 * af) / 32.0;
    k.a3_8 = (1.0 - 2.0 * af) / 16.0;
    k.a4_8 = (-1.0 + 2.0 * af) / 128.0;
    k.a0_6 = (11.0 + 10.0 * af) / 16.0;
    k.a1_6 = (15.0 + 34.0 * af) / 32.0;
    k.a2_6 = (-3.0 + 6.0 * af) / 16.0;
    k.a3_6 = (1.0 - 2.0 * af) / 32.0;
    return k;
}
static const double F_A0[5] = {15.0 / 16.0, 4.0 / 16.0, -6.0 / 16.0,
                               4.0 / 16.0, -1.0 / 16.0};
static const double F_B0[5] = {1.0 / 16.0, 3.0 / 4.0, 6.0 / 16.0, -4.0 / 16.0,
                               1.0 / 16.0};
static const double F_C0[5] = {-1.0 / 16.0, 4.0 / 16.0, 5.0 / 8.0, 4.0 / 16.0,
                               -1.0 / 16.0};

/* Filter.cpp:5-62 (the reference's wrap cases assume n >= 8) */
static void rhs_filter_periodic(double af, const double *f, int n, double *r)
{
    filt_coef k = filter_coefs(af);
    double c0 = k.a0_8, c1 = k.a1_8 / 2.0, c2 = k.a2_8 / 2.0,
           c3 = k.a3_8 / 2.0, c4 = k.a4_8 / 2.0;
    for (int i = 0; i < n; i++)
        r[i] = c4 * f[wrapi(i - 4, n)] + c3 * f[wrapi(i - 3, n)] +
               c2 * f[wrapi(i - 2, n)] + c1 * f[wrapi(i - 1, n)] +
               c0 * f[i] + c1 * f[wrapi(i + 1, n)] +
               c2 * f[wrapi(i + 2, n)] + c3 * f[wrapi(i + 3, n)] +
               c4 * f[wrapi(i + 4, n)];
}

/* Filter.cpp:64-96 */
static void rhs_filter_dirichlet(double af, const double *f, int n, double *r)
{
    filt_coef k = filter_coefs(af);
    double a0 = k.a0_6, a1 = k.a1_6 / 2.0, a2 = k.a2_6 / 2.0,
           a3 = k.a3_6 / 2.0;
    double c0 = k.a0_8, c1 = k.a1_8 / 2.0, c2 = k.a2_8 / 2.0,
           c3 = k.a3_8 / 2.0, c4 = k.a4_8 / 2.0;
    r[0] = F_A0[0] * f[0] + F_A0[1] * f[1] + F_A0[2] * f[2] + F_A0[3] * f[3] +
           F_A0[4] * f[4];
    r[1] = F_B0[0] * f[0] + F_B0[1] * f[1] + F_B0[2] * f[2] + F_B0[3] * f[3] +
           F_B0[4] * f[4];
    r[2] = F_C0[0] * f[0] + F_C0[1] * f[1] + F_C0[2] * f[2] + F_C0[3] * f[3] +
           F_C0[4] * f[4];
    r[3] = a3 * f[0] + a2 * f[1] + a1 * f[2] + a0 * f[3] + a1 * f[4] +
           a2 * f[5] + a3 * f[6];
    r[n - 1] = F_A0[0] * f[n - 1] + F_A0[1] * f[n - 2] + F_A0[2] * f[n - 3] +
               F_A0[3] * f[n - 4] + F_A0[4] * f[n - 5];
    r[n - 2] = F_B0[0] * f[n - 1] + F_B0[1] * f[n - 2] + F_B0[2] * f[n - 3] +
               F_B0[3] * f[n - 4] + F_B0[4] * f[n - 5];
    r[n - 3] = F_C0[0] * f[n - 1] + F_C0[1] * f[n - 2] + F_C0[2] * f[n - 3] +
               F_C0[3] * f[n - 4] + F_C0[4] * f[n - 5];
    r[n - 4] = a3 * f[n - 1] + a2 * f[n - 2] + a1 * f[n - 3] + a0 * f[n - 4] +
               a1 * f[n - 5] + a2 * f[n - 6] + a3 * f[n - 7];
    for (int i = 4; i < n - 4; i++)
        r[i] = c4 * f[i - 4] + c3 * f[i - 3] + c2 * f[i - 2] + c1 * f[i - 1] +
               c0 * f[i] + c1 * f[i + 1] + c2 * f[i + 2] + c3 * f[i + 3] +
               c4 * f[i + 4];
}

void or_line_rhs(int kind, int bctype, int n, double dh, double alpha_f,
                 const double *phi, double *rhs)
{
    int per = bctype == OR_PERIODIC_SOLVE;
    if (kind == 1)
        per ? rhs_d1_periodic(dh, phi, n, rhs) : rhs_d1_dirichlet(dh, phi, n, rhs);
    else if (kind == 2)
        per ? rhs_d2_periodic(dh, phi, n, rhs) : rhs_d2_dirichlet(dh, phi, n, rhs);
    else
        per ? rhs_filter_periodic(alpha_f, phi, n, rhs)
            : rhs_filter_dirichlet(alpha_f, phi, n, rhs);
}

/* Compact{1st,2nd}{Periodic,Dirichlet} (Derivatives.cpp:250-284) and
 * Filter{Periodic,Dirichlet} (Filter.cpp:98-117). */
void or_line_op(int kind, int bctype, int n, double dh, double alpha_f,
                const double *phi, double *out)
{
    double a[MAXLINE], b[MAXLINE], c[MAXLINE], r[MAXLINE], w[MAXLINE];
    or_line_lhs(kind, bctype, n, alpha_f, a, b, c);
    or_line_rhs(kind, bctype, n, dh, alpha_f, phi, r);
    if (bctype == OR_PERIODIC_SOLVE) {
        double al = kind == 1 ? D1_ALPHA : kind == 2 ? D2_ALPHA : alpha_f;
        or_cyclic(a, b, c, al, al, r, n, out);
    } else {
        or_solve_tri(a, b, c, r, out, w, n);
    }
}

/* line geometry for a field: base offset of line L and point stride */
static void line_geom(int dir, int Nx, int Ny, int Nz, int refLayout,
                      long L, long *base, long *stride, int *n)
{
    if (dir == OR_DIRX) {
        /* L = k*Ny + j */
        *base = L * Nx;
        *stride = 1;
        *n = Nx;
    } else if (dir == OR_DIRY) {
        long k = L / Nx, i = L % Nx;
        if (refLayout) { /* YZX: i*Nz*Ny + k*Ny + j (Macros.hpp:9) */
            *base = i * (long)Nz * Ny + k * Ny;
            *stride = 1;
        } else {
            *base = k * (long)Nx * Ny + i;
            *stride = Nx;
        }
        *n = Ny;
    } else {
        long j = L / Nx, i = L % Nx;
        if (refLayout) { /* ZXY: j*Nx*Nz + i*Nz + k (Macros.hpp:10) */
            *base = j * (long)Nx * Nz + i * Nz;
            *stride = 1;
        } else {
            *base = j * (long)Nx + i;
            *stride = (long)Nx * Ny;
        }
        *n = Nz;
    }
}

static long nlines(int dir, int Nx, int Ny, int Nz)
{
    return dir == OR_DIRX ? (long)Ny * Nz
         : dir == OR_DIRY ? (long)Nx * Nz : (long)Nx * Ny;
}

/* calc1stDerivField / calc2ndDerivField (Derivatives.cpp:3-96) and
 * filterField (Filter.cpp:129-171): the same operator on every line. */
void or_field_op(int kind, int dir, int bctype, int Nx, int Ny, int Nz,
                 double dh, double alpha_f, int refLayout, const double *in,
                 double *out)
{
    long nl = nlines(dir, Nx, Ny, Nz);
#pragma omp parallel for schedule(static)
    for (long L = 0; L < nl; L++) {
        double fin[MAXLINE], fout[MAXLINE];
        long base, stride;
        int n;
        line_geom(dir, Nx, Ny, Nz, refLayout, L, &base, &stride, &n);
        for (int p = 0; p < n; p++) fin[p] = in[base + p * stride];
        or_line_op(kind, bctype, n, dh, alpha_f, fin, fout);
        for (int p = 0; p < n; p++) out[base + p * stride] = fout[p];
    }
}

/* ------------------------------------------------------------------ */
/* Solver                                                               */
/* ------------------------------------------------------------------ */

enum {
    OR_D_Ux, OR_D_Uy, OR_D_Uz, OR_D_Vx, OR_D_Vy, OR_D_Vz, OR_D_Wx, OR_D_Wy,
    OR_D_Wz, OR_D_Uxx, OR_D_Uyy, OR_D_Uzz, OR_D_Vxx, OR_D_Vyy, OR_D_Vzz,
    OR_D_Wxx, OR_D_Wyy, OR_D_Wzz, OR_D_Uxy, OR_D_Uxz, OR_D_Uyz, OR_D_Vxy,
    OR_D_Vxz, OR_D_Vyz, OR_D_Wxy, OR_D_Wxz, OR_D_Wyz, OR_D_Tx, OR_D_Ty,
    OR_D_Tz, OR_D_Txx, OR_D_Tyy, OR_D_Tzz, OR_D_contX, OR_D_contY,
    OR_D_contZ, OR_D_mxX, OR_D_mxY, OR_D_mxZ, OR_D_myX, OR_D_myY, OR_D_myZ,
    OR_D_mzX, OR_D_mzY, OR_D_mzZ, OR_D_eX, OR_D_eY, OR_D_eZ, OR_D_COUNT
};

struct or_solver {
    or_config c;
    long N;
    double dx, dy, dz;
    double *x, *y, *z;
    /* IdealGas.hpp:35-47 */
    double R_gas, cp, T_ref, mu_ref;
    /* conserved sets: [set][var], set 0 = *1, 1 = *k, 2 = *k2, 3 = *2 */
    double *q[4][5];
    double *prim[8]; /* U V W T p mu Amu sos */
    double *d[OR_D_COUNT];
    double *tmp[4];
    /* sponge */
    int sponge;
    double *sigma, *spgAvg[5];
    int timeStep, filterTimeStep;
    double time, dt, dt_override;
    int rkStep, rkLast;
};

#define U_ (s->prim[0])
#define V_ (s->prim[1])
#define W_ (s->prim[2])
#define T_ (s->prim[3])
#define P_ (s->prim[4])
#define MU_ (s->prim[5])
#define AMU_ (s->prim[6])
#define SOS_ (s->prim[7])

void or_config_default(or_config *c, int Nx, int Ny, int Nz)
{
    memset(c, 0, sizeof(*c));
    c->Nx = Nx;
    c->Ny = Ny;
    c->Nz = Nz;
    c->Lx = c->Ly = c->Lz = 1.0;
    c->alpha_f = 0.49;
    c->mu_ref = 1.0;
    c->gamma = 1.4;
    c->Pr = 0.7;
    c->cfl = 0.5;
    c->filter_step = 1;
    c->rk = 4;
    c->sponge_frac[0] = 0.25;
    c->sponge_frac[1] = 0.125;
    c->sponge_frac[2] = 0.25;
    c->sponge_strength = 6.0;
    c->sponge_avgT = 15.0;
    c->sponge_epsP = 0.005;
}

static double *alloc_field(long n)
{
    double *p = (double *)calloc((size_t)n, sizeof(double));
    if (!p) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return p;
}

/* SpongeBC.hpp:35-154.  The reference allocates sigma without initialising
 * it and takes fmax with the formula; fresh pages read as zero, which is
 * what this restatement uses. */
static void build_sponge(or_solver *s)
{
    const or_config *c = &s->c;
    int Nx = c->Nx, Ny = c->Ny, Nz = c->Nz;
    double st = c->sponge_strength;
    double LX = c->sponge_frac[0] * c->Lx, LY = c->sponge_frac[1] * c->Ly,
           LZ = c->sponge_frac[2] * c->Lz;
    s->sigma = alloc_field(s->N);
    for (int v = 0; v < 5; v++) s->spgAvg[v] = alloc_field(s->N);
    for (int k = 0; k < Nz; k++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                long ii = (long)k * Nx * Ny + (long)j * Nx + i;
                double sg = s->sigma[ii], sx;
                if (c->bc_kind[0] == OR_SPONGE && s->x[i] < LX) {
                    sx = (LX - s->x[i]) / LX;
                    sg = fmax(st * (0.068 * pow(sx, 2.0) + 0.845 * pow(sx, 8.0)), sg);
                }
                if (c->bc_kind[1] == OR_SPONGE && s->x[i] > c->Lx - LX) {
                    sx = (s->x[i] - (c->Lx - LX)) / LX;
                    sg = fmax(st * (0.068 * pow(sx, 2.0) + 0.845 * pow(sx, 8.0)), sg);
                }
                if (c->bc_kind[2] == OR_SPONGE && s->y[j] < LY) {
                    sx = (LY - s->y[j]) / LY;
                    sg = fmax(st * (0.068 * pow(sx, 2.0) + 0.845 * pow(sx, 8.0)), sg);
                }
                if (c->bc_kind[3] == OR_SPONGE && s->y[j] > c->Ly - LY) {
                    sx = (s->y[j] - (c->Ly - LY)) / LY;
                    sg = fmax(st * (0.068 * pow(sx, 2.0) + 0.845 * pow(sx, 8.0)), sg);
                }
                if (c->bc_kind[4] == OR_SPONGE && s->z[k] < LZ) {
                    sx = (LZ - s->z[k]) / LZ;
                    sg = fmax(st * (0.068 * pow(sx, 2.0) + 0.845 * pow(sx, 8.0)), sg);
                }
                if (c->bc_kind[5] == OR_SPONGE && s->z[k] > c->Lz - LZ) {
                    sx = (s->z[k] - (c->Lz - LZ)) / LZ;
                    sg = fmax(st * (0.068 * pow(sx, 2.0) + 0.845 * pow(sx, 8.0)), sg);
                }
                s->sigma[ii] = sg;
            }
}

or_solver *or_create(const or_config *cfg)
{
    or_solver *s = (or_solver *)calloc(1, sizeof(or_solver));
    s->c = *cfg;
    const or_config *c = &s->c;
    /* BC.hpp:11 declares INLET, but CSolver_AWS handles no such face */
    for (int f = 0; f < 6; f++)
        if (c->bc_kind[f] == OR_INLET) {
            fprintf(stderr, "oracle: inlet boundaries have no reference handling\n");
            free(s);
            return NULL;
        }
    /* the wall value takes the 6 points next to the face */
    for (int f = 0; f < 6; f++) {
        int n = f < 2 ? c->Nx : f < 4 ? c->Ny : c->Nz;
        if ((c->bc_kind[f] == OR_ADIABATIC_WALL ||
             c->bc_kind[f] == OR_MOVING_ADIABATIC_WALL) && n < 7) {
            fprintf(stderr, "oracle: a wall direction needs at least 7 points\n");
            free(s);
            return NULL;
        }
    }
    s->N = (long)c->Nx * c->Ny * c->Nz;
    /* Domain.hpp:30-45 */
    s->x = alloc_field(c->Nx);
    s->y = alloc_field(c->Ny);
    s->z = alloc_field(c->Nz);
    for (int i = 0; i < c->Nx; i++) s->x[i] = ((double)i / ((double)c->Nx - 1.0)) * c->Lx;
    for (int j = 0; j < c->Ny; j++) s->y[j] = ((double)j / ((double)c->Ny - 1.0)) * c->Ly;
    for (int k = 0; k < c->Nz; k++) s->z[k] = ((double)k / ((double)c->Nz - 1.0)) * c->Lz;
    s->dx = s->x[1] - s->x[0];
    s->dy = s->y[1] - s->y[0];
    s->dz = s->z[1] - s->z[0];
    /* IdealGas.hpp:35-47 with p_ref = 1/gamma, rho_ref = T_ref = 1 */
    s->T_ref = 1.0;
    s->R_gas = (1.0 / c->gamma) / 1.0 / 1.0;
    s->cp = s->R_gas * c->gamma / (c->gamma - 1.0);
    s->mu_ref = c->mu_ref;
    for (int st = 0; st < 4; st++)
        for (int v = 0; v < 5; v++) s->q[st][v] = alloc_field(s->N);
    for (int v = 0; v < 8; v++) s->prim[v] = alloc_field(s->N);
    for (int v = 0; v < OR_D_COUNT; v++) s->d[v] = alloc_field(s->N);
    for (int v = 0; v < 4; v++) s->tmp[v] = alloc_field(s->N);
    s->sponge = 0;
    for (int f = 0; f < 6; f++)
        if (c->bc_kind[f] == OR_SPONGE) s->sponge = 1;
    if (s->sponge) build_sponge(s);
    s->dt_override = -1.0;
    return s;
}

void or_destroy(or_solver *s)
{
    if (!s) return;
    for (int st = 0; st < 4; st++)
        for (int v = 0; v < 5; v++) free(s->q[st][v]);
    for (int v = 0; v < 8; v++) free(s->prim[v]);
    for (int v = 0; v < OR_D_COUNT; v++) free(s->d[v]);
    for (int v = 0; v < 4; v++) free(s->tmp[v]);
    if (s->sponge) {
        free(s->sigma);
        for (int v = 0; v < 5; v++) free(s->spgAvg[v]);
    }
    free(s->x);
    free(s->y);
    free(s->z);
    free(s);
}

/* IdealGas.hpp:49-79 */
static inline double ig_mu(const or_solver *s, double T)
{
    return s->mu_ref * pow(T / s->T_ref, 0.76);
}
static inline double ig_amu(const or_solver *s, double T)
{
    return 0.76 * (s->mu_ref / pow(s->T_ref, 0.76)) * pow(T, 0.76 - 1.0);
}
static inline double ig_p(const or_solver *s, double rho, double rhoE,
                          double U, double V, double W)
{
    return (s->c.gamma - 1) * (rhoE - 0.5 * rho * (U * U + V * V + W * W));
}
static inline double ig_T(const or_solver *s, double rho, double p)
{
    return p / (rho * s->R_gas);
}
static inline double ig_sos(const or_solver *s, double rho, double p)
{
    return sqrt(s->c.gamma * p / rho);
}
static inline double ig_rhoE(const or_solver *s, double rho, double p,
                             double U, double V, double W)
{
    return p / (s->c.gamma - 1.0) + 0.5 * rho * (U * U + V * V + W * W);
}

/* Utils.hpp:52 */
double or_calc_neumann(double f1, double f2, double f3, double f4, double f5,
                       double f6)
{
    return (f1 * 360.0 - f2 * 450.0 + f3 * 400.0 - f4 * 225.0 + f5 * 72.0 -
            f6 * 10.0) /
           147.0;
}

static int is_wall(int kind)
{
    return kind == OR_ADIABATIC_WALL || kind == OR_MOVING_ADIABATIC_WALL;
}

/* One wall face f (X0, X1, Y0, Y1, Z0, Z1) of the state qs and the
 * primitive arrays: the wall velocity (zero, or the moving wall's
 * tangential components) in U, V, W and the momenta, and T from calcNeumann
 * of the 6 points inward.  pre = 1 (preStepBCHandling,
 * CSolver_AWS.cpp:553-826) also sets p = rho R T and rhoE on the face;
 * pre = 0 is the initial-condition form (:235-484), which recomputes p and
 * rhoE everywhere afterwards. */
static void wall_face(or_solver *s, int f, double **qs, int pre)
{
    const or_config *c = &s->c;
    const int Nx = c->Nx, Ny = c->Ny, Nz = c->Nz;
    const int dir = f / 2, hi = f % 2;
    const long sz = (long)Nx * Ny;
    const long stride = dir == 0 ? 1 : dir == 1 ? Nx : sz;
    const long in = hi ? -stride : stride;
    const int moving = c->bc_kind[f] == OR_MOVING_ADIABATIC_WALL;
    double uw[3] = {0.0, 0.0, 0.0};
    for (int m = 0; m < 3; m++)
        if (moving && m != dir) uw[m] = c->wall_vel[f][m];
    const int na = dir == 0 ? Ny : Nx, nb = dir == 2 ? Ny : Nz;
    for (int b = 0; b < nb; b++)
        for (int a = 0; a < na; a++) {
            int i, j, k;
            if (dir == 0) {
                i = hi ? Nx - 1 : 0; j = a; k = b;
            } else if (dir == 1) {
                i = a; j = hi ? Ny - 1 : 0; k = b;
            } else {
                i = a; j = b; k = hi ? Nz - 1 : 0;
            }
            long ip = (long)k * sz + (long)j * Nx + i;
            double rho = qs[0][ip];
            U_[ip] = uw[0];
            V_[ip] = uw[1];
            W_[ip] = uw[2];
            for (int m = 0; m < 3; m++)
                qs[1 + m][ip] = (moving && m != dir) ? rho * uw[m] : 0.0;
            T_[ip] = or_calc_neumann(T_[ip + in], T_[ip + 2 * in], T_[ip + 3 * in],
                                     T_[ip + 4 * in], T_[ip + 5 * in],
                                     T_[ip + 6 * in]);
            if (pre) {
                P_[ip] = rho * s->R_gas * T_[ip];
                qs[4][ip] = ig_rhoE(s, rho, P_[ip], U_[ip], V_[ip], W_[ip]);
            }
        }
}

/* the faces in the reference's order: adiabatic walls X0..Z1, then moving
 * walls X0..Z1 (an edge takes the last face's values) */
static int wall_faces(or_solver *s, double **qs, int pre)
{
    int any = 0;
    for (int f = 0; f < 6; f++)
        if (s->c.bc_kind[f] == OR_ADIABATIC_WALL) {
            wall_face(s, f, qs, pre);
            any = 1;
        }
    for (int f = 0; f < 6; f++)
        if (s->c.bc_kind[f] == OR_MOVING_ADIABATIC_WALL) {
            wall_face(s, f, qs, pre);
            any = 1;
        }
    return any;
}

/* preStepBCHandling, CSolver_AWS.cpp:553-826: on the stage's state (*1 at
 * the first stage, *k after) */
static void pre_step_bc(or_solver *s)
{
    wall_faces(s, s->rkStep == 1 ? s->q[0] : s->q[1], 1);
}

/* CSolver_AWS.cpp:199-504 */
void or_set_ic(or_solver *s, const double *rho0, const double *U0,
               const double *V0, const double *W0, const double *p0)
{
    long N = s->N;
    double **q1 = s->q[0];
#pragma omp parallel for
    for (long ip = 0; ip < N; ip++) {
        q1[0][ip] = rho0[ip];
        U_[ip] = U0[ip];
        V_[ip] = V0[ip];
        W_[ip] = W0[ip];
        P_[ip] = p0[ip];
        q1[1][ip] = q1[0][ip] * U_[ip];
        q1[2][ip] = q1[0][ip] * V_[ip];
        q1[3][ip] = q1[0][ip] * W_[ip];
    }
#pragma omp parallel for
    for (long ip = 0; ip < N; ip++) {
        q1[4][ip] = ig_rhoE(s, q1[0][ip], P_[ip], U_[ip], V_[ip], W_[ip]);
        T_[ip] = ig_T(s, q1[0][ip], P_[ip]);
        MU_[ip] = ig_mu(s, T_[ip]);
        AMU_[ip] = ig_amu(s, T_[ip]);
        SOS_[ip] = ig_sos(s, q1[0][ip], P_[ip]);
    }
    if (wall_faces(s, q1, 0)) {
        /* :486-495: pressure, speed of sound and energy everywhere from the
         * (wall-corrected) temperature */
#pragma omp parallel for
        for (long ip = 0; ip < N; ip++) {
            P_[ip] = q1[0][ip] * s->R_gas * T_[ip];
            SOS_[ip] = ig_sos(s, q1[0][ip], P_[ip]);
            q1[4][ip] = ig_rhoE(s, q1[0][ip], P_[ip], U_[ip], V_[ip], W_[ip]);
        }
    }
    if (s->sponge) {
        for (int v = 0; v < 5; v++)
            memcpy(s->spgAvg[v], q1[v], (size_t)N * sizeof(double));
    }
    s->time = 0.0;
    s->timeStep = 0;
    s->filterTimeStep = 0;
}

/* CSolver_AWS.cpp:507-551 (Q1: always CFL-based) */
void or_pre_step(or_solver *s)
{
    long N = s->N;
    double mx = -100000.0;
#pragma omp parallel for reduction(max : mx)
    for (long ip = 0; ip < N; ip++) {
        double u = (fabs(U_[ip]) + SOS_[ip]) / s->dx +
                   (fabs(V_[ip]) + SOS_[ip]) / s->dy +
                   (fabs(W_[ip]) + SOS_[ip]) / s->dz;
        if (u > mx) mx = u;
    }
    s->dt = s->dt_override > 0.0 ? s->dt_override : s->c.cfl / mx;
    if (s->timeStep == 0) {
        s->timeStep++;
        s->time = 0.0;
    } else {
        s->timeStep++;
        s->time += s->dt;
    }
}

static void fop(or_solver *s, int kind, int dir, const double *in, double *out)
{
    const or_config *c = &s->c;
    double dh = dir == 0 ? s->dx : dir == 1 ? s->dy : s->dz;
    or_field_op(kind, dir, c->bc_type[dir], c->Nx, c->Ny, c->Nz, dh,
                c->alpha_f, 0, in, out);
}

/* preStepDerivatives, CSolver_AWS.cpp:829-1243 */
static void pre_step_derivatives(or_solver *s)
{
    long N = s->N;
    double **qs = s->rkStep == 1 ? s->q[0] : s->q[1];
    double *rho = qs[0], *rU = qs[1], *rV = qs[2], *rW = qs[3], *rE = qs[4];
    double **d = s->d;
    double *t1 = s->tmp[0], *t2 = s->tmp[1], *t3 = s->tmp[2], *t4 = s->tmp[3];

    /* X block (:861-936): fluxes from the primitive arrays as they stand */
#pragma omp parallel for
    for (long ip = 0; ip < N; ip++) {
        t1[ip] = rU[ip] * U_[ip] + P_[ip];
        t2[ip] = rV[ip] * U_[ip];
        t3[ip] = rW[ip] * U_[ip];
        t4[ip] = rE[ip] * U_[ip] + U_[ip] * P_[ip];
    }
    fop(s, 1, 0, U_, d[OR_D_Ux]);
    fop(s, 1, 0, rU, d[OR_D_contX]);
    fop(s, 1, 0, V_, d[OR_D_Vx]);
    fop(s, 1, 0, W_, d[OR_D_Wx]);
    fop(s, 2, 0, U_, d[OR_D_Uxx]);
    fop(s, 2, 0, V_, d[OR_D_Vxx]);
    fop(s, 2, 0, W_, d[OR_D_Wxx]);
    fop(s, 1, 0, T_, d[OR_D_Tx]);
    fop(s, 2, 0, T_, d[OR_D_Txx]);
    fop(s, 1, 0, t1, d[OR_D_mxX]);
    fop(s, 1, 0, t2, d[OR_D_myX]);
    fop(s, 1, 0, t3, d[OR_D_mzX]);
    fop(s, 1, 0, t4, d[OR_D_eX]);

    /* Y block (:949-1087): primitives recomputed from the stage state */
#pragma omp parallel for
    for (long ip = 0; ip < N; ip++) {
        U_[ip] = rU[ip] / rho[ip];
        V_[ip] = rV[ip] / rho[ip];
        W_[ip] = rW[ip] / rho[ip];
        P_[ip] = ig_p(s, rho[ip], rE[ip], U_[ip], V_[ip], W_[ip]);
        T_[ip] = ig_T(s, rho[ip], P_[ip]);
        t1[ip] = rU[ip] * V_[ip];
        t2[ip] = rV[ip] * V_[ip] + P_[ip];
        t3[ip] = rW[ip] * V_[ip];
        t4[ip] = rE[ip] * V_[ip] + V_[ip] * P_[ip];
    }
    fop(s, 1, 1, U_, d[OR_D_Uy]);
    fop(s, 1, 1, V_, d[OR_D_Vy]);
    fop(s, 1, 1, W_, d[OR_D_Wy]);
    fop(s, 1, 1, d[OR_D_Ux], d[OR_D_Uxy]);
    fop(s, 1, 1, d[OR_D_Vx], d[OR_D_Vxy]);
    fop(s, 1, 1, d[OR_D_Wx], d[OR_D_Wxy]);
    fop(s, 1, 1, rV, d[OR_D_contY]);
    fop(s, 2, 1, U_, d[OR_D_Uyy]);
    fop(s, 2, 1, V_, d[OR_D_Vyy]);
    fop(s, 2, 1, W_, d[OR_D_Wyy]);
    fop(s, 1, 1, T_, d[OR_D_Ty]);
    fop(s, 2, 1, T_, d[OR_D_Tyy]);
    fop(s, 1, 1, t1, d[OR_D_mxY]);
    fop(s, 1, 1, t2, d[OR_D_myY]);
    fop(s, 1, 1, t3, d[OR_D_mzY]);
    fop(s, 1, 1, t4, d[OR_D_eY]);

    /* Z block (:1095-1223) */
#pragma omp parallel for
    for (long ip = 0; ip < N; ip++) {
        U_[ip] = rU[ip] / rho[ip];
        V_[ip] = rV[ip] / rho[ip];
        W_[ip] = rW[ip] / rho[ip];
        P_[ip] = ig_p(s, rho[ip], rE[ip], U_[ip], V_[ip], W_[ip]);
        T_[ip] = ig_T(s, rho[ip], P_[ip]);
        t1[ip] = rU[ip] * W_[ip];
        t2[ip] = rV[ip] * W_[ip];
        t3[ip] = rW[ip] * W_[ip] + P_[ip];
        t4[ip] = rE[ip] * W_[ip] + W_[ip] * P_[ip];
    }
    fop(s, 1, 2, U_, d[OR_D_Uz]);
    fop(s, 2, 2, U_, d[OR_D_Uzz]);
    fop(s, 1, 2, V_, d[OR_D_Vz]);
    fop(s, 2, 2, V_, d[OR_D_Vzz]);
    fop(s, 1, 2, W_, d[OR_D_Wz]);
    fop(s, 2, 2, W_, d[OR_D_Wzz]);
    fop(s, 1, 2, T_, d[OR_D_Tz]);
    fop(s, 2, 2, T_, d[OR_D_Tzz]);
    fop(s, 1, 2, d[OR_D_Ux], d[OR_D_Uxz]);
    fop(s, 1, 2, d[OR_D_Vx], d[OR_D_Vxz]);
    fop(s, 1, 2, d[OR_D_Wx], d[OR_D_Wxz]);
    fop(s, 1, 2, d[OR_D_Uy], d[OR_D_Uyz]);
    fop(s, 1, 2, d[OR_D_Vy], d[OR_D_Vyz]);
    fop(s, 1, 2, d[OR_D_Wy], d[OR_D_Wyz]);
    fop(s, 1, 2, rW, d[OR_D_contZ]);
    fop(s, 1, 2, t1, d[OR_D_mxZ]);
    fop(s, 1, 2, t2, d[OR_D_myZ]);
    fop(s, 1, 2, t3, d[OR_D_mzZ]);
    fop(s, 1, 2, t4, d[OR_D_eZ]);

    /* :1226-1233 restore the primitives from the stage state */
#pragma omp parallel for
    for (long ip = 0; ip < N; ip++) {
        U_[ip] = rU[ip] / rho[ip];
        V_[ip] = rV[ip] / rho[ip];
        W_[ip] = rW[ip] / rho[ip];
        P_[ip] = ig_p(s, rho[ip], rE[ip], U_[ip], V_[ip], W_[ip]);
        T_[ip] = ig_T(s, rho[ip], P_[ip]);
    }
}

/* solveContinuity .. solveEnergy, CSolver_AWS.cpp:1246-1504 */
static void solve_eqn_set(or_solver *s)
{
    long N = s->N;
    double **qs = s->rkStep == 1 ? s->q[0] : s->q[1];
    double **k2 = s->q[2];
    double **d = s->d;
    double dt = s->dt;
    double cpPr = s->cp / s->c.Pr;
    const double t43 = 4.0 / 3.0, t13 = 1.0 / 3.0, t23 = 2.0 / 3.0;
#pragma omp parallel for
    for (long ip = 0; ip < N; ip++) {
        double spg[5] = {0, 0, 0, 0, 0};
        if (s->sponge)
            for (int v = 0; v < 5; v++)
                spg[v] = OR_PLANT == 2 ? s->sigma[ip] * (qs[v][ip] - s->spgAvg[v][ip])
                                       : s->sigma[ip] * (s->spgAvg[v][ip] - qs[v][ip]);
        double Ux = d[OR_D_Ux][ip], Uy = d[OR_D_Uy][ip], Uz = d[OR_D_Uz][ip];
        double Vx = d[OR_D_Vx][ip], Vy = d[OR_D_Vy][ip], Vz = d[OR_D_Vz][ip];
        double Wx = d[OR_D_Wx][ip], Wy = d[OR_D_Wy][ip], Wz = d[OR_D_Wz][ip];
        double Uxx = d[OR_D_Uxx][ip], Uyy = d[OR_D_Uyy][ip], Uzz = d[OR_D_Uzz][ip];
        double Vxx = d[OR_D_Vxx][ip], Vyy = d[OR_D_Vyy][ip], Vzz = d[OR_D_Vzz][ip];
        double Wxx = d[OR_D_Wxx][ip], Wyy = d[OR_D_Wyy][ip], Wzz = d[OR_D_Wzz][ip];
        double Uxy = d[OR_D_Uxy][ip], Uxz = d[OR_D_Uxz][ip];
        double Vxy = d[OR_D_Vxy][ip], Vyz = d[OR_D_Vyz][ip];
        double Wxz = d[OR_D_Wxz][ip], Wyz = d[OR_D_Wyz][ip];
        double Tx = d[OR_D_Tx][ip], Ty = d[OR_D_Ty][ip], Tz = d[OR_D_Tz][ip];
        double mu = MU_[ip], amu = AMU_[ip];
        double U = U_[ip], V = V_[ip], W = W_[ip];
        double MuX = amu * Tx, MuY = amu * Ty, MuZ = amu * Tz;
        double r;

        /* :1269 */
        k2[0][ip] = dt * (-d[OR_D_contX][ip] - d[OR_D_contY][ip] -
                          d[OR_D_contZ][ip] + spg[0]);
        /* :1311-1316 */
        r = mu * (t43 * Uxx + Uyy + Uzz + t13 * Vxy + t13 * Wxz);
        r += t43 * MuX * (Ux - 0.5 * Vy - 0.5 * Wz) + MuY * (Uy + Vx) +
             MuZ * (Wx + Uz);
        r += -d[OR_D_mxX][ip] - d[OR_D_mxY][ip] - d[OR_D_mxZ][ip] + spg[1];
        k2[1][ip] = r * dt;
        /* :1360-1365 */
        r = mu * (t43 * Vyy + Vxx + Vzz + t13 * Uxy + t13 * Wyz);
        r += t43 * MuY * (Vy - 0.5 * Ux - 0.5 * Wz) + MuX * (Uy + Vx) +
             MuZ * (Wy + Vz);
        r += -d[OR_D_myX][ip] - d[OR_D_myY][ip] - d[OR_D_myZ][ip] + spg[2];
        k2[2][ip] = r * dt;
        /* :1408-1414 */
        r = mu * (t43 * Wzz + Wyy + Wxx + t13 * Uxz + t13 * Vyz);
        r += t43 * MuZ * (Wz - 0.5 * Ux - 0.5 * Vy) + MuX * (Wx + Uz) +
             MuY * (Wy + Vz);
        r += -d[OR_D_mzX][ip] - d[OR_D_mzY][ip] - d[OR_D_mzZ][ip] + spg[3];
        k2[3][ip] = r * dt;
        /* :1458-1495 (Q3: the sponge source enters twice) */
        double qtemp = cpPr * (MuX * Tx + MuY * Ty + MuZ * Tz + mu * d[OR_D_Txx][ip] +
                               mu * d[OR_D_Tyy][ip] + mu * d[OR_D_Tzz][ip]);
        double vt1 = mu * (U * (t43 * Uxx + Uyy + Uzz) + V * (Vxx + t43 * Vyy + Vzz) +
                           W * (Wxx + Wyy + t43 * Wzz) +
                           t43 * (Ux * Ux + Vy * Vy + Wz * Wz) + Uy * Uy + Uz * Uz +
                           Vx * Vx + Vz * Vz + Wx * Wx + Wy * Wy -
                           t43 * (Ux * Vy + Ux * Wz + Vy * Wz) +
                           2.0 * (Uy * Vx + Uz * Wx + Vz * Wy) +
                           t13 * (U * Vxy + U * Wxz + V * Uxy) +
                           t13 * (V * Wyz + W * Uxz + W * Vyz));
        double vt2 = t43 * (U * MuX * Ux + V * MuY * Vy + W * MuZ * Wz) -
                     t23 * U * MuX * (Vy + Wz) - t23 * V * MuY * (Ux + Wz) -
                     t23 * W * MuZ * (Ux + Vy) + U * MuY * (Uy + Vx) +
                     U * MuZ * (Uz + Wx) + V * MuX * (Uy + Vx) +
                     V * MuZ * (Vz + Wy) + W * MuX * (Uz + Wx) +
                     W * MuY * (Vz + Wy);
        double eul = -d[OR_D_eX][ip] - d[OR_D_eY][ip] - d[OR_D_eZ][ip] + spg[4];
        k2[4][ip] = dt * (qtemp + vt1 + vt2 + eul + (OR_PLANT == 1 ? 0.0 : spg[4]));
    }
}

/* postStepBCHandling, CSolver_AWS.cpp:1506-1670: wall planes keep the
 * normal continuity flux term only (X0..Z1 in turn), then sponge planes are
 * zeroed */
static void post_step_bc(or_solver *s)
{
    const or_config *c = &s->c;
    int Nx = c->Nx, Ny = c->Ny, Nz = c->Nz;
    double **k2 = s->q[2];
    const double dt = s->dt;
    for (int f = 0; f < 6; f++) {
        if (!is_wall(c->bc_kind[f])) continue;
        int dir = f / 2, hi = f % 2;
        const double *cont = s->d[dir == 0 ? OR_D_contX : dir == 1 ? OR_D_contY : OR_D_contZ];
        for (int k = 0; k < Nz; k++)
            for (int j = 0; j < Ny; j++)
                for (int i = 0; i < Nx; i++) {
                    int idx = dir == 0 ? i : dir == 1 ? j : k;
                    int n = dir == 0 ? Nx : dir == 1 ? Ny : Nz;
                    if (idx != (hi ? n - 1 : 0)) continue;
                    long ip = (long)k * Nx * Ny + (long)j * Nx + i;
                    k2[0][ip] = -dt * cont[ip];
                    for (int v = 1; v < 5; v++) k2[v][ip] = 0.0;
                }
    }
    for (int f = 0; f < 6; f++) {
        if (c->bc_kind[f] != OR_SPONGE || OR_PLANT == 5) continue;
        int dir = f / 2, hi = f % 2;
        for (int k = 0; k < Nz; k++)
            for (int j = 0; j < Ny; j++)
                for (int i = 0; i < Nx; i++) {
                    int idx = dir == 0 ? i : dir == 1 ? j : k;
                    int n = dir == 0 ? Nx : dir == 1 ? Ny : Nz;
                    if (idx != (hi ? n - 1 : 0)) continue;
                    long ip = (long)k * Nx * Ny + (long)j * Nx + i;
                    for (int v = 0; v < 5; v++) k2[v][ip] = 0.0;
                }
    }
}

void or_rhs(or_solver *s, int rkStep)
{
    s->rkStep = rkStep;
    pre_step_bc(s);
    pre_step_derivatives(s);
    solve_eqn_set(s);
    post_step_bc(s);
}

/* RK4.hpp:35-118 and TVDRK3.hpp:38-96 on plain arrays:
 * a[0..4] = *1, a[5..9] = *k, a[10..14] = *k2, a[15..19] = *2 */
void or_rk_update_arrays(int rk, int rkStep, long N, double **a)
{
    double **q1 = a, **qk = a + 5, **k2 = a + 10, **q2 = a + 15;
    if (rk == 4) {
        const double ca[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
        const double cb[3] = {1.0 / 2.0, 1.0 / 2.0, 1.0};
        int st = rkStep;
        for (int v = 0; v < 5; v++) {
#pragma omp parallel for
            for (long ip = 0; ip < N; ip++) {
                if (st == 1)
                    q2[v][ip] = q1[v][ip] + ca[0] * k2[v][ip];
                else
                    q2[v][ip] += ca[st - 1] * k2[v][ip];
                if (st < 4) qk[v][ip] = q1[v][ip] + cb[st - 1] * k2[v][ip];
            }
        }
    } else {
        const double a2 = 3.0 / 4.0, a3 = 1.0 / 3.0, b2 = 1.0 / 4.0,
                     b3 = 2.0 / 3.0, c1 = 1.0, c2 = 1.0 / 4.0, c3 = 2.0 / 3.0;
        int st = rkStep;
        for (int v = 0; v < 5; v++) {
#pragma omp parallel for
            for (long ip = 0; ip < N; ip++) {
                if (st == 1)
                    qk[v][ip] = q1[v][ip] + c1 * k2[v][ip];
                else if (st == 2)
                    qk[v][ip] = a2 * q1[v][ip] + b2 * qk[v][ip] + c2 * k2[v][ip];
                else
                    q2[v][ip] = a3 * q1[v][ip] + b3 * qk[v][ip] + c3 * k2[v][ip];
            }
        }
    }
}

static void update_conserved(or_solver *s)
{
    double *a[20];
    for (int set = 0; set < 4; set++)
        for (int v = 0; v < 5; v++) a[set * 5 + v] = s->q[set][v];
    or_rk_update_arrays(s->c.rk, s->rkStep, s->N, a);
}

/* filterConservedData, CSolver_AWS.cpp:1682-1893: round-robin direction
 * order (Q11), result into *1; *2 is scratch afterwards. */
static void filter_conserved(or_solver *s)
{
    long N = s->N;
    double **q1 = s->q[0], **q2 = s->q[3];
    if (s->timeStep % s->c.filter_step == 0) {
        s->filterTimeStep++;
        int order[3];
        int m = s->filterTimeStep % 3;
        if (m == 1) { order[0] = 0; order[1] = 1; order[2] = 2; }
        else if (m == 2) { order[0] = 1; order[1] = 2; order[2] = 0; }
        else { order[0] = 2; order[1] = 0; order[2] = 1; }
        for (int v = 0; v < 5; v++) {
            fop(s, 0, order[0], q2[v], q1[v]);
            fop(s, 0, order[1], q1[v], q2[v]);
            fop(s, 0, order[2], q2[v], q1[v]);
        }
    } else {
        for (int v = 0; v < 5; v++)
            memcpy(q1[v], q2[v], (size_t)N * sizeof(double));
    }
}

/* updateNonConservedData, CSolver_AWS.cpp:1895-1935 */
static void update_nonconserved(or_solver *s)
{
    long N = s->N;
    double **qs = s->rkLast ? s->q[0] : s->q[1];
#pragma omp parallel for
    for (long ip = 0; ip < N; ip++) {
        double rho = qs[0][ip];
        U_[ip] = qs[1][ip] / rho;
        V_[ip] = qs[2][ip] / rho;
        W_[ip] = qs[3][ip] / rho;
        P_[ip] = ig_p(s, rho, qs[4][ip], U_[ip], V_[ip], W_[ip]);
        T_[ip] = ig_T(s, rho, P_[ip]);
        MU_[ip] = ig_mu(s, T_[ip]);
        AMU_[ip] = ig_amu(s, T_[ip]);
        SOS_[ip] = ig_sos(s, rho, P_[ip]);
    }
}

/* RK4.hpp:130-165: one stage */
void or_sub_step(or_solver *s, int rkStep, int rkLast)
{
    s->rkStep = rkStep;
    s->rkLast = rkLast;
    pre_step_bc(s);
    pre_step_derivatives(s);
    solve_eqn_set(s);
    post_step_bc(s);
    update_conserved(s);
    if (s->rkLast) filter_conserved(s);
    update_nonconserved(s);
}

/* updateSponge, CSolver_AWS.cpp:1937-2031 */
void or_post_step(or_solver *s)
{
    if (!s->sponge) return;
    const or_config *c = &s->c;
    long N = s->N;
    double **q1 = s->q[0];
    double eps = OR_PLANT == 3 ? s->dt / c->sponge_avgT : 1.0 / (c->sponge_avgT / s->dt + 1.0);
    double spP = 1.0 / c->gamma;
    double epsP = OR_PLANT == 4 ? 1.0 - c->sponge_epsP : c->sponge_epsP;
    double **A = s->spgAvg;
#pragma omp parallel for
    for (long ip = 0; ip < N; ip++) {
        for (int v = 0; v < 5; v++) A[v][ip] += eps * (q1[v][ip] - A[v][ip]);
        double **K = OR_PLANT == 7 ? q1 : A;
        A[4][ip] = epsP * A[4][ip] +
                   (1.0 - epsP) *
                       (spP / (c->gamma - 1.0) +
                        0.5 * (K[1][ip] * K[1][ip] + K[2][ip] * K[2][ip] +
                               K[3][ip] * K[3][ip]) / K[0][ip]);
    }
    int Nx = c->Nx, Ny = c->Ny, Nz = c->Nz;
    for (int f = 0; f < 6; f++) {
        if (c->bc_kind[f] != OR_SPONGE || OR_PLANT == 6) continue;
        int dir = f / 2, hi = f % 2;
        for (int k = 0; k < Nz; k++)
            for (int j = 0; j < Ny; j++)
                for (int i = 0; i < Nx; i++) {
                    int idx = dir == 0 ? i : dir == 1 ? j : k;
                    int n = dir == 0 ? Nx : dir == 1 ? Ny : Nz;
                    if (idx != (hi ? n - 1 : 0)) continue;
                    long ip = (long)k * Nx * Ny + (long)j * Nx + i;
                    for (int v = 0; v < 5; v++) q1[v][ip] = A[v][ip];
                }
    }
}

double or_step(or_solver *s)
{
    int ns = s->c.rk == 4 ? 4 : 3;
    or_pre_step(s);
    for (int st = 1; st <= ns; st++) or_sub_step(s, st, st == ns);
    or_post_step(s);
    return s->dt;
}

double or_dt(const or_solver *s) { return s->dt; }
double or_time(const or_solver *s) { return s->time; }
int or_timestep(const or_solver *s) { return s->timeStep; }
void or_set_dt_override(or_solver *s, double dt) { s->dt_override = dt; }
double *or_field(or_solver *s, int set, int which) { return s->q[set - 1][which]; }
double *or_prim(or_solver *s, int which) { return s->prim[which]; }
double *or_deriv(or_solver *s, int which) { return s->d[which]; }
const double *or_sponge_sigma(or_solver *s) { return s->sponge ? s->sigma : NULL; }
double *or_sponge_avg(or_solver *s, int which) { return s->sponge ? s->spgAvg[which] : NULL; }

/* calcTaylorGreenQuantities, CSolver_AWS.cpp:2175-2257 (Q8: derivative
 * arrays are those of the last stage, rho and U those after the filter) */
void or_diag_tgv(or_solver *s, double out[7])
{
    long N = s->N;
    double **d = s->d;
    double *rho1 = s->q[0][0];
    double ens = 0, ens2mu = 0, diss = 0, ke = 0, dil = 0, mmu = 0;
    double NN = (double)N;
#pragma omp parallel for reduction(+ : ens, ens2mu, diss, ke, dil, mmu)
    for (long ip = 0; ip < N; ip++) {
        double Ux = d[OR_D_Ux][ip], Uy = d[OR_D_Uy][ip], Uz = d[OR_D_Uz][ip];
        double Vx = d[OR_D_Vx][ip], Vy = d[OR_D_Vy][ip], Vz = d[OR_D_Vz][ip];
        double Wx = d[OR_D_Wx][ip], Wy = d[OR_D_Wy][ip], Wz = d[OR_D_Wz][ip];
        double S00 = 0.5 * (Ux + Ux) - (1.0 / 3.0) * (Ux + Vy + Wz);
        double S11 = 0.5 * (Vy + Vy) - (1.0 / 3.0) * (Ux + Vy + Wz);
        double S22 = 0.5 * (Wz + Wz) - (1.0 / 3.0) * (Ux + Vy + Wz);
        double S01 = 0.5 * (Uy + Vx), S02 = 0.5 * (Uz + Wx), S12 = 0.5 * (Vz + Wy);
        double td = S00 * Ux;
        td += S01 * Uy;
        td += S02 * Uz;
        td += S01 * Vx;
        td += S11 * Vy;
        td += S12 * Vz;
        td += S02 * Wx;
        td += S12 * Wy;
        td += S22 * Wz;
        td *= 2 * MU_[ip];
        double uv = U_[ip] * U_[ip] + V_[ip] * V_[ip] + W_[ip] * W_[ip];
        double kin = rho1[ip] * uv;
        double wx = Wy - Vz, wy = Uz - Wx, wz = Vx - Uy;
        ens += (1.0 / 2.0) * rho1[ip] * (wx * wx + wy * wy + wz * wz);
        ens2mu += MU_[ip] * rho1[ip] * (wx * wx + wy * wy + wz * wz);
        diss += td;
        ke += kin / 2.0;
        dil += ((Ux + Vy + Wz) * (Ux + Vy + Wz)) / NN;
        mmu += MU_[ip] / NN;
    }
    out[0] = s->time;
    out[1] = ke;
    out[2] = diss;
    out[3] = ens;
    out[4] = ens2mu;
    out[5] = mmu;
    out[6] = sqrt(dil);
}

/* calcTurbulenceQuantities, CSolver_AWS.cpp:2033-2173 */
void or_diag_turb(or_solver *s, double out[11])
{
    long N = s->N;
    double **d = s->d;
    double *rho1 = s->q[0][0];
    double NN = (double)N;
    double mRho = 0, mMu = 0, mTD = 0, up = 0, uip = 0, urms = 0, mSOS = 0,
           t2d = 0, mKE = 0;
#pragma omp parallel for reduction(+ : mRho, mMu, mTD, up, uip, urms, mSOS, t2d, mKE)
    for (long ip = 0; ip < N; ip++) {
        double Ux = d[OR_D_Ux][ip], Uy = d[OR_D_Uy][ip], Uz = d[OR_D_Uz][ip];
        double Vx = d[OR_D_Vx][ip], Vy = d[OR_D_Vy][ip], Vz = d[OR_D_Vz][ip];
        double Wx = d[OR_D_Wx][ip], Wy = d[OR_D_Wy][ip], Wz = d[OR_D_Wz][ip];
        double S00 = 0.5 * (Ux + Ux) - (1.0 / 3.0) * (Ux + Vy + Wz);
        double S11 = 0.5 * (Vy + Vy) - (1.0 / 3.0) * (Ux + Vy + Wz);
        double S22 = 0.5 * (Wz + Wz) - (1.0 / 3.0) * (Ux + Vy + Wz);
        double S01 = 0.5 * (Uy + Vx), S02 = 0.5 * (Uz + Wx), S12 = 0.5 * (Vz + Wy);
        double td = S00 * Ux;
        td += S01 * Uy;
        td += S02 * Uz;
        td += S01 * Vx;
        td += S11 * Vy;
        td += S12 * Vz;
        td += S02 * Wx;
        td += S12 * Wy;
        td += S22 * Wz;
        td *= 2 * MU_[ip];
        double uv = U_[ip] * U_[ip] + V_[ip] * V_[ip] + W_[ip] * W_[ip];
        mRho += rho1[ip] / NN;
        mMu += MU_[ip] / NN;
        mTD += td / NN;
        up += (uv / 3.0) / NN;
        uip += ((Ux * Ux + Vy * Vy + Wz * Wz) / 3.0) / NN;
        urms += uv / NN;
        mSOS += SOS_[ip] / NN;
        t2d += (Ux * Ux) / NN;
        mKE += (rho1[ip] * uv) / NN / 2.0;
    }
    double rp = 0, dp = 0, vp = 0;
#pragma omp parallel for reduction(+ : rp, dp, vp)
    for (long ip = 0; ip < N; ip++) {
        double Ux = d[OR_D_Ux][ip], Uy = d[OR_D_Uy][ip], Uz = d[OR_D_Uz][ip];
        double Vx = d[OR_D_Vx][ip], Vy = d[OR_D_Vy][ip], Vz = d[OR_D_Vz][ip];
        double Wx = d[OR_D_Wx][ip], Wy = d[OR_D_Wy][ip], Wz = d[OR_D_Wz][ip];
        rp += ((rho1[ip] - mRho) * (rho1[ip] - mRho)) / NN;
        dp += ((Ux + Vy + Wz) * (Ux + Vy + Wz)) / NN;
        double wx = Wy - Vz, wy = Uz - Wx, wz = Vx - Uy;
        vp += (wx * wx + wy * wy + wz * wz) / NN;
    }
    rp = sqrt(rp);
    dp = sqrt(dp);
    vp = sqrt(vp);
    double mNu = mMu / mRho;
    mTD /= mRho;
    up = sqrt(up);
    uip = sqrt(uip);
    urms = sqrt(urms);
    double turbMach = urms / mSOS;
    double taylorMicro2 = up / uip;
    double taylorReyn3 = up * taylorMicro2 / mNu;
    double k0 = 8.0;
    double tau = sqrt(2.0 * M_PI) / k0 / up;
    double kolNu = pow(pow(mNu, 3.0) / mTD, 0.25);
    (void)t2d;
    out[0] = s->time;
    out[1] = tau;
    out[2] = taylorReyn3;
    out[3] = turbMach;
    out[4] = mTD;
    out[5] = up;
    out[6] = kolNu;
    out[7] = rp;
    out[8] = dp;
    out[9] = vp;
    out[10] = mKE;
}

/* ------------------------------------------------------------------ */
/* Aero-optics                                                          */
/* ------------------------------------------------------------------ */

/* AeroOptics::computeAO (AeroOptics.cpp:45-93) with generateBaseBeamDiff
 * (:3-24) and Utils.cpp's getBaseNodeIndex / linearInterpolation
 * (:242-332): for angle p of nA (a_p = p/(nA-1) * maxAngle degrees) and
 * every (x_i, z_k), the density interpolated trilinearly at Nline points of
 * the ray from (x_i, 0.2 Ly, z_k) to 0.8 Ly leaning by a_p in x (x wrapped
 * with fmod(., Lx)), integrated by the trapezoid rule.  out[p Nx Nz + k Nx
 * + i].  Grid coordinates as Domain.hpp:30-45: x_i = i/(Nx-1) Lx.
 * Last-plane rule: where the base index is the last point of an axis the
 * reference reads one point past its arrays, with a weight that is zero
 * (the query lies on that plane: z_k at k = Nz-1) or of rounding size; here
 * the neighbour's coordinate is the formula's next point and its value the
 * last point's own. */
static double ao_coord(int i, int n, double L) { return (((double)i) / ((double)n - 1.0)) * L; }

static double ao_interp(int Nx, int Ny, int Nz, double Lx, double Ly, double Lz,
                        const double *f, const double xp[3])
{
    const double dx = ao_coord(1, Nx, Lx), dy = ao_coord(1, Ny, Ly), dz = ao_coord(1, Nz, Lz);
    int b[3] = {(int)(xp[0] / dx), (int)(xp[1] / dy), (int)(xp[2] / dz)};
    const int n[3] = {Nx, Ny, Nz};
    const double L[3] = {Lx, Ly, Lz};
    double d[3];
    int i0[3], i1[3];
    for (int a = 0; a < 3; a++) {
        const double c0 = ao_coord(b[a], n[a], L[a]), c1 = ao_coord(b[a] + 1, n[a], L[a]);
        d[a] = (xp[a] - c0) / (c1 - c0);
        i0[a] = b[a];
        i1[a] = b[a] + 1 < n[a] ? b[a] + 1 : n[a] - 1;
    }
#define AO_F(ii, jj, kk) f[(long)(kk) * Nx * Ny + (long)(jj) * Nx + (ii)]
    const double c000 = AO_F(i0[0], i0[1], i0[2]), c001 = AO_F(i0[0], i0[1], i1[2]);
    const double c010 = AO_F(i0[0], i1[1], i0[2]), c011 = AO_F(i0[0], i1[1], i1[2]);
    const double c100 = AO_F(i1[0], i0[1], i0[2]), c101 = AO_F(i1[0], i0[1], i1[2]);
    const double c110 = AO_F(i1[0], i1[1], i0[2]), c111 = AO_F(i1[0], i1[1], i1[2]);
#undef AO_F
    const double c00 = c000 * (1 - d[0]) + c100 * d[0];
    const double c01 = c001 * (1 - d[0]) + c101 * d[0];
    const double c10 = c010 * (1 - d[0]) + c110 * d[0];
    const double c11 = c011 * (1 - d[0]) + c111 * d[0];
    const double c0 = c00 * (1 - d[1]) + c10 * d[1];
    const double c1 = c01 * (1 - d[1]) + c11 * d[1];
    return c0 * (1 - d[2]) + c1 * d[2];
}

void or_aero_optics(int Nx, int Ny, int Nz, double Lx, double Ly, double Lz, const double *rho,
                    int nA, double maxAngle, double *out)
{
    const double dx = ao_coord(1, Nx, Lx), dy = ao_coord(1, Ny, Ly), dz = ao_coord(1, Nz, Lz);
    const double ds = fmin(dx, fmin(dy, dz));
    const double lower = 0.2 * Ly, upper = 0.8 * Ly, totalDY = upper - lower;
    for (int ap = 0; ap < nA; ap++) {
        const double ang = (((double)ap / ((double)(nA - 1))) * maxAngle) * M_PI / 180.0;
        const double Ldiag = totalDY / cos(ang), totalDX = Ldiag * sin(ang);
        const int Nline = (int)ceil(Ldiag / ds) + 1;
        const double dsLine = Ldiag / (double)(Nline - 1);
        const double dxLine = totalDX / (double)(Nline - 1), dyLine = totalDY / (double)(Nline - 1);
#pragma omp parallel for collapse(2)
        for (int k = 0; k < Nz; k++)
            for (int i = 0; i < Nx; i++) {
                /* the ray's points: sp = 0 .. Nline-1 */
#define AO_AT(sp)                                                                          \
    ao_interp(Nx, Ny, Nz, Lx, Ly, Lz, rho,                                                 \
              (const double[3]){fmod(ao_coord(i, Nx, Lx) + (double)((sp) * dxLine), Lx),   \
                                lower + (double)((sp) * dyLine), ao_coord(k, Nz, Lz)})
                /* trapezoid rule in the reference's order: the two ends, then
                 * the interior points in turn */
                double acc = AO_AT(0) * dsLine / 2.0 + AO_AT(Nline - 1) * dsLine / 2.0;
                for (int sp = 1; sp < Nline - 1; sp++) acc += AO_AT(sp) * dsLine;
#undef AO_AT
                out[(long)ap * Nx * Nz + (long)k * Nx + i] = acc;
            }
    }
}
