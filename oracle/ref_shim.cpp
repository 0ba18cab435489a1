// ref_shim.cpp -- C entry points over the reference's own classes, compiled
// together with the reference's Utils.cpp / Derivatives.cpp / Filter.cpp
// (from where they lie in the reference checkout) into oracle/_ref/libref.so.
// TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle against the
// reference's compiled code.  Nothing here ships in the product.
//
// Only header-only reference code is used beyond those three sources:
// Domain.hpp, BC.hpp, IdealGas.hpp, SpongeBC.hpp, RK4.hpp, TVDRK3.hpp,
// AbstractCSolver.hpp.  CSolver_AWS.cpp is not buildable here (it needs
// libpng's png.h through PngWriter.hpp), so the RHS assembly is pinned by
// other means.
#include <malloc.h>
#include <iostream>
#include <sstream>

#include <cstring>
#include <vector>

#include "AeroOptics.hpp"
#include "Domain.hpp"
#include "BC.hpp"
#include "Utils.hpp"
#include "Derivatives.hpp"
#include "Filter.hpp"
#include "IdealGas.hpp"
#include "SpongeBC.hpp"
#include "RK4.hpp"
#include "TVDRK3.hpp"

namespace {

// the reference constructors print progress to std::cout; keep test output
// readable by discarding it while they run
struct QuietCout {
    std::stringstream sink;
    std::streambuf *old;
    QuietCout() : old(std::cout.rdbuf(sink.rdbuf())) {}
    ~QuietCout() { std::cout.rdbuf(old); }
};

// a solver with only the state arrays set, enough for
// RK4/TVDRK3::updateConservedData (RK4.hpp:35-118, TVDRK3.hpp:38-96)
struct ArraySolver : public AbstractCSolver {
    void initializeSolverData() {}
    void setInitialConditions() {}
    void preStep() {}
    void preSubStep() {}
    void solveEqnSet() {}
    void postSubStep() {}
    void updateData() {}
    void postStep() {}
};

}  // namespace

extern "C" {

void ref_solve_tri(double *a, double *b, double *c, double *d, double *x,
                   int n)
{
    double *work = new double[n];
    solveTri(a, b, c, d, x, work, n);
    delete[] work;
}

void ref_cyclic(double *a, double *b, double *c, double alpha, double beta,
                double *r, int n, double *x)
{
    cyclic(a, b, c, alpha, beta, r, n, x);
}

// kind: 1 = calc1stDerivField, 2 = calc2ndDerivField, 0 = filterField.
// Input/output use the reference's layout contract for dir.
void ref_field_op(int kind, int dir, int bctype, int Nx, int Ny, int Nz,
                  double Lx, double Ly, double Lz, double alphaF, double *in,
                  double *out)
{
    QuietCout q;
    Domain dom(Nx, Ny, Nz, Lx, Ly, Lz);
    BC::BCType bt = bctype == 0 ? BC::PERIODIC_SOLVE : BC::DIRICHLET_SOLVE;
    Derivatives::Direct d = dir == 0   ? Derivatives::DIRX
                            : dir == 1 ? Derivatives::DIRY
                                       : Derivatives::DIRZ;
    if (kind == 0) {
        Filter f(alphaF, &dom, bt, d);
        f.filterField(in, out);
    } else {
        Derivatives der(&dom, bt, d);
        if (kind == 1)
            der.calc1stDerivField(in, out);
        else
            der.calc2ndDerivField(in, out);
    }
}

// Domain.hpp:16-52
void ref_domain(int Nx, int Ny, int Nz, double Lx, double Ly, double Lz,
                double *x, double *y, double *z, double *dxyz)
{
    QuietCout q;
    Domain dom(Nx, Ny, Nz, Lx, Ly, Lz);
    for (int i = 0; i < Nx; i++) x[i] = dom.x[i];
    for (int j = 0; j < Ny; j++) y[j] = dom.y[j];
    for (int k = 0; k < Nz; k++) z[k] = dom.z[k];
    dxyz[0] = dom.dx;
    dxyz[1] = dom.dy;
    dxyz[2] = dom.dz;
}

// IdealGas.hpp:35-79 pointwise relations, over arrays of length n.
// out: p, T, mu, Amu, sos, rhoE (rhoE recomputed from rho, p, U, V, W)
void ref_idealgas(int n, double mu_ref, const double *rho, const double *rhoU,
                  const double *rhoV, const double *rhoW, const double *rhoE,
                  double *out_p, double *out_T, double *out_mu,
                  double *out_Amu, double *out_sos, double *out_rhoE)
{
    QuietCout q;
    Domain dom(8, 8, 8, 1.0, 1.0, 1.0);
    IdealGas ig(&dom, mu_ref);
    for (int i = 0; i < n; i++) {
        double U = ig.solveU(rho[i], rhoU[i]);
        double V = ig.solveU(rho[i], rhoV[i]);
        double W = ig.solveU(rho[i], rhoW[i]);
        double p = ig.solvep(rho[i], rhoE[i], U, V, W);
        double T = ig.solveT(rho[i], p);
        out_p[i] = p;
        out_T[i] = T;
        out_mu[i] = ig.solveMu(T);
        out_Amu[i] = ig.solveAmu(T);
        out_sos[i] = ig.solveSOS(rho[i], p);
        out_rhoE[i] = ig.solverhoE(rho[i], p, U, V, W);
    }
}

// SpongeBC.hpp:35-154: sigma for the given face kinds (0 periodic,
// 1 sponge).  The reference allocates sigma with new double[] and reads it
// (fmax with the formula) before writing it.  While the constructor runs,
// glibc's M_PERTURB makes every allocation come back zero-filled (the value a
// fresh page holds), so the result is defined.
void ref_sponge_sigma(int Nx, int Ny, int Nz, double Lx, double Ly,
                      double Lz, const int *kinds, double *sigma_out)
{
    QuietCout q;
    Domain dom(Nx, Ny, Nz, Lx, Ly, Lz);
    IdealGas ig(&dom, 1.0);
    BC::BCKind kk[6];
    for (int f = 0; f < 6; f++) kk[f] = kinds[f] == 1 ? BC::SPONGE : BC::PERIODIC;
    BC bc(BC::PERIODIC_SOLVE, kk[0], kk[1], BC::PERIODIC_SOLVE, kk[2], kk[3],
          BC::PERIODIC_SOLVE, kk[4], kk[5]);
    mallopt(M_PERTURB, 0xff);  // allocations filled with ~0xff = 0x00
    SpongeBC spg(&dom, &ig, &bc);
    mallopt(M_PERTURB, 0);
    long N = (long)Nx * Ny * Nz;
    for (long i = 0; i < N; i++) sigma_out[i] = spg.sigma[i];
}

// RK4::updateConservedData / TVDRK3::updateConservedData on caller arrays.
// arrays: [set][var] with set 0 = *1, 1 = *k, 2 = *k2, 3 = *2
void ref_rk_update(int rk, int rkStep, int Nx, int Ny, int Nz,
                   double **arrays)
{
    QuietCout q;
    Domain dom(Nx, Ny, Nz, 1.0, 1.0, 1.0);
    ArraySolver cs;
    cs.dom = &dom;
    cs.useTiming = false;
    cs.rkStep = rkStep;
    double **a = arrays;
    cs.rho1 = a[0];  cs.rhoU1 = a[1];  cs.rhoV1 = a[2];  cs.rhoW1 = a[3];  cs.rhoE1 = a[4];
    cs.rhok = a[5];  cs.rhoUk = a[6];  cs.rhoVk = a[7];  cs.rhoWk = a[8];  cs.rhoEk = a[9];
    cs.rhok2 = a[10]; cs.rhoUk2 = a[11]; cs.rhoVk2 = a[12]; cs.rhoWk2 = a[13]; cs.rhoEk2 = a[14];
    cs.rho2 = a[15]; cs.rhoU2 = a[16]; cs.rhoV2 = a[17]; cs.rhoW2 = a[18]; cs.rhoE2 = a[19];
    if (rk == 4) {
        RK4 r(&cs);
        r.updateConservedData();
    } else {
        TVDRK3 r(&cs);
        r.updateConservedData();
    }
}

// Utils.hpp:52
double ref_calc_neumann(double f1, double f2, double f3, double f4, double f5, double f6)
{
    return calcNeumann(f1, f2, f3, f4, f5, f6);
}


// AeroOptics::computeAO (AeroOptics.cpp:45-93) of the reference on the
// density rho (XYZ).  computeAO also appends its results to "<angle>.OPL"
// files in the working directory (writeOPLFiles); callers run it in a
// scratch directory.  Where a ray point lies on the last plane of an axis
// the reference reads one point past its coordinate and field arrays (with
// weight zero); to keep those reads defined the domain's coordinate arrays
// are replaced by copies one point longer (the same formula,
// Domain.hpp:30-45) and the density is padded with zeros by a plane and a
// row.
void ref_aero_optics(int Nx, int Ny, int Nz, double Lx, double Ly, double Lz, const double *rho,
                     int numAngles, double maxAngle, double *out)
{
    QuietCout q;
    Domain dom(Nx, Ny, Nz, Lx, Ly, Lz);
    const int n[3] = {Nx, Ny, Nz};
    const double L[3] = {Lx, Ly, Lz};
    double **axes[3] = {&dom.x, &dom.y, &dom.z};
    std::vector<std::vector<double>> ext(3);
    for (int a = 0; a < 3; a++) {
        ext[a].resize(n[a] + 1);
        for (int i = 0; i <= n[a]; i++) ext[a][i] = (((double)i) / ((double)n[a] - 1.0)) * L[a];
        delete[] *axes[a];
        *axes[a] = ext[a].data();
    }
    const long N = (long)Nx * Ny * Nz;
    std::vector<double> padded(N + (long)Nx * Ny + Nx + 1, 0.0);
    std::memcpy(padded.data(), rho, N * sizeof(double));
    IdealGas ig(&dom, 1.0);
    AeroOptics ao(&dom, &ig, padded.data(), nullptr, nullptr, numAngles, maxAngle);
    ao.computeAO();
    std::memcpy(out, ao.OPL, (size_t)numAngles * Nx * Nz * sizeof(double));
    delete[] ao.OPL;
    // the coordinate arrays belong to `ext`; Domain has no destructor
    for (int a = 0; a < 3; a++) *axes[a] = nullptr;
}

}  // extern "C"
