/*
 * cs3d.hpp -- C++ drop-in for the solver classes of emathew1/3DCompactSolver,
 * implemented over the C ABI of cs3d.h (header-only; link
 * lib3dcompact_hip.so).
 *
 * Same class names, constructors, public members and call sequence as the
 * reference:
 *   Domain        Domain.hpp:6-52         TimeStepping  TimeStepping.hpp:6-45
 *   BC            BC.hpp:6-27             IdealGas      IdealGas.hpp:8-55
 *   Derivatives   Derivatives.hpp:13-183  Filter        Filter.hpp:11-115
 *   AbstractCSolver AbstractCSolver.hpp:9-43
 *   AbstractRK    AbstractRK.hpp:6-14     RK4 / TVDRK3  RK4.hpp:7-170, TVDRK3.hpp:7-139
 *   CSolver_AWS   CSolver_AWS.hpp:119 (constructor), CSolver_AWS.cpp (phases)
 * The classes live in namespace cs3d; define CS3D_DROPIN_GLOBAL before the
 * include to also bring them into the global namespace, in place of the
 * reference's headers.
 *
 * What differs from the reference, by design:
 *  - CSolver_AWS keeps its state on the GPU.  The conserved quadruples
 *    (rho1 ... rhoE2) are host mirrors, filled by syncHostState() and by the
 *    diagnostics/dump calls; the IC arrays rho0 ... W0 are read once by
 *    setInitialConditions() (the reference deletes them there; here they
 *    stay valid until the solver is destroyed).
 *  - The RK stage update (AbstractRK::updateConservedData) is fused into the
 *    last sweep of each stage on the GPU; RK4/TVDRK3::updateConservedData
 *    are kept for the call sequence and do nothing.  A time integrator of
 *    the caller's own (an AbstractRK subclass that updates rho1 ... rhoE2)
 *    would act on host mirrors the GPU never reads back, so
 *    CSolver_AWS::setInitialConditions aborts unless cs3d::RK4 or
 *    cs3d::TVDRK3 has named the scheme through setRKOrder.
 *  - Errors: the reference returns void and prints; here a failed library
 *    call prints the library's message to std::cerr and aborts.
 */
#ifndef CS3D_HPP
#define CS3D_HPP

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "cs3d.h"

namespace cs3d {

inline void check(int rc, const char *what)
{
    if (rc != CS3D_OK) {
        std::cerr << " > cs3d error in " << what << ": " << cs3d_last_error() << std::endl;
        std::abort();
    }
}

// Domain.hpp:6-52: x_i = i L / (N - 1)
class Domain {
  public:
    int Nx, Ny, Nz, N;
    double *x, *y, *z;
    double Lx, Ly, Lz;
    double dx, dy, dz;

    Domain(int Nx, int Ny, int Nz, double Lx, double Ly, double Lz)
        : Nx(Nx), Ny(Ny), Nz(Nz), N(Nx * Ny * Nz), Lx(Lx), Ly(Ly), Lz(Lz),
          xs(Nx), ys(Ny), zs(Nz)
    {
        for (int i = 0; i < Nx; i++) xs[i] = ((double)i / ((double)Nx - 1.0)) * Lx;
        for (int j = 0; j < Ny; j++) ys[j] = ((double)j / ((double)Ny - 1.0)) * Ly;
        for (int k = 0; k < Nz; k++) zs[k] = ((double)k / ((double)Nz - 1.0)) * Lz;
        x = xs.data();
        y = ys.data();
        z = zs.data();
        dx = x[1] - x[0];
        dy = y[1] - y[0];
        dz = z[1] - z[0];
    }

  private:
    std::vector<double> xs, ys, zs;
};

// BC.hpp:6-27
class BC {
  public:
    enum BCType { PERIODIC_SOLVE = CS3D_PERIODIC_SOLVE, DIRICHLET_SOLVE = CS3D_DIRICHLET_SOLVE };
    enum BCKind {
        PERIODIC = CS3D_BC_PERIODIC,
        SPONGE = CS3D_BC_SPONGE,
        ADIABATIC_WALL = CS3D_BC_ADIABATIC_WALL,
        MOVING_ADIABATIC_WALL = CS3D_BC_MOVING_ADIABATIC_WALL,
        INLET = CS3D_BC_INLET
    };
    BCType bcXType, bcYType, bcZType;
    BCKind bcX0, bcX1, bcY0, bcY1, bcZ0, bcZ1;

    BC(BCType bcXType, BCKind bcX0, BCKind bcX1, BCType bcYType, BCKind bcY0, BCKind bcY1,
       BCType bcZType, BCKind bcZ0, BCKind bcZ1)
        : bcXType(bcXType), bcYType(bcYType), bcZType(bcZType), bcX0(bcX0), bcX1(bcX1),
          bcY0(bcY0), bcY1(bcY1), bcZ0(bcZ0), bcZ1(bcZ1)
    {
    }
};

// TimeStepping.hpp:6-45.  The reference's calcDtFromCFL always takes the
// CFL branch (CSolver_AWS.cpp:530-534); CONST_DT is kept for the signature.
class TimeStepping {
  public:
    enum TimeSteppingType { CONST_DT, CONST_CFL };
    TimeSteppingType timeSteppingType;
    double CFL;
    double dt = 0.0;
    int maxTimeStep;
    double maxTime;
    int filterStep;
    int checkStep;
    int dumpStep;

    TimeStepping(TimeSteppingType timeSteppingType, double CFL, int maxTimeStep, double maxTime,
                 int filterStep, int checkStep, int dumpStep)
        : timeSteppingType(timeSteppingType), CFL(CFL), maxTimeStep(maxTimeStep),
          maxTime(maxTime), filterStep(filterStep), checkStep(checkStep), dumpStep(dumpStep)
    {
    }
};

// IdealGas.hpp:8-55: the constants the solver reads (gamma and Pr may be
// set after construction, as the reference's drivers do with ig->Pr)
class IdealGas {
  public:
    double gamma = 1.4, Pr = 0.7;
    double p_ref, rho_ref = 1.0, T_ref = 1.0, mu_ref;
    IdealGas(Domain *, double mu_ref) : p_ref(1.0 / 1.4), mu_ref(mu_ref) {}
};

// Derivatives.hpp:13-183 over cs3d_ops.  Fields are in the direction's own
// layout (DIRX: XYZ, DIRY: YZX, DIRZ: ZXY; Derivatives.cpp:13,27,40), FP64,
// on the host.  calc*FieldDevice take device pointers in the same layout.
class Derivatives {
  public:
    enum Direct { DIRX = CS3D_DIRX, DIRY = CS3D_DIRY, DIRZ = CS3D_DIRZ };
    int Nx, Ny, Nz, N;
    double dx, dy, dz;
    Direct currentDir;
    BC::BCType bcType;

    Derivatives(Domain *dom, BC::BCType bcType, Direct currentDir)
        : Nx(dom->Nx), Ny(dom->Ny), Nz(dom->Nz), N(dom->N), dx(dom->dx), dy(dom->dy),
          dz(dom->dz), currentDir(currentDir), bcType(bcType),
          lineN(currentDir == DIRX ? dom->Nx : currentDir == DIRY ? dom->Ny : dom->Nz),
          lineL(currentDir == DIRX ? dom->Lx : currentDir == DIRY ? dom->Ly : dom->Lz)
    {
        check(cs3d_ops_create(Nx, Ny, Nz, dom->Lx, dom->Ly, dom->Lz, (int)bcType,
                              (int)currentDir, 0.0, CS3D_FP64, &ops),
              "Derivatives");
    }
    ~Derivatives()
    {
        cs3d_ops_destroy(ops);
        cs3d_ops_destroy(lineOps);
    }
    Derivatives(const Derivatives &) = delete;
    Derivatives &operator=(const Derivatives &) = delete;

    // Derivatives.cpp:3-96
    void calc1stDerivField(double *dataIn, double *dataOut)
    {
        check(cs3d_ops_apply(ops, CS3D_DERIV1, CS3D_LAYOUT_REFERENCE, dataIn, dataOut, CS3D_HOST),
              "calc1stDerivField");
    }
    void calc2ndDerivField(double *dataIn, double *dataOut)
    {
        check(cs3d_ops_apply(ops, CS3D_DERIV2, CS3D_LAYOUT_REFERENCE, dataIn, dataOut, CS3D_HOST),
              "calc2ndDerivField");
    }
    void calc1stDerivFieldDevice(const double *dataIn, double *dataOut)
    {
        check(cs3d_ops_apply(ops, CS3D_DERIV1, CS3D_LAYOUT_REFERENCE, dataIn, dataOut,
                             CS3D_DEVICE),
              "calc1stDerivFieldDevice");
    }
    void calc2ndDerivFieldDevice(const double *dataIn, double *dataOut)
    {
        check(cs3d_ops_apply(ops, CS3D_DERIV2, CS3D_LAYOUT_REFERENCE, dataIn, dataOut,
                             CS3D_DEVICE),
              "calc2ndDerivFieldDevice");
    }
    void synchronize() { check(cs3d_ops_synchronize(ops), "Derivatives::synchronize"); }

    // the per-line entry points on one contiguous line of the direction's
    // length (Derivatives.hpp:170-171, 180-183; Derivatives.cpp:98-117,
    // 250-284): the bc-specific forms take the same route, which follows
    // bcType.  The stencil-only helpers multRHS* have no counterpart (they
    // are internal to the reference's solves; the GPU path forms B f inside
    // its line kernels).
    void calc1stDeriv(double *phi, double *dphi) { line(CS3D_DERIV1, phi, dphi); }
    void calc2ndDeriv(double *phi, double *dphi) { line(CS3D_DERIV2, phi, dphi); }
    void Compact1stPeriodic(double *phi, double *dphidx) { line(CS3D_DERIV1, phi, dphidx); }
    void Compact2ndPeriodic(double *phi, double *dphidx) { line(CS3D_DERIV2, phi, dphidx); }
    void Compact1stDirichlet(double *phi, double *dphidx) { line(CS3D_DERIV1, phi, dphidx); }
    void Compact2ndDirichlet(double *phi, double *dphidx) { line(CS3D_DERIV2, phi, dphidx); }

  private:
    void line(int kind, const double *in, double *out)
    {
        if (!lineOps)
            check(cs3d_ops_create(lineN, 1, 1, lineL, 1.0, 1.0, (int)bcType, CS3D_DIRX, 0.0,
                                  CS3D_FP64, &lineOps),
                  "Derivatives (line)");
        check(cs3d_ops_apply(lineOps, kind, CS3D_LAYOUT_XYZ, in, out, CS3D_HOST),
              "Derivatives (line)");
    }
    cs3d_ops *ops = nullptr;
    cs3d_ops *lineOps = nullptr;  // one line of the direction, made on first use
    int lineN;
    double lineL;
};

// Filter.hpp:11-115 over cs3d_ops (filterField, Filter.cpp:129-171)
class Filter {
  public:
    int Nx, Ny, Nz, N;
    double alphaF;
    Derivatives::Direct currentDir;
    BC::BCType bcType;

    Filter(double alphaF, Domain *dom, BC::BCType bcType, Derivatives::Direct currentDir)
        : Nx(dom->Nx), Ny(dom->Ny), Nz(dom->Nz), N(dom->N), alphaF(alphaF),
          currentDir(currentDir), bcType(bcType),
          lineN(currentDir == Derivatives::DIRX ? dom->Nx
                : currentDir == Derivatives::DIRY ? dom->Ny : dom->Nz),
          lineL(currentDir == Derivatives::DIRX ? dom->Lx
                : currentDir == Derivatives::DIRY ? dom->Ly : dom->Lz)
    {
        check(cs3d_ops_create(Nx, Ny, Nz, dom->Lx, dom->Ly, dom->Lz, (int)bcType,
                              (int)currentDir, alphaF, CS3D_FP64, &ops),
              "Filter");
    }
    ~Filter()
    {
        cs3d_ops_destroy(ops);
        cs3d_ops_destroy(lineOps);
    }
    Filter(const Filter &) = delete;
    Filter &operator=(const Filter &) = delete;

    void filterField(double *dataIn, double *dataOut)
    {
        check(cs3d_ops_apply(ops, CS3D_FILTER, CS3D_LAYOUT_REFERENCE, dataIn, dataOut, CS3D_HOST),
              "filterField");
    }
    void filterFieldDevice(const double *dataIn, double *dataOut)
    {
        check(cs3d_ops_apply(ops, CS3D_FILTER, CS3D_LAYOUT_REFERENCE, dataIn, dataOut,
                             CS3D_DEVICE),
              "filterFieldDevice");
    }

    // one contiguous line of the direction's length (Filter.hpp:111-114,
    // Filter.cpp:98-127); the multRHS* stencil helpers have no counterpart
    void compactFilter(double *phi, double *phiF) { line(phi, phiF); }
    void FilterPeriodic(double *phi, double *phiF) { line(phi, phiF); }
    void FilterDirichlet(double *phi, double *phiF) { line(phi, phiF); }

  private:
    void line(const double *in, double *out)
    {
        if (!lineOps)
            check(cs3d_ops_create(lineN, 1, 1, lineL, 1.0, 1.0, (int)bcType, CS3D_DIRX, alphaF,
                                  CS3D_FP64, &lineOps),
                  "Filter (line)");
        check(cs3d_ops_apply(lineOps, CS3D_FILTER, CS3D_LAYOUT_XYZ, in, out, CS3D_HOST),
              "Filter (line)");
    }
    cs3d_ops *ops = nullptr;
    cs3d_ops *lineOps = nullptr;
    int lineN;
    double lineL;
};

// AeroOptics.hpp:20-74 over cs3d_aero_optics_field / cs3d_aero_optics:
// optical-path integrals of a density along slanted rays
// (AeroOptics.cpp:45-93).  A standalone object integrates the host array it
// was given (rho); the one CSolver_AWS owns integrates the solver's state on
// the device.
class AeroOptics {
  public:
    Domain *domain;
    IdealGas *idealGas;
    int Nx, Ny, Nz, N;
    double dx, dy, dz, ds;
    int Nline = 0;
    double dxLine = 0.0, dyLine = 0.0, dsLine = 0.0;
    double *rho, *p, *T;
    int numAngles;
    double maxAngle, lowerIBound, upperIBound;
    double totalDX = 0.0, totalDY, Ldiag = 0.0;
    double *OPL;
    std::string outDir = ".";

    AeroOptics(Domain *domain, IdealGas *idealGas, double *rho, double *p, double *T,
               int numAngles, double maxAngle)
        : domain(domain), idealGas(idealGas), Nx(domain->Nx), Ny(domain->Ny), Nz(domain->Nz),
          N(domain->N), dx(domain->dx), dy(domain->dy), dz(domain->dz),
          ds(std::fmin(dx, std::fmin(dy, dz))), rho(rho), p(p), T(T), numAngles(numAngles),
          maxAngle(maxAngle), lowerIBound(0.2 * domain->Ly), upperIBound(0.8 * domain->Ly),
          totalDY(upperIBound - lowerIBound),
          OPL(new double[(size_t)numAngles * domain->Nx * domain->Nz])
    {
    }
    ~AeroOptics() { delete[] OPL; }
    AeroOptics(const AeroOptics &) = delete;
    AeroOptics &operator=(const AeroOptics &) = delete;

    // AeroOptics.cpp:3-24: the ray geometry of one angle (for callers that
    // read Ldiag, Nline ...; the integrals compute their own)
    void generateBaseBeamDiff(int angleNumber)
    {
        const double ang = ((double)angleNumber / ((double)(numAngles - 1))) * maxAngle * M_PI / 180.0;
        Ldiag = totalDY / std::cos(ang);
        totalDX = Ldiag * std::sin(ang);
        Nline = (int)std::ceil(Ldiag / ds) + 1;
        dsLine = Ldiag / (double)(Nline - 1);
        dxLine = totalDX / (double)(Nline - 1);
        dyLine = totalDY / (double)(Nline - 1);
    }
    // AeroOptics.cpp:27-43: appends each angle's Nx*Nz integrals to
    // "<angle>.OPL" (binary FP64)
    void writeOPLFiles()
    {
        for (int ap = 0; ap < numAngles; ap++) {
            const double ang = ((double)ap / ((double)(numAngles - 1))) * maxAngle;
            char buff[64];
            std::snprintf(buff, sizeof(buff), "%4.2f", ang);
            std::ofstream fp(outDir + "/" + buff + ".OPL", std::ios::binary | std::ios::app);
            fp.write((const char *)&OPL[(size_t)ap * Nx * Nz], sizeof(double) * Nx * Nz);
        }
    }
    // AeroOptics.cpp:45-93
    void computeAO()
    {
        if (solver)
            check(cs3d_aero_optics(solver, 1, numAngles, maxAngle, OPL), "computeAO");
        else
            check(cs3d_aero_optics_field(Nx, Ny, Nz, domain->Lx, domain->Ly, domain->Lz, rho,
                                         numAngles, maxAngle, OPL),
                  "computeAO");
        writeOPLFiles();
    }

    // set by CSolver_AWS: integrate its state's density on the device
    cs3d_solver *solver = nullptr;
};

// AbstractCSolver.hpp:9-43, with one extension: the RK driver tells the
// solver its order before setInitialConditions (the GPU state is built
// there and fuses the stage update).
class AbstractCSolver {
  public:
    int rkStep = 0;
    bool rkLast = false;
    bool endFlag = false;
    bool useTiming = false;

    double *rho0 = nullptr, *p0 = nullptr, *U0 = nullptr, *V0 = nullptr, *W0 = nullptr;

    double *rho1 = nullptr, *rhok = nullptr, *rhok2 = nullptr, *rho2 = nullptr;
    double *rhoU1 = nullptr, *rhoUk = nullptr, *rhoUk2 = nullptr, *rhoU2 = nullptr;
    double *rhoV1 = nullptr, *rhoVk = nullptr, *rhoVk2 = nullptr, *rhoV2 = nullptr;
    double *rhoW1 = nullptr, *rhoWk = nullptr, *rhoWk2 = nullptr, *rhoW2 = nullptr;
    double *rhoE1 = nullptr, *rhoEk = nullptr, *rhoEk2 = nullptr, *rhoE2 = nullptr;

    Domain *dom = nullptr;
    BC *bc = nullptr;
    TimeStepping *ts = nullptr;
    IdealGas *ig = nullptr;
    Derivatives *derivX = nullptr, *derivY = nullptr, *derivZ = nullptr;
    Filter *filtX = nullptr, *filtY = nullptr, *filtZ = nullptr;

    virtual ~AbstractCSolver() {}
    virtual void initializeSolverData() = 0;
    virtual void setInitialConditions() = 0;
    virtual void preStep() = 0;
    virtual void preSubStep() = 0;
    virtual void solveEqnSet() = 0;
    virtual void postSubStep() = 0;
    virtual void updateData() = 0;
    virtual void postStep() = 0;
    virtual void setRKOrder(int) {}
};

// AbstractRK.hpp:6-14
class AbstractRK {
  public:
    AbstractCSolver *cs = nullptr;
    virtual ~AbstractRK() {}
    virtual void executeSolverLoop() = 0;
    virtual void updateConservedData() = 0;
};

// RK4.hpp:7-170: the reference's loop, stage by stage
class RK4 : public AbstractRK {
  public:
    double a1 = 1.0 / 6.0, a2 = 1.0 / 3.0, a3 = 1.0 / 3.0, a4 = 1.0 / 6.0;
    double b1 = 1.0 / 2.0, b2 = 1.0 / 2.0, b3 = 1.0;
    explicit RK4(AbstractCSolver *cs) { this->cs = cs; }
    void updateConservedData() override {}
    void executeSolverLoop() override
    {
        cs->setRKOrder(4);
        cs->setInitialConditions();
        while (!cs->endFlag) {
            cs->rkLast = false;
            cs->preStep();
            for (int s = 1; s <= 4; s++) {
                cs->rkStep = s;
                cs->rkLast = s == 4;
                cs->preSubStep();
                cs->solveEqnSet();
                cs->postSubStep();
                updateConservedData();
                cs->updateData();
            }
            cs->postStep();
        }
    }
};

// TVDRK3.hpp:7-139
class TVDRK3 : public AbstractRK {
  public:
    explicit TVDRK3(AbstractCSolver *cs) { this->cs = cs; }
    void updateConservedData() override {}
    void executeSolverLoop() override
    {
        cs->setRKOrder(3);
        cs->setInitialConditions();
        while (!cs->endFlag) {
            cs->rkLast = false;
            cs->preStep();
            for (int s = 1; s <= 3; s++) {
                cs->rkStep = s;
                cs->rkLast = s == 3;
                cs->preSubStep();
                cs->solveEqnSet();
                cs->postSubStep();
                updateConservedData();
                cs->updateData();
            }
            cs->postStep();
        }
    }
};

// CSolver_AWS (CSolver_AWS.hpp:119, CSolver_AWS.cpp) on the GPU.  The
// pre/sub/post phases map onto cs3d_pre_step / cs3d_sub_step /
// cs3d_post_step; solveEqnSet runs the whole stage (derivatives, RHS,
// boundary handling and the RK update).
class CSolver_AWS : public AbstractCSolver {
  public:
    double alphaF, mu_ref;
    int blocksize;
    double time = 0.0, dt = 0.0;
    int timeStep = 0;
    // moving-wall velocities (CSolver_AWS.hpp:105-107, zero by default)
    double X0WallV = 0.0, X0WallW = 0.0, X1WallV = 0.0, X1WallW = 0.0;
    double Y0WallU = 0.0, Y0WallW = 0.0, Y1WallU = 0.0, Y1WallW = 0.0;
    double Z0WallU = 0.0, Z0WallV = 0.0, Z1WallU = 0.0, Z1WallV = 0.0;
    // GPU-path options (no reference counterpart)
    int precision = CS3D_FP64;       // CS3D_FP32: FP32 bulk arithmetic
    int compensated = 0;             // FP32 only: state and RK sums as hi + lo pairs
    bool tgvDiagnostics = false;     // append taylorgreen.out every step
    bool turbDiagnostics = false;    // append turbdata.out every step
    std::string outDir = ".";
    // shearLayerInfoCalc's files every 25 steps, as the reference's postStep
    // writes them for every run; fixShearQ12 separates v'w' from u'w' (Q12)
    bool shearLayerOutput = true, fixShearQ12 = false;
    // AeroOptics, as the reference builds it (CSolver_AWS.hpp:161-165: 5
    // angles to 70 degrees) and runs it every aoWriteStep steps from
    // postStep (CSolver_AWS.cpp:2941-2942), appending "<angle>.OPL" files
    // in outDir; here on the state's density (the reference passes rho2,
    // which after a filtered step holds the density filtered along two of
    // the three directions).  aeroOpticsOutput = false skips it.
    AeroOptics *ao = nullptr;
    int aoWriteStep = 1;
    bool aeroOpticsOutput = true;

    CSolver_AWS(Domain *dom, BC *bc, TimeStepping *ts, double alphaF, double mu_ref,
                int blocksize, bool useTiming)
        : alphaF(alphaF), mu_ref(mu_ref), blocksize(blocksize)
    {
        this->dom = dom;
        this->bc = bc;
        this->ts = ts;
        this->useTiming = useTiming;
        ig = new IdealGas(dom, mu_ref);
        initializeSolverData();
        ao = new AeroOptics(dom, ig, rho2, nullptr, nullptr, 5, 70.0);
    }
    ~CSolver_AWS() override
    {
        delete ao;
        if (h) cs3d_destroy(h);
        delete ig;
    }
    CSolver_AWS(const CSolver_AWS &) = delete;
    CSolver_AWS &operator=(const CSolver_AWS &) = delete;

    // CSolver_AWS.cpp:3-196: host arrays (IC inputs and the state mirrors)
    void initializeSolverData() override
    {
        const size_t N = (size_t)dom->N;
        store.assign(25 * N, 0.0);
        double *p = store.data();
        double **slots[25] = {&rho0,  &p0,    &U0,    &V0,     &W0,    &rho1,   &rhok,
                              &rhok2, &rho2,  &rhoU1, &rhoUk,  &rhoUk2, &rhoU2, &rhoV1,
                              &rhoVk, &rhoVk2, &rhoV2, &rhoW1, &rhoWk, &rhoWk2, &rhoW2,
                              &rhoE1, &rhoEk, &rhoEk2, &rhoE2};
        for (int i = 0; i < 25; i++) *slots[i] = p + i * N;
    }

    void setRKOrder(int order) override { rkOrder = order; }

    // CSolver_AWS.cpp:199-504
    void setInitialConditions() override
    {
        if (rkOrder != 3 && rkOrder != 4) {
            std::cerr << " > CSolver_AWS advances its state on the GPU inside solveEqnSet; "
                         "drive it with cs3d::RK4 or cs3d::TVDRK3 (they call setRKOrder), not "
                         "an AbstractRK of your own"
                      << std::endl;
            std::abort();
        }
        cs3d_config c;
        check(cs3d_config_default(&c, dom->Nx, dom->Ny, dom->Nz), "cs3d_config_default");
        c.Lx = dom->Lx;
        c.Ly = dom->Ly;
        c.Lz = dom->Lz;
        c.bc_type[0] = bc->bcXType;
        c.bc_type[1] = bc->bcYType;
        c.bc_type[2] = bc->bcZType;
        const int kinds[6] = {bc->bcX0, bc->bcX1, bc->bcY0, bc->bcY1, bc->bcZ0, bc->bcZ1};
        for (int f = 0; f < 6; f++) c.bc_kind[f] = kinds[f];
        c.cfl = ts->CFL;
        c.filter_step = ts->filterStep;
        c.max_steps = ts->maxTimeStep;
        c.max_time = ts->maxTime;
        c.alpha_f = alphaF;
        c.mu_ref = mu_ref;
        c.gamma = ig->gamma;
        c.Pr = ig->Pr;
        c.rk = rkOrder == 3 ? CS3D_TVDRK3 : CS3D_RK4;
        c.precision = precision;
        c.compensated = compensated;
        const double wv[6][3] = {{0.0, X0WallV, X0WallW}, {0.0, X1WallV, X1WallW},
                                 {Y0WallU, 0.0, Y0WallW}, {Y1WallU, 0.0, Y1WallW},
                                 {Z0WallU, Z0WallV, 0.0}, {Z1WallU, Z1WallV, 0.0}};
        for (int f = 0; f < 6; f++)
            for (int m = 0; m < 3; m++) c.wall_vel[f][m] = wv[f][m];
        if (h) cs3d_destroy(h);
        h = nullptr;
        check(cs3d_create(&c, &h), "setInitialConditions (create)");
        if (restartStep >= 0) loadRestartFields(restartStep);
        check(cs3d_set_initial_conditions(h, rho0, U0, V0, W0, p0), "setInitialConditions");
        if (restartStep >= 0) applyRestart(restartStep);
        check(cs3d_set_diagnostics(h, tgvDiagnostics || turbDiagnostics), "set_diagnostics");
        ao->solver = h;
        if (useTiming) check(cs3d_profile_enable(h, 1), "profile_enable");
        endFlag = false;
    }

    // CSolver_AWS.cpp:2897-2904: the step-0 dump, then calcDtFromCFL
    // (:507-551)
    void preStep() override
    {
        if (timeStep == 0) dumpSolution();
        check(cs3d_pre_step(h), "preStep");
    }
    void preSubStep() override {}
    void solveEqnSet() override { check(cs3d_sub_step(h, rkStep, rkLast ? 1 : 0), "solveEqnSet"); }
    void postSubStep() override {}
    void updateData() override {}

    // CSolver_AWS.cpp:2930-2946: diagnostics, updateSponge, checkSolution,
    // dumpSolution, checkEnd
    void postStep() override
    {
        check(cs3d_get_time(h, &time, &dt, &timeStep), "get_time");
        ts->dt = dt;
        if (shearLayerOutput) shearLayerInfoCalc();
        check(cs3d_post_step(h), "postStep");
        if (tgvDiagnostics) calcTaylorGreenQuantities();
        if (turbDiagnostics) calcTurbulenceQuantities();
        checkSolution();
        dumpSolution();
        if (aeroOpticsOutput && aoWriteStep > 0 && timeStep % aoWriteStep == 0) {
            ao->outDir = outDir;
            ao->computeAO();
        }
        if (timeStep >= ts->maxTimeStep || time >= ts->maxTime) endFlag = true;
    }

    // Gradients by the discrete adjoint (no reference counterpart: the
    // reference has no adjoint).  keGradient advances nsteps steps of the
    // solver's scheme (RK4 or TVDRK3) as the time loop does and returns J,
    // the final volume-averaged kinetic
    // energy, with dJ/dU0, dJ/dV0, dJ/dW0 (host arrays of N doubles, XYZ;
    // NULL skips one) at fixed initial rho and p, states recomputed from a
    // checkpoint every `checkpoint` steps.  rhsVJP gives
    // (d(dt*RHS)/dQ)^T lam at the state (set 1) or stage state (set 2).
    // Any boundary set, either precision, the compensated mode (cs3d.h).
    double keGradient(int nsteps, int checkpoint, double *gU, double *gV, double *gW)
    {
        double J = 0.0;
        check(cs3d_adjoint_ke_gradient(h, nsteps, checkpoint, &J, gU, gV, gW), "keGradient");
        check(cs3d_get_time(h, &time, &dt, &timeStep), "get_time");
        return J;
    }
    void rhsVJP(int set, const double *const lam[5], double *const out[5])
    {
        check(cs3d_rhs_vjp(h, set, lam[0], lam[1], lam[2], lam[3], lam[4], out[0], out[1], out[2],
                           out[3], out[4]),
              "rhsVJP");
    }
    // frees the adjoint's device memory
    void adjointRelease() { check(cs3d_adjoint_release(h), "adjointRelease"); }

    // calcTaylorGreenQuantities (CSolver_AWS.cpp:2175-2257): appends a line
    // to taylorgreen.out in the reference's format (:2240-2248)
    void calcTaylorGreenQuantities()
    {
        double o[7];
        check(cs3d_diag_taylor_green(h, o), "calcTaylorGreenQuantities");
        std::ofstream f(outDir + "/taylorgreen.out", std::ios::app);
        f << std::setprecision(17);
        for (int i = 0; i < 7; i++) f << o[i] << (i < 6 ? " " : "");
        f << std::endl;
    }
    // shearLayerInfoCalc (CSolver_AWS.cpp:2259-2456): every 25 steps, x-z
    // plane statistics per y index appended to rhoprime2.out ... rhoAvg.out
    // (a line: time, then the Ny values, a blank after each)
    void shearLayerInfoCalc()
    {
        if (timeStep % 25 != 0) return;
        const int Ny = dom->Ny;
        std::vector<double> o(11 * (size_t)Ny);
        check(cs3d_diag_shear_layer(h, fixShearQ12 ? 1 : 0, o.data()), "shearLayerInfoCalc");
        const char *files[11] = {"rhoprime2", "uprime2", "vprime2", "wprime2", "pprime2", "tprime2",
                                 "uvprime",   "uwprime", "vwprime", "uAvg",    "rhoAvg"};
        for (int f = 0; f < 11; f++) {
            std::ofstream out(outDir + "/" + files[f] + ".out", std::ios::app);
            out << std::setprecision(17) << time << " ";
            for (int j = 0; j < Ny; j++) out << o[(size_t)f * Ny + j] << " ";
            out << std::endl;
        }
    }

    // calcTurbulenceQuantities (CSolver_AWS.cpp:2033-2173): turbdata.out in
    // the reference's format (:2154-2164, a blank after every column)
    void calcTurbulenceQuantities()
    {
        double o[11];
        check(cs3d_diag_turbulence(h, o), "calcTurbulenceQuantities");
        std::ofstream f(outDir + "/turbdata.out", std::ios::app);
        f << std::setprecision(17);
        for (int i = 0; i < 11; i++) f << o[i] << " ";
        f << std::endl;
    }

    // checkSolution (CSolver_AWS.cpp:2459-2491): every checkStep steps,
    // step, time, dt and the ranges of the state (getRange, Utils.cpp:334-349)
    void checkSolution()
    {
        if (ts->checkStep <= 0 || timeStep % ts->checkStep != 0) return;
        const size_t N = (size_t)dom->N;
        std::vector<double> prim(8 * N);
        double *P[8];
        for (int i = 0; i < 8; i++) P[i] = prim.data() + i * N;
        check(cs3d_get_conserved(h, 1, rho1, rhoU1, rhoV1, rhoW1, rhoE1), "checkSolution");
        check(cs3d_get_primitives(h, P[0], P[1], P[2], P[3], P[4], P[5], P[6], P[7]),
              "checkSolution");
        std::cout << std::endl
                  << "-------------------------------------------------" << std::endl
                  << " Step = " << timeStep << ", time = " << time << ", dt = " << dt
                  << std::endl
                  << "-------------------------------------------------" << std::endl;
        const char *names[9] = {"RHO", "U", "V", "W", "P", "T", "mu", "RHOE", "SOS"};
        const double *data[9] = {rho1, P[0], P[1], P[2], P[3], P[4], P[5], rhoE1, P[7]};
        for (int v = 0; v < 9; v++) {
            double lo = 1e300, hi = -1e300;
            for (size_t i = 0; i < N; i++) {
                lo = std::fmin(lo, data[v][i]);
                hi = std::fmax(hi, data[v][i]);
            }
            std::cout << "  Range of " << names[v] << ": " << lo << ":" << hi << std::endl;
        }
        std::cout << std::endl;
    }

    // fills the host mirrors rho1 ... rhoE2 from the device state
    void syncHostState()
    {
        check(cs3d_get_conserved(h, 1, rho1, rhoU1, rhoV1, rhoW1, rhoE1), "sync *1");
        check(cs3d_get_conserved(h, 2, rhok, rhoUk, rhoVk, rhoWk, rhoEk), "sync *k");
        check(cs3d_get_conserved(h, 3, rhok2, rhoUk2, rhoVk2, rhoWk2, rhoEk2), "sync *k2");
        check(cs3d_get_conserved(h, 4, rho2, rhoU2, rhoV2, rhoW2, rhoE2), "sync *2");
    }

    // dumpSolution (CSolver_AWS.cpp:2493-2535): every dumpStep steps, raw
    // FP64 XYZ arrays rho.out.<step> ... rhoE.out.<step>; in addition the
    // run state restartSolution needs (restart.<step>, text) and, with
    // sponges, their averages and stale-plane copy (spongeAvg*.out.<step>,
    // spongeStale*.out.<step>)
    void dumpSolution()
    {
        if (ts->dumpStep <= 0 || timeStep % ts->dumpStep != 0) return;
        check(cs3d_get_conserved(h, 1, rho1, rhoU1, rhoV1, rhoW1, rhoE1), "dumpSolution");
        double *data[5] = {rho1, rhoU1, rhoV1, rhoW1, rhoE1};
        for (int v = 0; v < 5; v++) writeField(fieldName("", v), data[v]);
        if (hasSponge()) {
            std::vector<double> buf(5 * (size_t)dom->N);
            double *b[5];
            for (int v = 0; v < 5; v++) b[v] = buf.data() + v * (size_t)dom->N;
            for (int set = 5; set <= 6; set++) {
                check(cs3d_get_conserved(h, set, b[0], b[1], b[2], b[3], b[4]), "dumpSolution");
                for (int v = 0; v < 5; v++)
                    writeField(fieldName(set == 5 ? "spongeAvg" : "spongeStale", v), b[v]);
            }
        }
        cs3d_run_state rs;
        check(cs3d_get_run_state(h, &rs), "dumpSolution");
        std::ofstream f(outDir + "/restart." + std::to_string(timeStep));
        f << std::setprecision(17) << "timeStep " << rs.timeStep << "\n"
          << "filterTimeStep " << rs.filterTimeStep << "\n"
          << "time " << rs.time << "\n"
          << "dt " << rs.dt << "\n"
          << "cfl_max " << rs.cfl_max << "\n"
          << "rho_ref " << rs.rho_ref << "\n"
          << "rhoE_ref " << rs.rhoE_ref << "\n"
          << "stale " << rs.stale << "\n";
    }

    // No reference counterpart: with restartStep >= 0, setInitialConditions
    // continues from the files dumpSolution wrote at that step in outDir
    // instead of reading the IC arrays.
    int restartStep = -1;

  private:
    // the dumped state as IC arrays (for the handle's set-up)
    void loadRestartFields(int step)
    {
        const size_t N = (size_t)dom->N;
        double *data[5] = {rho1, rhoU1, rhoV1, rhoW1, rhoE1};
        for (int v = 0; v < 5; v++) readField(fieldName("", v, step), data[v]);
        const double g = ig->gamma;
        for (size_t i = 0; i < N; i++) {
            const double r = rho1[i], u = rhoU1[i] / r, v = rhoV1[i] / r, w = rhoW1[i] / r;
            rho0[i] = r;
            U0[i] = u;
            V0[i] = v;
            W0[i] = w;
            p0[i] = (g - 1.0) * (rhoE1[i] - 0.5 * r * (u * u + v * v + w * w));
        }
    }

    // the dumped run state and fields into the handle
    void applyRestart(int step)
    {
        const size_t N = (size_t)dom->N;
        cs3d_run_state rs{};
        std::ifstream f(outDir + "/restart." + std::to_string(step));
        std::string key;
        while (f >> key) {
            if (key == "timeStep") f >> rs.timeStep;
            else if (key == "filterTimeStep") f >> rs.filterTimeStep;
            else if (key == "time") f >> rs.time;
            else if (key == "dt") f >> rs.dt;
            else if (key == "cfl_max") f >> rs.cfl_max;
            else if (key == "rho_ref") f >> rs.rho_ref;
            else if (key == "rhoE_ref") f >> rs.rhoE_ref;
            else if (key == "stale") f >> rs.stale;
        }
        check(cs3d_set_run_state(h, &rs), "restartSolution");
        check(cs3d_set_conserved(h, 1, rho1, rhoU1, rhoV1, rhoW1, rhoE1), "restartSolution");
        if (hasSponge()) {
            std::vector<double> buf(5 * N);
            double *b[5];
            for (int v = 0; v < 5; v++) b[v] = buf.data() + v * N;
            for (int set = 5; set <= 6; set++) {
                for (int v = 0; v < 5; v++)
                    readField(fieldName(set == 5 ? "spongeAvg" : "spongeStale", v, step), b[v]);
                check(cs3d_set_conserved(h, set, b[0], b[1], b[2], b[3], b[4]), "restartSolution");
            }
        }
        // setting the state recomputed the CFL maximum; the dumped one stands
        check(cs3d_set_run_state(h, &rs), "restartSolution");
        timeStep = rs.timeStep;
        time = rs.time;
        dt = rs.dt;
    }

  public:

    cs3d_solver *handle() { return h; }

  private:
    bool hasSponge() const
    {
        const int kinds[6] = {bc->bcX0, bc->bcX1, bc->bcY0, bc->bcY1, bc->bcZ0, bc->bcZ1};
        for (int k : kinds)
            if (k == BC::SPONGE) return true;
        return false;
    }
    std::string fieldName(const char *prefix, int v, int step = -1) const
    {
        const char *names[5] = {"rho", "rhoU", "rhoV", "rhoW", "rhoE"};
        return outDir + "/" + prefix + names[v] + ".out." +
               std::to_string(step < 0 ? timeStep : step);
    }
    void writeField(const std::string &fn, const double *d) const
    {
        FILE *fp = std::fopen(fn.c_str(), "wb");
        if (!fp || std::fwrite(d, sizeof(double), (size_t)dom->N, fp) != (size_t)dom->N) {
            std::cerr << " > cannot write " << fn << std::endl;
            std::abort();
        }
        std::fclose(fp);
    }
    void readField(const std::string &fn, double *d) const
    {
        FILE *fp = std::fopen(fn.c_str(), "rb");
        if (!fp || std::fread(d, sizeof(double), (size_t)dom->N, fp) != (size_t)dom->N) {
            std::cerr << " > cannot read " << fn << std::endl;
            std::abort();
        }
        std::fclose(fp);
    }

    cs3d_solver *h = nullptr;
    int rkOrder = 0;  // set by RK4 / TVDRK3 through setRKOrder
    std::vector<double> store;
};

}  // namespace cs3d

#ifdef CS3D_DROPIN_GLOBAL
using cs3d::AbstractCSolver;
using cs3d::AeroOptics;
using cs3d::AbstractRK;
using cs3d::BC;
using cs3d::CSolver_AWS;
using cs3d::Derivatives;
using cs3d::Domain;
using cs3d::Filter;
using cs3d::IdealGas;
using cs3d::RK4;
using cs3d::TimeStepping;
using cs3d::TVDRK3;
#endif

#endif
