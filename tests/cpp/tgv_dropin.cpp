// Taylor-Green driver written against the C++ drop-in classes exactly as a
// reference driver is written against the reference's (3DCompact.cpp:40-110,
// taylorgreen_vortex drivers): Domain, BC, TimeStepping, CSolver_AWS, RK4.
// Usage: tgv_dropin N steps precision(0 FP64 | 1 FP32) outdir [dumpStep
// [restartStep]]: steps is the last step (maxTimeStep); with restartStep
// the run continues from the files a dump at that step left in outdir.
// Prints the final volume-averaged kinetic energy.
#define CS3D_DROPIN_GLOBAL
#include "cs3d.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>

int main(int argc, char **argv)
{
    const int N = argc > 1 ? std::atoi(argv[1]) : 32;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 3;
    const int prec = argc > 3 ? std::atoi(argv[3]) : 0;
    const char *out = argc > 4 ? argv[4] : ".";
    const int dumpStep = argc > 5 ? std::atoi(argv[5]) : 1000000;
    const int restart = argc > 6 ? std::atoi(argv[6]) : -1;
    const double L = 2.0 * M_PI - 2.0 * M_PI / N;
    Domain *dom = new Domain(N, N, N, L, L, L);
    BC *bc = new BC(BC::PERIODIC_SOLVE, BC::PERIODIC, BC::PERIODIC, BC::PERIODIC_SOLVE,
                    BC::PERIODIC, BC::PERIODIC, BC::PERIODIC_SOLVE, BC::PERIODIC, BC::PERIODIC);
    TimeStepping *ts = new TimeStepping(TimeStepping::CONST_CFL, 0.5, steps, 1e30, 1, 1000000,
                                        dumpStep);
    const double mu_ref = (1.0 / 1600.0) / std::pow(100.0, 0.76);
    CSolver_AWS *cs = new CSolver_AWS(dom, bc, ts, 0.49, mu_ref, 8, false);
    cs->ig->Pr = 0.71;
    cs->precision = prec;
    cs->tgvDiagnostics = true;
    cs->outDir = out;
    cs->restartStep = restart;
    // Taylor-Green initial condition (Mach 0.1, isothermal density)
    const double p0 = 1.0 / (1.4 * 0.1 * 0.1);
    for (int k = 0; k < N; k++)
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++) {
                const int ip = k * N * N + j * N + i;
                const double x = dom->x[i], y = dom->y[j], z = dom->z[k];
                cs->U0[ip] = std::sin(x) * std::cos(y) * std::cos(z);
                cs->V0[ip] = -std::cos(x) * std::sin(y) * std::cos(z);
                cs->W0[ip] = 0.0;
                cs->p0[ip] = p0 + (std::cos(2 * x) + std::cos(2 * y)) * (std::cos(2 * z) + 2.0) / 16.0;
                cs->rho0[ip] = cs->p0[ip] / p0;
            }
    AbstractRK *rk = new RK4(cs);
    rk->executeSolverLoop();
    double ke = 0.0;
    cs3d::check(cs3d_kinetic_energy(cs->handle(), &ke), "kinetic_energy");
    std::printf("steps %d time %.17g ke %.17g\n", cs->timeStep, cs->time, ke);
    delete rk;
    delete cs;
    delete ts;
    delete bc;
    delete dom;
    return 0;
}
