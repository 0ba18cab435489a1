// dropin_extras.cpp -- the drop-in's per-line operators, AeroOptics and
// gradient API (include/cs3d.hpp), written against the reference's class
// interface.  usage: dropin_extras <outdir>
//   <outdir>/lines.bin: for dir x, y, z and bc periodic, Dirichlet: the
//     line input (sin + cos mix, length N_dir), calc1stDeriv, calc2ndDeriv,
//     compactFilter of it (FP64)
//   <outdir>/rho.bin and the "<angle>.OPL" files of a standalone AeroOptics
//     on that density
//   <outdir>/grad.bin: J then dJ/dU0 of keGradient over 2 steps of a
//     16^3 Taylor-Green run (with the OPL files of its postStep calls)
#define CS3D_DROPIN_GLOBAL
#include "cs3d.hpp"

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

static void put(FILE *f, const std::vector<double> &v) { std::fwrite(v.data(), sizeof(double), v.size(), f); }

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    const std::string out = argv[1];
    Domain *dom = new Domain(16, 12, 10, 3.0, 2.0, 1.5);
    FILE *fl = std::fopen((out + "/lines.bin").c_str(), "wb");
    for (int d = 0; d < 3; d++)
        for (int bc = 0; bc < 2; bc++) {
            const BC::BCType t = bc == 0 ? BC::PERIODIC_SOLVE : BC::DIRICHLET_SOLVE;
            const Derivatives::Direct dir = (Derivatives::Direct)d;
            Derivatives der(dom, t, dir);
            Filter flt(0.49, dom, t, dir);
            const int n = d == 0 ? dom->Nx : d == 1 ? dom->Ny : dom->Nz;
            std::vector<double> f(n), d1(n), d2(n), ff(n);
            for (int i = 0; i < n; i++) f[i] = std::sin(0.7 * i) + 0.3 * std::cos(1.9 * i + d);
            der.calc1stDeriv(f.data(), d1.data());
            der.calc2ndDeriv(f.data(), d2.data());
            flt.compactFilter(f.data(), ff.data());
            put(fl, f);
            put(fl, d1);
            put(fl, d2);
            put(fl, ff);
        }
    std::fclose(fl);

    // a standalone AeroOptics on a host density
    std::vector<double> rho(dom->N);
    for (int k = 0; k < dom->Nz; k++)
        for (int j = 0; j < dom->Ny; j++)
            for (int i = 0; i < dom->Nx; i++)
                rho[(size_t)k * dom->Nx * dom->Ny + j * dom->Nx + i] =
                    1.0 + 0.1 * std::sin(dom->x[i]) * std::cos(2 * dom->y[j]) + 0.05 * dom->z[k];
    FILE *fr = std::fopen((out + "/rho.bin").c_str(), "wb");
    put(fr, rho);
    std::fclose(fr);
    IdealGas ig(dom, 1.0);
    AeroOptics ao(dom, &ig, rho.data(), nullptr, nullptr, 5, 70.0);
    ao.outDir = out;
    ao.computeAO();

    // the gradient API on a small Taylor-Green run
    const int N = 16;
    const double L = 2 * M_PI - 2 * M_PI / N;
    Domain *d2 = new Domain(N, N, N, L, L, L);
    BC *bcp = new BC(BC::PERIODIC_SOLVE, BC::PERIODIC, BC::PERIODIC, BC::PERIODIC_SOLVE,
                     BC::PERIODIC, BC::PERIODIC, BC::PERIODIC_SOLVE, BC::PERIODIC, BC::PERIODIC);
    TimeStepping *ts = new TimeStepping(TimeStepping::CONST_CFL, 0.5, 100, 100.0, 1, 0, 0);
    CSolver_AWS *cs = new CSolver_AWS(d2, bcp, ts, 0.49, (1.0 / 1600) / std::pow(100.0, 0.76), 8,
                                      false);
    cs->ig->Pr = 0.71;
    cs->outDir = out;
    cs->shearLayerOutput = false;
    const double p0 = 1.0 / (1.4 * 0.01);
    for (int k = 0; k < N; k++)
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++) {
                const size_t ip = (size_t)k * N * N + j * N + i;
                const double x = d2->x[i], y = d2->y[j], z = d2->z[k];
                cs->U0[ip] = std::sin(x) * std::cos(y) * std::cos(z);
                cs->V0[ip] = -std::cos(x) * std::sin(y) * std::cos(z);
                cs->W0[ip] = 0.0;
                cs->p0[ip] = p0 + (std::cos(2 * x) + std::cos(2 * y)) * (std::cos(2 * z) + 2.0) / 16.0;
                cs->rho0[ip] = cs->p0[ip] / p0;
            }
    cs->setRKOrder(4);
    cs->setInitialConditions();
    std::vector<double> gU(d2->N);
    const double J = cs->keGradient(2, 1, gU.data(), nullptr, nullptr);
    FILE *fg = std::fopen((out + "/grad.bin").c_str(), "wb");
    std::fwrite(&J, sizeof(double), 1, fg);
    put(fg, gU);
    std::fclose(fg);
    cs->adjointRelease();
    std::printf("ok J %.17g\n", J);
    delete cs;
    return 0;
}
