// solver.cpp -- host side of the GPU solver behind the C ABI: device state,
// the stage/step orchestration of CSolver_AWS + RK4/TVDRK3
// (CSolver_AWS.cpp:2897-2946, RK4.hpp:120-170, TVDRK3.hpp:98-139), and
// host transfers.  Everything stays on one HIP stream; the only calls that
// wait for the device are the ones that return host data.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cs3d.h"
#include "capi_common.hpp"
#include "plan_dev.hpp"
#include "solver.hpp"
#include "adjoint.hpp"
#include "aero_optics.hpp"
#include "solver_kernels.hpp"

namespace cs3d {

static void hc(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(hip_msg(e, what));
}

// Workgroup size caps of the sweeps and the filter sweeps.  The environment
// variables CS3D_SWEEP_THREADS, CS3D_SWEEP_X_THREADS, CS3D_FILTER_THREADS
// and CS3D_FILTER_X_THREADS set them for tuning; values outside [64, cap]
// are ignored.
static int env_cap(const char *name, int cap, int dflt = 0)
{
    if (dflt <= 0) dflt = cap;
    const char *v = std::getenv(name);
    if (!v) return dflt;
    int x = std::atoi(v);
    return (x >= 64 && x <= cap) ? x : dflt;
}

template <typename real>
struct SolverT : SolverBase {
    // Y/Z sweeps want 16+ lines per workgroup (64-byte rows of x); the X
    // sweep reads whole lines and runs best with the smallest workgroups:
    // one line, or as many as fill one wavefront when a line has fewer than
    // 64 segments (P).  Independent workgroups of one wavefront overlap
    // their load and solve phases on a CU (256^3 FP32: 64 threads 6% faster
    // than 256).  Lines of more than 64 segments take the number of lines
    // that keeps the most lanes busy on a CU: a workgroup occupies whole
    // wavefronts, and a CU holds 8 x-sweep wavefronts in FP32 (4 in FP64:
    // their registers), e.g. 3 lines of 80 segments (4 wavefronts, two
    // workgroups per CU) where one line leaves a wavefront 3/4 idle
    // (320-point lines: x sweep 9% faster in FP32).
    static int x_lines_per_group(int P)
    {
        if (P <= 64) return std::max(1, 64 / std::max(P, 1));
        const int cu_waves = sizeof(real) == 4 ? 8 : 4;
        int best = 1, best_lanes = 0;
        for (int lx = 1; lx * P <= X_SWEEP_THREADS; lx++) {
            const int waves = (lx * P + 63) / 64;
            const int lanes = (cu_waves / waves) * lx * P;
            if (lanes > best_lanes) {
                best = lx;
                best_lanes = lanes;
            }
        }
        return best;
    }
    static int sweep_max_threads(int dir, int P)
    {
        if (dir == 0)
            return env_cap("CS3D_SWEEP_X_THREADS", X_SWEEP_THREADS,
                           std::min(X_SWEEP_THREADS, std::max(64, x_lines_per_group(P) * P)));
        return env_cap("CS3D_SWEEP_THREADS", sizeof(real) == 4 ? 1024 : 512);
    }
    static int filter_max_threads(int dir)
    {
        // rows (x): small workgroups (4 lines of 256 points) overlap their
        // phases best, as for the x sweep (FP64 filter x sweep 10% faster
        // than at 512 threads)
        if (dir == 0) return env_cap("CS3D_FILTER_X_THREADS", 512, 128);
        // columns: FP32 at 1024 lanes (32 lines); FP64 at 256 (8 lines,
        // two workgroups per CU whose phases overlap: 256^3 FP64 filter
        // sweeps 26% faster than one 512-lane workgroup per CU; FP32 at
        // 512 lanes made the filter 7% faster but the x sweep after it 1.4%
        // slower)
        return env_cap("CS3D_FILTER_THREADS", sizeof(real) == 4 ? 1024 : 512,
                       sizeof(real) == 4 ? 1024 : 256);
    }
    // the adjoint's x kernels take one line per workgroup, as the x sweep
    static int adj_x_threads(int P)
    {
        return env_cap("CS3D_ADJ_X_THREADS", ADJ_X_THREADS, std::min(ADJ_X_THREADS, std::max(64, P)));
    }

    cs3d_config cfg;
    int Nx, Ny, Nz;
    long N;
    double dxyz[3];
    PlanDev<real> D1[3], D2[3], FL[3];
    // the x filter on the derivative segmentation, fused into the last
    // stage's x sweep when x lines are uniform and periodic (fx_ok; in FP64
    // only on lines of 64 segments, whose filter carries come from the
    // wavefront scan: through the LDS ring its 4-point plan needs 49 terms;
    // CS3D_NO_FILTER_FUSION in the environment turns it off)
    PlanDev<real> FLX;
    // the forward z sweep's plans with 3-point uniform segments (z3) where the
    // periodic z lines divide by 3 but not by 4; every other use of the z
    // derivatives (standalone operators, adjoint) keeps D1[2], D2[2]
    PlanDev<real> Z31, Z32;
    bool z3 = false;
    const PlanDev<real> &sw1(int d) const { return d == 2 && z3 ? Z31 : D1[d]; }
    const PlanDev<real> &sw2(int d) const { return d == 2 && z3 ? Z32 : D2[d]; }
    bool fx_ok = false;
    LaunchGeom sweepG[3], filterG[3];
    // the column sweeps' geometry for their variants with a smaller
    // workgroup limit (col_small)
    LaunchGeom sweepG_small[3];
    LineMap lm[3];
    Fields<real> F;
    Phys<real> ph;
    StageArgs base;
    hipStream_t st = nullptr;
    void *mem = nullptr;
    StepScalars *dsc = nullptr;
    double *dpart = nullptr, *dout = nullptr;
    double *dfpart = nullptr;  // CFL partials of the last filter sweep
    int nred = 0;
    bool sponge = false;
    bool ic_done = false;
    bool diag_on = false, diag_valid = false;
    bool stale_pending = false;  // a sponge update happened since the last stage 1
    // optional per-launch timing with HIP events on the solver stream
    struct EvPair {
        hipEvent_t a, b;
        int cls;
    };
    bool prof = false;
    std::vector<EvPair> evpool;
    size_t evused = 0;
    int timeStep = 0, filterTimeStep = 0;
    double rho_ref = 1.0, rhoE_ref = 0.0;
    double R_gas = 1.0, cp = 1.0;

    explicit SolverT(const cs3d_config &c) : cfg(c)
    {
        Nx = c.Nx;
        Ny = c.Ny;
        Nz = c.Nz;
        N = (long)Nx * Ny * Nz;
        // kernels address a field with 32-bit byte offsets
        if ((unsigned long long)N * sizeof(real) > 0xFFFFFFFFull)
            throw std::invalid_argument("grid too large: every field must stay below 4 GiB");
        const int dims[3] = {Nx, Ny, Nz};
        const double Ls[3] = {c.Lx, c.Ly, c.Lz};
        for (int d = 0; d < 3; d++) dxyz[d] = (1.0 / ((double)dims[d] - 1.0)) * Ls[d];
        // FP64 only on lines of 64 segments: through the LDS ring its 4-point
        // filter needs 49 carry terms and the fused sweep costs as much as
        // the separate one (measured 256^3 with the ring: FP32 step -1.2%,
        // FP64 +0.2%)
        if (c.bc_type[0] == CS3D_PERIODIC_SOLVE && Nx % SEG_DERIV == 0 &&
            (sizeof(real) == 4 || Nx == 64 * SEG_DERIV) && !std::getenv("CS3D_NO_FILTER_FUSION")) {
            make_plan(FLX, OP_FILTER, 1, Nx, dxyz[0], c.alpha_f, SEG_DERIV);
            fx_ok = FLX.host.uniform && FLX.host.M == SEG_DERIV;
        }
        for (int d = 0; d < 3; d++) {
            // Domain.hpp:30-45
            dxyz[d] = (1.0 / ((double)dims[d] - 1.0)) * Ls[d];
            int per = c.bc_type[d] == CS3D_PERIODIC_SOLVE;
            make_plan(D1[d], OP_D1, per, dims[d], dxyz[d], c.alpha_f);
            make_plan(D2[d], OP_D2, per, dims[d], dxyz[d], c.alpha_f);
            make_plan(FL[d], OP_FILTER, per, dims[d], dxyz[d], c.alpha_f);
            if (D1[d].host.P != D2[d].host.P || D1[d].host.uniform != D2[d].host.uniform)
                throw std::invalid_argument("inconsistent derivative plans");
            if (d == 2 && per && dims[2] % SEG_DERIV != 0 && dims[2] % SEG_DERIV3 == 0 &&
                !std::getenv("CS3D_NO_Z3")) {
                make_plan(Z31, OP_D1, 1, dims[2], dxyz[2], c.alpha_f, SEG_DERIV3);
                make_plan(Z32, OP_D2, 1, dims[2], dxyz[2], c.alpha_f, SEG_DERIV3);
                z3 = Z31.host.uniform && Z32.host.uniform && Z31.host.M == SEG_DERIV3 &&
                     Z31.host.P == Z32.host.P;
            }
            const OpHost &s1 = sw1(d).host, &s2 = sw2(d).host;
            const int P = s1.P;
            int kr = std::max({s1.K, s2.K, kc_deriv<real>()});
            if (d == 0 && fx_ok) {
                fx_ok = D1[0].host.uniform && FLX.host.P == P;
                if (fx_ok && P != 64) kr = std::max({kr, FLX.host.K, kc_filter_x<real>()});
            }
            sweepG[d] = line_geom(P, kr, sweep_max_threads(d, P), sweep_planes<real>(d),
                                  sizeof(real), map_of(d), SWEEP_STAGE_FIELDS * s1.M, d > 0);
            if (sweepG[d].LX < 1) throw std::invalid_argument("line too long for the fused sweeps");
            sweepG_small[d] = sweepG[d];
            if (d > 0) {
                // FP32 table y sweeps at a quarter size: three workgroups of
                // 8 lines per CU keep more lines in flight than one of 16
                // (shear layer 512 x 128 x 42: y sweeps 81 -> 67 us; the
                // table z sweeps, two of 32 lines per CU, stay at half size)
                const int div = sizeof(real) == 4 && d == 1 ? 4 : 2;
                sweepG_small[d] = line_geom(P, kr,
                                            std::min(sweep_max_threads(d, P), sweep_threads<real>() / div),
                                            SWEEP_PLANES, sizeof(real), MAP_COL,
                                            SWEEP_STAGE_FIELDS * s1.M, true);
                if (sweepG_small[d].LX < 1)
                    throw std::invalid_argument("line too long for the fused sweeps");
            }
            // (the compensated filter sweeps run at most 512 lanes; so do
            // the FP32 table filter columns, two workgroups per CU: shear
            // layer 512 x 128 x 42 filter sweeps 8% faster than at 1024)
            const bool fsmall = c.compensated || (d > 0 && !FL[d].host.uniform);
            filterG[d] = line_geom(FL[d].host.P, std::max(FL[d].host.K, kc_filter<real>()),
                                   fsmall ? std::min(filter_max_threads(d), 512)
                                          : filter_max_threads(d),
                                   5, sizeof(real), map_of(d));
            if (filterG[d].LX < 1) throw std::invalid_argument("line too long for the filter");
            lm[d] = xyz_lines(d, Nx, Ny, Nz);
        }
        // IdealGas.hpp:35-47: p_ref = 1/gamma, rho_ref = T_ref = 1
        R_gas = (1.0 / c.gamma) / 1.0 / 1.0;
        cp = R_gas * c.gamma / (c.gamma - 1.0);
        for (int f = 0; f < 6; f++) sponge |= c.bc_kind[f] == CS3D_BC_SPONGE;

        std::memset(&F, 0, sizeof(F));
        // compensated mode (FP32): low parts of *1, *2 and, for TVDRK3, *k
        const bool cmp = c.compensated != 0;
        if (cmp && sizeof(real) != 4)
            throw std::invalid_argument("the compensated mode is an FP32 mode");
        int nf = 15 + 4 + 4 + 10 + 5 + 4 + (sponge ? 15 : 0) + (cmp ? (c.rk == 3 ? 15 : 10) : 0);
        // fields start on 512-byte boundaries (the vector and interleaved
        // accesses need 16/32-byte alignment whatever N is)
        const size_t Np = ((size_t)N + 63) / 64 * 64;
        hc(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
        hc(hipMalloc(&mem, (size_t)nf * Np * sizeof(real)), "hipMalloc(fields)");
        hc(hipMemsetAsync(mem, 0, (size_t)nf * Np * sizeof(real), st), "hipMemset(fields)");
        real *p = (real *)mem;
        auto take = [&](int width = 1) {
            real *r = p;
            p += (size_t)width * Np;
            return r;
        };
        // a conserved set: quads (rho', rhoU, rhoV, rhoW) and rhoE'
        auto take_set = [&]() {
            QSet<real> q5;
            q5.m = take(4);
            q5.e = take();
            return q5;
        };
        F.q = take_set();
        F.qk = F.qkw = take_set();
        F.acc = take_set();
        if (cmp) {
            F.qlo = take_set();
            F.acclo = take_set();
            if (c.rk == 3) F.qklo = take_set();
        }
        // interleaved intermediates: a pair field takes two fields' space,
        // a quad four
        F.dYut = take(2);
        F.dYvw = take(2);
        F.dZuw = take(2);
        F.dZvt = take(2);
        F.Ry4 = take(4);
        F.RyE = take();
        F.Rz4 = take(4);
        F.RzE = take();
        for (int v = 0; v < 5; v++) F.k2[v] = take();
        for (int v = 0; v < 4; v++) F.diag[v] = take();
        if (sponge) {
            F.sigma = take();
            F.avg = take_set();
            F.qstale = take_set();
            for (int v = 0; v < 4; v++) F.mix[v] = take();
        }
        hc(hipMalloc(&dsc, sizeof(StepScalars)), "hipMalloc(scalars)");
        hc(hipMemsetAsync(dsc, 0, sizeof(StepScalars), st), "hipMemset(scalars)");
        nred = (int)std::min<long>(1024, (N + 255) / 256);
        hc(hipMalloc(&dpart, (size_t)nred * RED_SLOTS_HOST * sizeof(double)), "hipMalloc(partials)");
        int nfb = 0;
        for (int d = 0; d < 3; d++) nfb = std::max(nfb, filter_blocks(d));
        hc(hipMalloc(&dfpart, (size_t)nfb * 2 * sizeof(double)), "hipMalloc(partials)");
        hc(hipMalloc(&dout, 2 * RED_SLOTS_HOST * sizeof(double)), "hipMalloc(reductions)");

        std::memset(&base, 0, sizeof(base));
        base.rk = c.rk;
        base.Nx = Nx;
        base.Ny = Ny;
        base.Nz = Nz;
        base.compensated = cmp ? 1 : 0;
        for (int f = 0; f < 6; f++) {
            base.sponge_face[f] = c.bc_kind[f] == CS3D_BC_SPONGE;
            base.wall_face[f] = is_wall(f);
        }
        if (sponge) upload_sigma();
        hc(hipStreamSynchronize(st), "init");
    }

    ~SolverT() override
    {
        if (st) (void)hipStreamSynchronize(st);
        for (auto &e : evpool) {
            (void)hipEventDestroy(e.a);
            (void)hipEventDestroy(e.b);
        }
        for (int d = 0; d < 3; d++) {
            D1[d].release();
            D2[d].release();
            FL[d].release();
        }
        FLX.release();
        Z31.release();
        Z32.release();
        for (int d = 0; d < 3; d++) {
            D1T[d].release();
            D2T[d].release();
            FLT[d].release();
        }
        for (void *p : adjalloc) (void)hipFree(p);
        if (mem) (void)hipFree(mem);
        if (dsc) (void)hipFree(dsc);
        if (dpart) (void)hipFree(dpart);
        if (dfpart) (void)hipFree(dfpart);
        if (dout) (void)hipFree(dout);
        if (st) (void)hipStreamDestroy(st);
    }

    // SpongeBC.hpp:35-154 (sigma starts at zero; see the sponge note there)
    void upload_sigma()
    {
        std::vector<double> x(Nx), y(Ny), z(Nz);
        for (int i = 0; i < Nx; i++) x[i] = ((double)i / ((double)Nx - 1.0)) * cfg.Lx;
        for (int j = 0; j < Ny; j++) y[j] = ((double)j / ((double)Ny - 1.0)) * cfg.Ly;
        for (int k = 0; k < Nz; k++) z[k] = ((double)k / ((double)Nz - 1.0)) * cfg.Lz;
        double LX = cfg.sponge_frac[0] * cfg.Lx, LY = cfg.sponge_frac[1] * cfg.Ly,
               LZ = cfg.sponge_frac[2] * cfg.Lz, st_ = cfg.sponge_strength;
        auto ramp = [&](double s) { return st_ * (0.068 * std::pow(s, 2.0) + 0.845 * std::pow(s, 8.0)); };
        std::vector<real> sg((size_t)N);
        for (int k = 0; k < Nz; k++)
            for (int j = 0; j < Ny; j++)
                for (int i = 0; i < Nx; i++) {
                    double v = 0.0;
                    if (cfg.bc_kind[0] == CS3D_BC_SPONGE && x[i] < LX) v = std::fmax(ramp((LX - x[i]) / LX), v);
                    if (cfg.bc_kind[1] == CS3D_BC_SPONGE && x[i] > cfg.Lx - LX)
                        v = std::fmax(ramp((x[i] - (cfg.Lx - LX)) / LX), v);
                    if (cfg.bc_kind[2] == CS3D_BC_SPONGE && y[j] < LY) v = std::fmax(ramp((LY - y[j]) / LY), v);
                    if (cfg.bc_kind[3] == CS3D_BC_SPONGE && y[j] > cfg.Ly - LY)
                        v = std::fmax(ramp((y[j] - (cfg.Ly - LY)) / LY), v);
                    if (cfg.bc_kind[4] == CS3D_BC_SPONGE && z[k] < LZ) v = std::fmax(ramp((LZ - z[k]) / LZ), v);
                    if (cfg.bc_kind[5] == CS3D_BC_SPONGE && z[k] > cfg.Lz - LZ)
                        v = std::fmax(ramp((z[k] - (cfg.Lz - LZ)) / LZ), v);
                    sg[(size_t)k * Nx * Ny + (size_t)j * Nx + i] = (real)v;
                }
        hc(hipMemcpy(F.sigma, sg.data(), (size_t)N * sizeof(real), hipMemcpyHostToDevice), "sigma");
    }

    void build_phys()
    {
        ph.gm1 = (real)(cfg.gamma - 1.0);
        ph.gamma = (real)cfg.gamma;
        ph.iR = (real)(1.0 / R_gas);
        double p_ref = (cfg.gamma - 1.0) * rhoE_ref;
        double T0s = p_ref / (rho_ref * R_gas);
        ph.rho_ref = (real)rho_ref;
        ph.rhoE_ref = (real)rhoE_ref;
        ph.p_ref = (real)p_ref;
        ph.T0s = (real)T0s;
        ph.RT0s = (real)(R_gas * T0s);
        ph.h_ref = (real)(rhoE_ref + p_ref);
        ph.mu_ref = (real)cfg.mu_ref;
        ph.iTref_ig = (real)1.0;
        ph.cpPr = (real)(cp / cfg.Pr);
        ph.idx = (real)(1.0 / dxyz[0]);
        ph.idy = (real)(1.0 / dxyz[1]);
        ph.idz = (real)(1.0 / dxyz[2]);
    }

    // ---------------------------------------------------------------- IC
    bool is_wall(int f) const
    {
        return cfg.bc_kind[f] == CS3D_BC_ADIABATIC_WALL ||
               cfg.bc_kind[f] == CS3D_BC_MOVING_ADIABATIC_WALL;
    }
    bool any_wall() const
    {
        for (int f = 0; f < 6; f++)
            if (is_wall(f)) return true;
        return false;
    }
    // the wall faces in the reference's order: adiabatic X0..Z1, then
    // moving X0..Z1
    std::vector<int> wall_order() const
    {
        std::vector<int> o;
        for (int kind : {CS3D_BC_ADIABATIC_WALL, CS3D_BC_MOVING_ADIABATIC_WALL})
            for (int f = 0; f < 6; f++)
                if (cfg.bc_kind[f] == kind) o.push_back(f);
        return o;
    }
    // wall velocity of face f: a moving wall's tangential components
    void wall_velocity(int f, double vel[3]) const
    {
        for (int m = 0; m < 3; m++)
            vel[m] = (cfg.bc_kind[f] == CS3D_BC_MOVING_ADIABATIC_WALL && m != f / 2)
                         ? cfg.wall_vel[f][m]
                         : 0.0;
    }

    void set_ic(const double *rho0, const double *U0, const double *V0, const double *W0,
                const double *p0) override
    {
        // setInitialConditions (CSolver_AWS.cpp:207-234)
        std::vector<double> q[5];
        for (int v = 0; v < 5; v++) q[v].resize((size_t)N);
        for (long i = 0; i < N; i++) {
            double rho = rho0[i], U = U0[i], V = V0[i], W = W0[i], p = p0[i];
            q[0][i] = rho;
            q[1][i] = rho * U;
            q[2][i] = rho * V;
            q[3][i] = rho * W;
            q[4][i] = p / (cfg.gamma - 1.0) + 0.5 * rho * (U * U + V * V + W * W);
        }
        if (any_wall()) ic_walls(q, rho0, U0, V0, W0, p0);
        double srho = 0.0, sE = 0.0;
        for (long i = 0; i < N; i++) {
            srho += q[0][i];
            sE += q[4][i];
        }
        rho_ref = srho / (double)N;
        rhoE_ref = sE / (double)N;
        build_phys();
        upload_state(q, 0);
        if (sponge) copy_set(F.avg, F.q);
        StepScalars z;
        std::memset(&z, 0, sizeof(z));
        hc(hipMemcpyAsync(dsc, &z, sizeof(z), hipMemcpyHostToDevice, st), "scalars");
        timeStep = 0;
        filterTimeStep = 0;
        diag_valid = false;
        stale_pending = false;
        hc(launch_reduce<real>(F, ph, N, 0, dpart, nred, dout, dsc, st), "reduce");
        hc(hipStreamSynchronize(st), "set_ic");
        ic_done = true;
    }

    // the wall branches of setInitialConditions (CSolver_AWS.cpp:235-495):
    // on each wall face in turn the wall velocity and T from calcNeumann
    // (Utils.hpp:52) of the six points inward, then p = rho R T and rhoE
    // everywhere from the corrected primitives
    void ic_walls(std::vector<double> *q, const double *rho0, const double *U0, const double *V0,
                  const double *W0, const double *p0)
    {
        std::vector<double> U(U0, U0 + N), V(V0, V0 + N), W(W0, W0 + N), T((size_t)N);
        for (long i = 0; i < N; i++) T[i] = p0[i] / (rho0[i] * R_gas);
        const long sxy = (long)Nx * Ny;
        for (int f : wall_order()) {
            const int dir = f / 2, hi = f % 2;
            double vel[3];
            wall_velocity(f, vel);
            const long stride = dir == 0 ? 1 : dir == 1 ? Nx : sxy;
            const long in = hi ? -stride : stride;
            for (int k = 0; k < Nz; k++)
                for (int j = 0; j < Ny; j++)
                    for (int i = 0; i < Nx; i++) {
                        const int idx = dir == 0 ? i : dir == 1 ? j : k;
                        const int n = dir == 0 ? Nx : dir == 1 ? Ny : Nz;
                        if (idx != (hi ? n - 1 : 0)) continue;
                        const long ip = k * sxy + (long)j * Nx + i;
                        U[ip] = vel[0];
                        V[ip] = vel[1];
                        W[ip] = vel[2];
                        for (int m = 0; m < 3; m++) q[1 + m][ip] = q[0][ip] * vel[m];
                        T[ip] = (T[ip + in] * 360.0 - T[ip + 2 * in] * 450.0 +
                                 T[ip + 3 * in] * 400.0 - T[ip + 4 * in] * 225.0 +
                                 T[ip + 5 * in] * 72.0 - T[ip + 6 * in] * 10.0) /
                                147.0;
                    }
        }
        for (long i = 0; i < N; i++) {
            const double rho = q[0][i], p = rho * R_gas * T[i];
            q[4][i] = p / (cfg.gamma - 1.0) + 0.5 * rho * (U[i] * U[i] + V[i] * V[i] + W[i] * W[i]);
        }
    }

    // preStepBCHandling (CSolver_AWS.cpp:553-826) on the stage state
    void pre_step_walls(const StageArgs &sa)
    {
        for (int f : wall_order()) {
            WallFace w;
            w.dir = f / 2;
            w.hi = f % 2;
            wall_velocity(f, w.vel);
            w.R = R_gas;
            timed(5, [&] { hc(launch_wall_face<real>(F, ph, sa, w, st), "walls"); });
        }
    }

    // q holds full conserved variables (FP64); stored in perturbation form
    void upload_state(const std::vector<double> *q, int which)
    {
        const double *in[5];
        for (int v = 0; v < 5; v++) in[v] = q[v].data();
        set_conserved_raw(which == 0 ? 1 : which == 1 ? 2 : which == 2 ? 3 : 4, in);
    }

    // device copy of a conserved set
    void copy_set(const QSet<real> &dst, const QSet<real> &src)
    {
        hc(hipMemcpyAsync(dst.m, src.m, 4 * (size_t)N * sizeof(real), hipMemcpyDeviceToDevice, st),
           "copy");
        hc(hipMemcpyAsync(dst.e, src.e, (size_t)N * sizeof(real), hipMemcpyDeviceToDevice, st),
           "copy");
    }

    // ---------------------------------------------------------------- timing
    template <typename Fn>
    void timed(int cls, Fn launch)
    {
        if (!prof) {
            launch();
            return;
        }
        if (evused == evpool.size()) {
            EvPair e;
            hc(hipEventCreate(&e.a), "hipEventCreate");
            hc(hipEventCreate(&e.b), "hipEventCreate");
            evpool.push_back(e);
        }
        EvPair &e = evpool[evused++];
        e.cls = cls;
        hc(hipEventRecord(e.a, st), "hipEventRecord");
        launch();
        hc(hipEventRecord(e.b, st), "hipEventRecord");
    }

    void profile_enable(int on) override
    {
        prof = on != 0;
        evused = 0;
    }

    void profile_report(double *ms, int *counts) override
    {
        for (int c = 0; c < PROF_CLASSES; c++) {
            ms[c] = 0.0;
            counts[c] = 0;
        }
        hc(hipStreamSynchronize(st), "sync");
        for (size_t i = 0; i < evused; i++) {
            float t = 0.0f;
            hc(hipEventElapsedTime(&t, evpool[i].a, evpool[i].b), "hipEventElapsedTime");
            ms[evpool[i].cls] += t;
            counts[evpool[i].cls]++;
        }
        evused = 0;
    }

    // ---------------------------------------------------------------- steps
    void pre_step() override
    {
        need_ic();
        timed(5, [&] { hc(launch_dt(dsc, cfg.cfl, cfg.dt_fixed, st), "dt"); });
        timeStep++;
    }

    // fuse: the last stage of a filtering step runs the x filter in its
    // x sweep (fx_ok), writing the x-filtered state to *1
    void sweeps(int rkStep, int mode, int diag, bool fuse = false)
    {
        StageArgs sa = base;
        sa.rkStep = rkStep;
        sa.mode = mode;
        sa.diag = diag;
        sa.stale = (rkStep == 1 && stale_pending) ? 1 : 0;
        if (any_wall()) pre_step_walls(sa);
        // y, then z, then x (the last sweep assembles and updates)
        for (int d : {1, 2, 0})
            timed(d, [&] {
                const OpHost &h = sw1(d).host;
                const bool small = d > 0 && col_small<real>(d, h.uniform, sa.stale);
                const LaunchGeom &g = small ? sweepG_small[d] : sweepG[d];
                hc(launch_sweep<real>(d, sw1(d).dev, sw2(d).dev, (fuse && d == 0) ? FLX.dev : nullptr,
                                      h.uniform, h.M, g, lm[d], F, ph, sa, dsc, st),
                   "sweep");
            });
    }

    // workgroups of a filter sweep along dir
    int filter_blocks(int dir) const { return (lm[dir].nlines + filterG[dir].LX - 1) / filterG[dir].LX; }

    // in_lo/out_lo: the sets' low parts in compensated mode (null m: none)
    void filter_sweep(int dir, const QSet<real> &in, const QSet<real> &out, double *part = nullptr,
                      const QSet<real> &in_lo = QSet<real>{}, const QSet<real> &out_lo = QSet<real>{})
    {
        FiltIO<real> io{};
        io.set = 1;
        io.qin = in;
        io.qout = out;
        io.qin_lo = in_lo;
        io.qout_lo = out_lo;
        filter_launch(dir, io, part);
    }
    // five separate fields (the adjoint's cotangents)
    void filter_sweep(int dir, real *const *in, real *const *out)
    {
        FiltIO<real> io{};
        io.set = 0;
        for (int v = 0; v < 5; v++) {
            io.sep.in[v] = in[v];
            io.sep.out[v] = out[v];
        }
        filter_launch(dir, io, nullptr);
    }
    void filter_launch(int dir, const FiltIO<real> &io, double *part)
    {
        const PlanDev<real> &pl = FL[dir];
        timed(3, [&] {
            hc(launch_filter5<real>(pl.dev, pl.host.uniform, pl.host.M, filterG[dir],
                                    dir == 0 ? MAP_ROW : MAP_COL, lm[dir], io, ph, part, st),
               "filter");
        });
    }

    // filterConservedData (CSolver_AWS.cpp:1682-1893): *2 -> *1.  With
    // cfl set, the last sweep also leaves the CFL partials of its output;
    // returns their number of workgroups (0 when it did not).
    int filter_conserved(bool cfl, bool fused = false)
    {
        if (timeStep % cfg.filter_step == 0) {
            filterTimeStep++;
            int o[3];
            filter_order(filterTimeStep % 3, o, fused);
            if (!fused) filter_sweep(o[0], F.acc, F.q, nullptr, F.acclo, F.qlo);
            filter_sweep(o[1], F.q, F.acc, nullptr, F.qlo, F.acclo);
            // the partials need 2 doubles per lane of the sweep's LDS
            const bool fuse = cfl && filterG[o[2]].shm >= 2 * sizeof(double) *
                                                           (size_t)filterG[o[2]].threads;
            filter_sweep(o[2], F.acc, F.q, fuse ? dfpart : nullptr, F.acclo, F.qlo);
            return fuse ? filter_blocks(o[2]) : 0;
        }
        // the memcpy of *2 into *1 becomes a swap of the two buffers
        std::swap(F.q, F.acc);
        std::swap(F.qlo, F.acclo);
        return 0;
    }

    void sub_step(int rkStep, int rkLast) override
    {
        need_ic();
        int diag = (diag_on && rkLast) ? 1 : 0;
        const bool fuse = fx_ok && rkLast && timeStep % cfg.filter_step == 0;
        sweeps(rkStep, 0, diag, fuse);
        if (rkStep == 1) stale_pending = false;
        if (rkLast) {
            // calcDtFromCFL's maximum for the next step, and the step's
            // diagnostics, from the filtered state before the sponge update;
            // without diagnostics the last filter sweep forms the partials
            const int nb = filter_conserved(!diag, fuse);
            timed(4, [&] {
                if (nb)
                    hc(launch_reduce_cfl(dfpart, nb, dout, dsc, st), "reduce");
                else
                    hc(launch_reduce<real>(F, ph, N, diag, dpart, nred, dout, dsc, st), "reduce");
            });
            diag_valid = diag != 0;
        }
    }

    void post_step() override
    {
        need_ic();
        if (sponge) {
            StageArgs sa = base;
            timed(5, [&] {
                hc(launch_sponge<real>(F, ph, sa, dsc, cfg.sponge_avgT, cfg.sponge_epsP,
                                       1.0 / cfg.gamma, st),
                   "sponge");
            });
            stale_pending = true;
        }
    }

    void step(int nsteps) override
    {
        int ns = cfg.rk == CS3D_RK4 ? 4 : 3;
        for (int n = 0; n < nsteps; n++) {
            pre_step();
            for (int s = 1; s <= ns; s++) sub_step(s, s == ns);
            post_step();
        }
    }

    void rhs(int rkStep) override
    {
        need_ic();
        sweeps(rkStep, 1, 0);
    }

    int check_finite() override
    {
        StepScalars h;
        hc(hipMemcpyAsync(&h, dsc, sizeof(h), hipMemcpyDeviceToHost, st), "scalars");
        hc(hipStreamSynchronize(st), "sync");
        return h.nonfinite ? CS3D_ERR_NONFINITE : CS3D_OK;
    }

    void need_ic() const
    {
        if (!ic_done) throw std::logic_error("initial conditions not set");
    }

    // ---------------------------------------------------------------- host data
    void get_conserved(int set, double *const *out) override
    {
        need_ic();
        hc(hipStreamSynchronize(st), "sync");
        if (set == 3) {
            // dt*RHS: five separate fields, no reference offset
            std::vector<real> buf((size_t)N);
            for (int v = 0; v < 5; v++) {
                if (!out[v]) continue;
                hc(hipMemcpy(buf.data(), F.k2[v], (size_t)N * sizeof(real), hipMemcpyDeviceToHost),
                   "download");
                for (long i = 0; i < N; i++) out[v][i] = (double)buf[i];
            }
            return;
        }
        const QSet<real> &src = set_of(set);
        std::vector<real> m(4 * (size_t)N), e((size_t)N);
        hc(hipMemcpy(m.data(), src.m, m.size() * sizeof(real), hipMemcpyDeviceToHost), "download");
        hc(hipMemcpy(e.data(), src.e, e.size() * sizeof(real), hipMemcpyDeviceToHost), "download");
        // compensated mode: hi + lo
        std::vector<real> ml, el;
        const QSet<real> *lo = lo_of(set);
        if (lo) {
            ml.resize(m.size());
            el.resize(e.size());
            hc(hipMemcpy(ml.data(), lo->m, ml.size() * sizeof(real), hipMemcpyDeviceToHost), "download");
            hc(hipMemcpy(el.data(), lo->e, el.size() * sizeof(real), hipMemcpyDeviceToHost), "download");
        }
        for (int v = 0; v < 4; v++) {
            if (!out[v]) continue;
            const double off = v == 0 ? rho_ref : 0.0;
            for (long i = 0; i < N; i++)
                out[v][i] = (double)m[4 * i + v] + (lo ? (double)ml[4 * i + v] : 0.0) + off;
        }
        if (out[4])
            for (long i = 0; i < N; i++) out[4][i] = (double)e[i] + (lo ? (double)el[i] : 0.0) + rhoE_ref;
    }

    // the low parts of set 1, 2 or 4 in compensated mode, else null
    const QSet<real> *lo_of(int set) const
    {
        const QSet<real> *l = set == 1 ? &F.qlo : set == 2 ? &F.qklo : set == 4 ? &F.acclo : nullptr;
        return (l && l->m) ? l : nullptr;
    }

    // the device fields of a conserved set (1, 2, 4, 5, 6; see cs3d.h)
    const QSet<real> &set_of(int set) const
    {
        if ((set == 5 || set == 6) && !sponge)
            throw std::invalid_argument("sets 5 and 6 exist with sponge faces only");
        if (set < 1 || set > 6 || set == 3) throw std::invalid_argument("bad conserved set");
        return set == 1 ? F.q : set == 2 ? F.qk : set == 4 ? F.acc : set == 5 ? F.avg : F.qstale;
    }

    void set_conserved(int set, const double *const *in) override
    {
        need_ic();
        set_conserved_raw(set, in);
        if (set == 1) hc(launch_reduce<real>(F, ph, N, 0, dpart, nred, dout, dsc, st), "reduce");
    }

    // full conserved values in (FP64), stored in perturbation form
    void set_conserved_raw(int set, const double *const *in)
    {
        hc(hipStreamSynchronize(st), "sync");
        for (int v = 0; v < 5; v++)
            if (!in[v]) throw std::invalid_argument("set_conserved needs all five fields");
        if (set == 3) {
            std::vector<real> buf((size_t)N);
            for (int v = 0; v < 5; v++) {
                for (long i = 0; i < N; i++) buf[i] = (real)in[v][i];
                hc(hipMemcpy(F.k2[v], buf.data(), (size_t)N * sizeof(real), hipMemcpyHostToDevice),
                   "upload");
            }
            return;
        }
        const QSet<real> &dst = set_of(set);
        const QSet<real> *lo = lo_of(set);
        std::vector<real> m(4 * (size_t)N), e((size_t)N), ml, el;
        if (lo) {
            ml.resize(m.size());
            el.resize(e.size());
        }
        // hi = the value rounded, lo = the rounding residual (compensated)
        auto put = [&](double x, real &hi, real *lop) {
            hi = (real)x;
            if (lop) *lop = (real)(x - (double)hi);
        };
        for (long i = 0; i < N; i++) {
            for (int v = 0; v < 4; v++)
                put(in[v][i] - (v == 0 ? rho_ref : 0.0), m[4 * i + v], lo ? &ml[4 * i + v] : nullptr);
            put(in[4][i] - rhoE_ref, e[i], lo ? &el[i] : nullptr);
        }
        hc(hipMemcpy(dst.m, m.data(), m.size() * sizeof(real), hipMemcpyHostToDevice), "upload");
        hc(hipMemcpy(dst.e, e.data(), e.size() * sizeof(real), hipMemcpyHostToDevice), "upload");
        if (lo) {
            hc(hipMemcpy(lo->m, ml.data(), ml.size() * sizeof(real), hipMemcpyHostToDevice), "upload");
            hc(hipMemcpy(lo->e, el.data(), el.size() * sizeof(real), hipMemcpyHostToDevice), "upload");
        }
    }

    // shearLayerInfoCalc (CSolver_AWS.cpp:2259-2456): x-z plane means and
    // second moments per y index, out = 11 rows of Ny values in the order of
    // its files.  Q12: the reference adds v'w' into uwprime and leaves
    // vwprime zero; fix_q12 gives each its own.
    void shear_profiles(int fix_q12, double *out) override
    {
        need_ic();
        const size_t nrow = (size_t)Ny * Nz * 9;
        double *drows = nullptr, *dmean = nullptr;
        hc(hipMalloc(&drows, nrow * sizeof(double)), "hipMalloc(profiles)");
        hc(hipMalloc(&dmean, (size_t)Ny * 6 * sizeof(double)), "hipMalloc(profiles)");
        std::vector<double> rows(nrow), mean((size_t)Ny * 6), mom((size_t)Ny * 9);
        auto planes = [&](int pass, std::vector<double> &dst, int nq) {
            hc(launch_plane_rows<real>(F, ph, Nx, Ny, Nz, pass, dmean, drows, st), "profiles");
            hc(hipMemcpyAsync(rows.data(), drows, nrow * sizeof(double), hipMemcpyDeviceToHost, st),
               "profiles");
            hc(hipStreamSynchronize(st), "sync");
            const double inv = 1.0 / ((double)Nx * Nz);
            for (int j = 0; j < Ny; j++)
                for (int q = 0; q < nq; q++) {
                    double s = 0.0;
                    for (int k = 0; k < Nz; k++) s += rows[((size_t)k * Ny + j) * 9 + q];
                    dst[(size_t)j * nq + q] = s * inv;
                }
        };
        planes(0, mean, 6);
        hc(hipMemcpyAsync(dmean, mean.data(), mean.size() * sizeof(double), hipMemcpyHostToDevice,
                          st),
           "profiles");
        planes(1, mom, 9);
        (void)hipFree(drows);
        (void)hipFree(dmean);
        for (int j = 0; j < Ny; j++) {
            const double *m = &mom[(size_t)j * 9];
            double *o = out;
            for (int q = 0; q < 6; q++) o[q * Ny + j] = m[q];  // rho' u' v' w' p' T' squared
            o[6 * Ny + j] = m[6];                               // u'v'
            o[7 * Ny + j] = fix_q12 ? m[7] : m[7] + m[8];       // u'w' (+ v'w', Q12)
            o[8 * Ny + j] = fix_q12 ? m[8] : 0.0;               // v'w'
            o[9 * Ny + j] = mean[(size_t)j * 6 + 1];            // uAvg
            o[10 * Ny + j] = mean[(size_t)j * 6 + 0];           // rhoAvg
        }
    }

    void shear_profiles_noalloc(int fix_q12, double *out) override
    {
        need_ic();
        if ((long)Ny * Nz > SHEAR_MAX_ROWS || Ny > SHEAR_MAX_NY)
            throw std::invalid_argument("grid too large for the profiles without allocation: Ny * Nz "
                                        "must be at most " + std::to_string(SHEAR_MAX_ROWS) +
                                        " and Ny at most " + std::to_string(SHEAR_MAX_NY));
        hc(shear_profiles_static<real>(F, ph, Nx, Ny, Nz, fix_q12, out, st), "profiles");
    }

    // AeroOptics::computeAO (AeroOptics.cpp:45-93) on the density of set 1
    // (the state) or set 4 (*2, the array the reference passes,
    // CSolver_AWS.hpp:165)
    void aero_optics(int set, int nA, double maxAngle, double *opl) override
    {
        need_ic();
        if (set != 1 && set != 4) throw std::invalid_argument("aero optics: set must be 1 or 4");
        const std::string bad = ao_check(Nx, Ny, Nz, cfg.Lx, cfg.Ly, cfg.Lz, nA, maxAngle);
        if (!bad.empty()) throw std::invalid_argument("aero optics: " + bad);
        AoDomain d{Nx, Ny, Nz, cfg.Lx, cfg.Ly, cfg.Lz};
        AoAngles g;
        ao_geometry(nA, maxAngle, cfg.Ly, std::fmin(dxyz[0], std::fmin(dxyz[1], dxyz[2])), g);
        const size_t nout = (size_t)nA * Nx * Nz;
        double *dout = nullptr;
        hc(hipMalloc(&dout, nout * sizeof(double)), "hipMalloc(optical paths)");
        hipError_t e = launch_opl_set<real>((set == 1 ? F.q : F.acc).m, rho_ref, d, g, dout, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(opl, dout, nout * sizeof(double), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        (void)hipFree(dout);
        hc(e, "optical paths");
    }

    void get_run_state(cs3d_run_state *rs) override
    {
        need_ic();
        StepScalars h;
        hc(hipMemcpyAsync(&h, dsc, sizeof(h), hipMemcpyDeviceToHost, st), "scalars");
        hc(hipStreamSynchronize(st), "sync");
        rs->timeStep = timeStep;
        rs->filterTimeStep = filterTimeStep;
        rs->time = h.time;
        rs->dt = h.dt;
        rs->cfl_max = h.cfl_max;
        rs->rho_ref = rho_ref;
        rs->rhoE_ref = rhoE_ref;
        rs->stale = stale_pending ? 1 : 0;
    }

    void set_run_state(const cs3d_run_state &rs) override
    {
        need_ic();
        if (rs.timeStep < 0 || rs.filterTimeStep < 0 || !(rs.rho_ref > 0.0) || !(rs.dt >= 0.0) ||
            !(rs.cfl_max >= 0.0))
            throw std::invalid_argument("set_run_state: bad run state");
        if (rs.stale && !sponge) throw std::invalid_argument("set_run_state: stale planes without sponges");
        StepScalars h;
        hc(hipMemcpyAsync(&h, dsc, sizeof(h), hipMemcpyDeviceToHost, st), "scalars");
        hc(hipStreamSynchronize(st), "sync");
        h.timeStep = rs.timeStep;
        h.time = rs.time;
        h.dt = rs.dt;
        if (rs.cfl_max > 0.0) h.cfl_max = rs.cfl_max;
        hc(hipMemcpyAsync(dsc, &h, sizeof(h), hipMemcpyHostToDevice, st), "scalars");
        timeStep = rs.timeStep;
        filterTimeStep = rs.filterTimeStep;
        stale_pending = rs.stale != 0;
        rho_ref = rs.rho_ref;
        rhoE_ref = rs.rhoE_ref;
        build_phys();
        diag_valid = false;
        hc(hipStreamSynchronize(st), "sync");
    }

    void get_primitives(double *const *out) override
    {
        need_ic();
        std::vector<double> q[5];
        double *o[5];
        for (int v = 0; v < 5; v++) {
            q[v].resize((size_t)N);
            o[v] = q[v].data();
        }
        get_conserved(1, o);
        // updateNonConservedData (CSolver_AWS.cpp:1895-1935)
        for (long i = 0; i < N; i++) {
            double rho = q[0][i], U = q[1][i] / rho, V = q[2][i] / rho, W = q[3][i] / rho;
            double p = (cfg.gamma - 1.0) * (q[4][i] - 0.5 * rho * (U * U + V * V + W * W));
            double T = p / (rho * R_gas);
            double mu = cfg.mu_ref * std::pow(T / 1.0, 0.76);
            double amu = 0.76 * (cfg.mu_ref / std::pow(1.0, 0.76)) * std::pow(T, 0.76 - 1.0);
            double sos = std::sqrt(cfg.gamma * p / rho);
            const double vals[8] = {U, V, W, p, T, mu, amu, sos};
            for (int k = 0; k < 8; k++)
                if (out[k]) out[k][i] = vals[k];
        }
    }

    void get_time(double *time, double *dt, int *ts) override
    {
        StepScalars h;
        hc(hipMemcpyAsync(&h, dsc, sizeof(h), hipMemcpyDeviceToHost, st), "scalars");
        hc(hipStreamSynchronize(st), "sync");
        if (time) *time = h.time;
        if (dt) *dt = h.dt;
        if (ts) *ts = h.timeStep;
    }

    void set_diag(int on) override { diag_on = on != 0; }

    int diag_tgv(double out[7]) override
    {
        if (!diag_valid) return CS3D_ERR_STATE;
        double r[RED_SLOTS_HOST];
        StepScalars h;
        hc(hipMemcpyAsync(r, dout, sizeof(r), hipMemcpyDeviceToHost, st), "diag");
        hc(hipMemcpyAsync(&h, dsc, sizeof(h), hipMemcpyDeviceToHost, st), "scalars");
        hc(hipStreamSynchronize(st), "sync");
        out[0] = h.time;
        out[1] = r[1];
        out[2] = r[2];
        out[3] = r[3];
        out[4] = r[4];
        out[5] = r[6] / (double)N;
        out[6] = std::sqrt(r[5] / (double)N);
        return h.nonfinite ? CS3D_ERR_NONFINITE : CS3D_OK;
    }

    // calcTurbulenceQuantities (CSolver_AWS.cpp:2033-2173) from the post-step
    // sums; the density variance uses the stored perturbation rho' (no
    // cancellation against rho_ref)
    int diag_turb(double out[11]) override
    {
        if (!diag_valid) return CS3D_ERR_STATE;
        double r[RED_SLOTS_HOST];
        StepScalars h;
        hc(hipMemcpyAsync(r, dout, sizeof(r), hipMemcpyDeviceToHost, st), "diag");
        hc(hipMemcpyAsync(&h, dsc, sizeof(h), hipMemcpyDeviceToHost, st), "scalars");
        hc(hipStreamSynchronize(st), "sync");
        const double NN = (double)N;
        const double mRp = r[8] / NN;
        const double mRho = rho_ref + mRp;
        const double rhoprime = std::sqrt(std::fmax(r[9] / NN - mRp * mRp, 0.0));
        const double mMu = r[6] / NN;
        const double mTD = (r[2] / NN) / mRho;
        const double up = std::sqrt(r[10] / NN / 3.0);
        const double uip = std::sqrt(r[11] / NN);
        const double urms = std::sqrt(r[10] / NN);
        const double mSOS = r[12] / NN;
        const double mNu = mMu / mRho;
        const double taylorMicro2 = up / uip;
        const double k0 = 8.0;  // spectrum peak of the initial field (:2130)
        out[0] = h.time;
        out[1] = std::sqrt(2.0 * M_PI) / k0 / up;           // tau
        out[2] = up * taylorMicro2 / mNu;                   // taylorReyn3
        out[3] = urms / mSOS;                               // turbMach
        out[4] = mTD;                                       // meanTurbDiss
        out[5] = up;                                        // uprime
        out[6] = std::pow(std::pow(mNu, 3.0) / mTD, 0.25);  // kolNu
        out[7] = rhoprime;
        out[8] = std::sqrt(r[5] / NN);                      // dilprime
        out[9] = std::sqrt(r[13] / NN);                     // vortprime
        out[10] = r[1] / NN;                                // meanKineticEng
        return h.nonfinite ? CS3D_ERR_NONFINITE : CS3D_OK;
    }

    int kinetic_energy(double *ke) override
    {
        need_ic();
        double r[RED_SLOTS_HOST];
        hc(launch_reduce<real>(F, ph, N, 2, dpart, nred, dout + RED_SLOTS_HOST, dsc, st), "reduce");
        hc(hipMemcpyAsync(r, dout + RED_SLOTS_HOST, sizeof(r), hipMemcpyDeviceToHost, st), "ke");
        hc(hipStreamSynchronize(st), "sync");
        *ke = r[1] / (double)N;
        return std::isfinite(*ke) ? CS3D_OK : CS3D_ERR_NONFINITE;
    }

    // ------------------------------------------------------------------
    // Discrete adjoint.  The transpose of the discrete stage map.  On
    // periodic grids without sponges or walls, five line passes per
    // vector-Jacobian product (adjoint.hpp): the y and z derivative fields of the stage
    // state recomputed, the x derivatives with the cotangent seeds and all
    // x-direction transposed terms, then the z and y transposed terms
    // (periodic D1 is skew-adjoint, D2 and the filter self-adjoint), each
    // pass chaining its share back to the conserved state.  RK4 and TVDRK3
    // are reversed stage by stage with the forward dt sequence frozen; the stage states of
    // each checkpoint segment are replayed with the fused forward stages and
    // kept.  Cotangents are conserved sets (a quad and the energy).  Other
    // grids (non-periodic lines, sponges, walls) take the general path
    // below (vjp_general).

    std::vector<void *> adjalloc;
    bool adj_ready = false;
    AdjRec<real> arec{};
    AdjSeeds<real> aseed{};
    QSet<real> aresx{}, aresxz{};
    QSet<real> lamb{}, la5{}, sc5{}, vb5{}, vb5b{}, qb5{};
    real *gvel[3] = {};
    LaunchGeom adjG[5];
    std::vector<QSet<real>> adjpool;  // gradient states kept across calls

    real *adj_alloc_field(int width = 1)
    {
        void *p = nullptr;
        hc(hipMalloc(&p, (size_t)width * N * sizeof(real)), "hipMalloc(adjoint)");
        adjalloc.push_back(p);
        return (real *)p;
    }
    QSet<real> adj_alloc_set()
    {
        QSet<real> a;
        a.m = adj_alloc_field(4);
        a.e = adj_alloc_field();
        return a;
    }

    static int adj_pass_dir(int pass)
    {
        return pass == ADJ_X ? 0 : (pass == ADJ_ZREC || pass == ADJ_Z) ? 2 : 1;
    }

    // General path (non-periodic lines or sponge faces; CS3D_ADJ_GENERAL in
    // the environment takes it on periodic grids too): the transposed
    // operators' plans, the work fields of one product (GenAdj), the
    // cotangents of the sponge averages and of the stale planes, and per
    // step of a segment the state on the stale planes, the averages and the
    // filtered state before the sponge update.
    bool adj_general = false;
    PlanDev<real> D1T[3], D2T[3], FLT[3];
    GenAdj<real> gadj{};
    QSet<real> gl_avg{}, gl_stale{};
    const QSet<real> *gcur_S = nullptr;
    QSet<real> gw_q{};
    std::vector<QSet<real>> gw_keep;
    // walls beside sponges: the corrected pre-reset set of a stale first
    // stage and that set before each face
    QSet<real> gw_qS{};
    std::vector<QSet<real>> gw_keepS;

    void adj_require()
    {
        need_ic();
        bool general = sponge || any_wall() || std::getenv("CS3D_ADJ_GENERAL") != nullptr;
        for (int d = 0; d < 3; d++) general |= cfg.bc_type[d] != CS3D_PERIODIC_SOLVE;
        if (adj_ready) {
            if (general != adj_general) throw std::logic_error("adjoint path changed");
            return;
        }
        adj_general = general;
        if (general) {
            // the transposed operators address quad components with 32-bit
            // byte offsets
            if ((unsigned long long)N * 4 * sizeof(real) > 0xFFFFFFFFull)
                throw std::invalid_argument("grid too large for the general adjoint path");
            const int dims[3] = {Nx, Ny, Nz};
            for (int d = 0; d < 3; d++) {
                const int per = cfg.bc_type[d] == CS3D_PERIODIC_SOLVE;
                // z on 3-point segments where the forward z sweep has them
                const int seg = d == 2 && z3 ? SEG_DERIV3 : 0;
                make_plan(D1T[d], OP_D1, per, dims[d], dxyz[d], cfg.alpha_f, seg, true);
                make_plan(D2T[d], OP_D2, per, dims[d], dxyz[d], cfg.alpha_f, seg, true);
                make_plan(FLT[d], OP_FILTER, per, dims[d], dxyz[d], cfg.alpha_f, 0, true);
            }
            GenAdj<real> &g = gadj;
            for (int k = 0; k < 4; k++) g.pY[k] = adj_alloc_field();
            for (int k = 0; k < 4; k++) g.pX[k] = sponge ? adj_alloc_field() : g.pY[k];
            for (int k = 0; k < 12; k++) g.d1[k] = adj_alloc_field();
            for (int k = 0; k < 12; k++) g.d2[k] = adj_alloc_field();
            for (int k = 0; k < 6; k++) g.mx[k] = adj_alloc_field();
            for (int k = 0; k < 4; k++) g.s[k] = adj_alloc_field();
            for (int k = 0; k < G1_COUNT; k++) g.g1[k] = adj_alloc_field();
            for (int k = 0; k < 5; k++) g.vs[k] = adj_alloc_field();
            for (int d = 0; d < 3; d++) g.vc[d] = any_wall() ? adj_alloc_field() : g.vs[0];
            for (int k = 0; k < 4; k++) g.bY[k] = adj_alloc_field();
            for (int k = 0; k < 4; k++) g.bX[k] = adj_alloc_field();
            for (int d = 0; d < 3; d++)
                for (int c = 0; c < 5; c++) g.F[d][c] = adj_alloc_field();
            g.sa = base;
            g.sigma = sponge ? F.sigma : nullptr;
            if (sponge) {
                gl_avg = adj_alloc_set();
                gl_stale = adj_alloc_set();
            }
            // walls: the corrected stage state and the state before each face
            if (any_wall()) {
                gw_q = adj_alloc_set();
                gw_keep.clear();
                for (size_t f = 0; f < wall_order().size(); f++) gw_keep.push_back(adj_alloc_set());
                if (sponge) {
                    gw_qS = adj_alloc_set();
                    gw_keepS.clear();
                    for (size_t f = 0; f < wall_order().size(); f++)
                        gw_keepS.push_back(adj_alloc_set());
                }
            }
        } else {
            for (int pass = ADJ_YREC; pass <= ADJ_Y; pass++) {
                const int dir = adj_pass_dir(pass);
                const OpHost &h1 = D1[dir].host, &h2 = D2[dir].host;
                if (h1.P != h2.P || h1.M != h2.M || h1.uniform != h2.uniform)
                    throw std::logic_error("derivative plans of one direction differ in segmentation");
                const int kr = std::max(std::max(h1.K, h2.K), kc_deriv<real>());
                LaunchGeom &g = adjG[pass];
                if (pass == ADJ_X) {
                    g = line_geom(h1.P, kr, adj_x_threads(h1.P), ADJ_X_PLANES, sizeof(real), MAP_ROW);
                    while (g.LX > 0 && adj_x_lds(h1.P, kr, g.LX, h1.M, sizeof(real)) > LDS_BYTES) g.LX--;
                    g.threads = g.LX * h1.P;
                    g.shm = g.LX > 0 ? adj_x_lds(h1.P, kr, g.LX, h1.M, sizeof(real)) : 0;
                }
                else if (pass == ADJ_YREC || pass == ADJ_ZREC)
                    g = line_geom(h1.P, kr,
                                  env_cap("CS3D_ADJ_THREADS", adj_rec_threads<real>(h1.uniform != 0)),
                                  ADJ_R_PLANES, sizeof(real), MAP_COL);
                else
                    g = line_geom(h1.P, kr, env_cap("CS3D_ADJ_THREADS", adj_col_threads<real>()),
                                  ADJ_T_PLANES, sizeof(real), MAP_COL);
                if (g.LX < 1) throw std::invalid_argument("line too long for the adjoint line kernels");
            }
            arec.dYut = adj_alloc_field(2);
            arec.dYvw = adj_alloc_field(2);
            arec.dY2 = adj_alloc_field(4);
            arec.dZ1 = adj_alloc_field(4);
            arec.dZ2 = adj_alloc_field(4);
            arec.mYZ = adj_alloc_field(2);
            aseed.sq = adj_alloc_field(4);
            aseed.gz = adj_alloc_field(4);
            aseed.gy = adj_alloc_field(4);
            aseed.gs = adj_alloc_field(2);
            aseed.mzz = adj_alloc_field(2);
            aresx = adj_alloc_set();
            aresxz = adj_alloc_set();
        }
        QSet<real> *sets[6] = {&lamb, &la5, &sc5, &vb5, &vb5b, &qb5};
        for (auto p : sets) *p = adj_alloc_set();
        for (auto &p : gvel) p = adj_alloc_field();
        adj_ready = true;
    }

    int map_of(int d) const { return d == 0 ? MAP_ROW : MAP_COL; }

    // out (+)= coef op^T(in) along direction d (op: a transposed plan)
    void opT(PlanDev<real> &pl, int d, const real *in, real *out, real coef, int acc)
    {
        hc(launch_field_opT<real>(pl, map_of(d), lm[d], in, out, coef, acc, st), "transposed operator");
    }
    void opF(PlanDev<real> &pl, int d, const real *in, real *out)
    {
        hc(launch_field_op<real>(pl, map_of(d), lm[d], in, out, st), "operator");
    }

    // The general-path product (adjoint.hpp: GenAdj): stale marks the first
    // stage after a sponge update, whose x block read the stale planes
    // gcur_S.  (The product needs no averages: the sponge source is linear
    // in the state, with slope -sigma.)
    void vjp_general(const QSet<real> &Q, const QSet<real> &v, const AdjVjp<real> &a, bool stale)
    {
        GenAdj<real> g = gadj;
        g.q = Q;
        g.v = v;
        g.cv = a.cv;
        g.stale = stale ? 1 : 0;
        if (stale) {
            if (!gcur_S) throw std::logic_error("adjoint: stale stage without its planes");
            g.qstale = *gcur_S;
            g.res_stale = gl_stale;
        }
        if (sponge) g.resA = gl_avg;
        hc(launch_gadj<real>(GADJ_PRIMS, g, a, ph, N, st), "adjoint prims");
        // derivative fields, one batched launch per direction: x derivatives
        // of the x block's set; then y, with the mixed Uxy = D_y Ux and
        // Vxy = D_y Vx; then z, with Uxz, Vyz, Wxz, Wyz = D_z of Ux, Vy, Wx, Wy
        {
            OpBatch<real> b{};
            for (int f = 0; f < 4; f++) {
                b.in[f] = g.pX[f];
                b.out[f] = g.d1[3 * f];
                b.out[4 + f] = g.d2[3 * f];
            }
            hc(launch_ops_fwd<real>(D1[0], D2[0], map_of(0), lm[0], 4, 4, b, st), "x derivatives");
            b = OpBatch<real>{};
            for (int f = 0; f < 4; f++) {
                b.in[f] = g.pY[f];
                b.out[f] = g.d1[3 * f + 1];
                b.out[6 + f] = g.d2[3 * f + 1];
            }
            b.in[4] = g.d1[0];
            b.out[4] = g.mx[0];
            b.in[5] = g.d1[3];
            b.out[5] = g.mx[2];
            hc(launch_ops_fwd<real>(D1[1], D2[1], map_of(1), lm[1], 6, 4, b, st), "y derivatives");
            b = OpBatch<real>{};
            for (int f = 0; f < 4; f++) {
                b.in[f] = g.pY[f];
                b.out[f] = g.d1[3 * f + 2];
                b.out[8 + f] = g.d2[3 * f + 2];
            }
            const int src[4] = {0, 4, 6, 7}, dst[4] = {1, 3, 4, 5};
            for (int k = 0; k < 4; k++) {
                b.in[4 + k] = g.d1[src[k]];
                b.out[4 + k] = g.mx[dst[k]];
            }
            hc(launch_ops_fwd<real>(sw1(2), sw2(2), map_of(2), lm[2], 8, 4, b, st), "z derivatives");
        }
        hc(launch_gadj<real>(GADJ_SEEDS, g, a, ph, N, st), "adjoint seeds");
        // transposed terms, two batched launches per direction, z, then y,
        // then x: the primitive cotangents (b += D1^T of a first-derivative
        // seed + D2^T of a second-derivative one), then the mixed terms and
        // the flux cotangents.  A shared first-derivative seed (g1XY for Uy
        // and Vx, g1XZ for Uz and Wx, g1YZ for Vz and Wy) is read along its
        // first direction before the mixed terms of its other use are added
        // to it.
        const real c13 = real(1) / real(3), c43 = real(4) / real(3);
        real *const *G = g.g1;
        real *s0 = g.s[0], *s1 = g.s[1], *s2 = g.s[2], *gq = g.s[3];
        // pair launch: out[f] += D1^T(a[f]) + ca[f] D2^T(e[f])
        auto pair = [&](int d, real *const *out, real *const (&a)[4], real *const (&e)[4],
                        const real (&ca)[4]) {
            OpBatch<real> b{};
            for (int f = 0; f < 4; f++) {
                b.in[f] = a[f];
                b.coef[f] = 1;
                b.in[4 + f] = e[f];
                b.coef[4 + f] = ca[f];
                b.out[f] = out[f];
                b.acc[f] = 1;
            }
            hc(launch_ops_T<real>(D1T[d], D2T[d], map_of(d), lm[d], 4, 4, b, st),
               "transposed terms");
        };
        // list launch: the mixed terms (+= c13 D1^T), then F[d][c] = -D1^T(vs)
        auto list = [&](int d, int nm, real *const *min, real *const *mout) {
            OpBatch<real> b{};
            for (int k = 0; k < nm; k++) {
                b.in[k] = min[k];
                b.out[k] = mout[k];
                b.coef[k] = c13;
                b.acc[k] = 1;
            }
            for (int c = 0; c < 5; c++) {
                b.in[nm + c] = c ? g.vs[c] : g.vc[d];
                b.out[nm + c] = g.F[d][c];
                b.coef[nm + c] = -1;
                b.acc[nm + c] = 0;
            }
            hc(launch_ops_T<real>(D1T[d], D2T[d], map_of(d), lm[d], nm + 5, 0, b, st),
               "transposed terms");
        };
        {
            real *const a[4] = {G[G1_XZ], G[G1_YZ], G[G1_WZ], G[G1_TZ]};
            real *const e[4] = {s0, s1, s2, gq};
            const real ca[4] = {1, 1, c43, 1};
            pair(2, g.bY, a, e, ca);
            // Vyz, Uxz, Wxz, Wyz
            real *const mi[4] = {s2, s2, s0, s1};
            real *const mo[4] = {G[G1_VY], G[G1_UX], G[G1_XZ], G[G1_YZ]};
            list(2, 4, mi, mo);
        }
        {
            real *const a[4] = {G[G1_XY], G[G1_VY], G[G1_YZ], G[G1_TY]};
            real *const e[4] = {s0, s1, s2, gq};
            const real ca[4] = {1, c43, 1, 1};
            pair(1, g.bY, a, e, ca);
            // Uxy, Vxy
            real *const mi[2] = {s1, s0};
            real *const mo[2] = {G[G1_UX], G[G1_XY]};
            list(1, 2, mi, mo);
        }
        {
            real *const a[4] = {G[G1_UX], G[G1_XY], G[G1_XZ], G[G1_TX]};
            real *const e[4] = {s0, s1, s2, gq};
            const real ca[4] = {c43, 1, 1, 1};
            pair(0, g.bX, a, e, ca);
            list(0, 0, nullptr, nullptr);
        }
        hc(launch_gadj<real>(GADJ_CHAIN, g, a, ph, N, st), "adjoint chain");
    }

    // F^T along d of every field of a set: the quads' components as strided
    // lines (in != out)
    void filter_set_T(int d, const QSet<real> &in, const QSet<real> &out)
    {
        LineMap l4 = lm[d];
        l4.s1 *= 4;
        l4.s2 *= 4;
        l4.sp *= 4;
        for (int k = 0; k < 4; k++)
            hc(launch_field_opT<real>(FLT[d], map_of(d), l4, in.m + k, out.m + k, 1, 0, st),
               "transposed filter");
        opT(FLT[d], d, in.e, out.e, 1, 0);
    }

    // The cotangent res = (d(dt R)/dQ)^T lam of the stage map at state Q,
    // with v = dt * lam given; a holds the epilogue (AdjVjp).
    // preStepBCHandling on a copy: dst = W(src), the state before each face
    // kept in gw_keep for wall_vjp.  srcS (a stale first stage): the
    // pre-reset set is corrected with it into gw_qS, as the forward's first
    // stage after a sponge update does (the wall temperatures from its
    // sponge-plane values, the face values written to it there), and kept
    // before each face in gw_keepS
    void wall_apply(const QSet<real> &src, const QSet<real> &dst, const QSet<real> *srcS = nullptr)
    {
        copy_set(dst, src);
        Fields<real> Fw = F;
        Fw.q = dst;
        Fw.qlo = QSet<real>{};
        StageArgs sa = base;
        sa.rkStep = 1;
        sa.stale = 0;
        if (srcS) {
            copy_set(gw_qS, *srcS);
            Fw.qstale = gw_qS;
            sa.stale = 1;
        }
        const std::vector<int> order = wall_order();
        for (size_t i = 0; i < order.size(); i++) {
            copy_set(gw_keep[i], dst);
            if (srcS) copy_set(gw_keepS[i], gw_qS);
            hc(launch_wall_face<real>(Fw, ph, sa, wall_face_of(order[i]), st), "walls");
        }
    }
    // lam (in place): from the cotangent of W(src) to that of src, for the
    // src of the last wall_apply; lamS (with a stale wall_apply): the
    // pre-reset set's cotangent likewise, in place
    void wall_vjp(const QSet<real> &lam, const QSet<real> *lamS = nullptr)
    {
        const std::vector<int> order = wall_order();
        for (size_t i = order.size(); i-- > 0;)
            hc(launch_wall_face_vjp<real>(gw_keep[i], lam, lamS ? gw_keepS[i] : QSet<real>{},
                                          lamS ? *lamS : QSet<real>{}, ph, base,
                                          wall_face_of(order[i]), st),
               "walls reverse");
    }
    // the product at the corrected state of the last wall_apply; stale: its
    // x block read the corrected pre-reset set gw_qS, whose share of the
    // cotangent goes to gl_stale
    void vjp_corrected(const QSet<real> &v, const AdjVjp<real> &f, bool stale)
    {
        const QSet<real> *keep = gcur_S;
        if (stale) gcur_S = &gw_qS;
        vjp(gw_q, v, f, stale);
        gcur_S = keep;
    }
    WallFace wall_face_of(int f) const
    {
        WallFace w;
        w.dir = f / 2;
        w.hi = f % 2;
        wall_velocity(f, w.vel);
        w.R = R_gas;
        return w;
    }

    // The product J^T(W(Q)) (cv v) into out, then taken back through the
    // wall correction: (d(dt R(W(Q)))/dQ)^T v.  QS (a stale first stage):
    // the pre-reset set, corrected with Q and held fixed (its cotangent is
    // left in gl_stale)
    void vjp_walls(const QSet<real> &Q, const QSet<real> &v, real cv, const QSet<real> &out,
                   const QSet<real> *QS = nullptr)
    {
        wall_apply(Q, gw_q, QS);
        AdjVjp<real> f{};
        f.cv = cv;
        f.out = out;
        vjp_corrected(v, f, QS != nullptr);
        wall_vjp(out, QS ? &gl_stale : nullptr);
    }

    void vjp(const QSet<real> &Q, const QSet<real> &v, AdjVjp<real> a, bool stale = false)
    {
        if (a.vn.m && (a.vn.m == v.m || a.vn.e == v.e))
            throw std::logic_error("adjoint: next cotangent aliases v");
        if (adj_general) {
            vjp_general(Q, v, a, stale);
            return;
        }
        if (stale) throw std::logic_error("adjoint: a stale stage on the periodic path");
        a.q = Q;
        a.v = v;
        a.rec = arec;
        a.sd = aseed;
        a.resx = aresx;
        a.resxz = aresxz;
        for (int pass = ADJ_YREC; pass <= ADJ_Y; pass++) {
            const int dir = adj_pass_dir(pass);
            hc(launch_adj_pass<real>(pass, D1[dir].dev, D2[dir].dev, D1[dir].host.uniform,
                                     D1[dir].host.M, adjG[pass], lm[dir], a, ph, st),
               "adjoint pass");
        }
    }

    // the direction order filter_conserved uses for filter count m
    // filterConservedData's round robin (filter count m: X Y Z, Y Z X, Z X Y);
    // fused: x first (in the last x sweep), the other two in their order.
    // Filters along different axes commute, so the orders differ by
    // rounding only.
    static void filter_order(int m, int o[3], bool fused = false)
    {
        if (m == 1) { o[0] = 0; o[1] = 1; o[2] = 2; }
        else if (m == 2) { o[0] = 1; o[1] = 2; o[2] = 0; }
        else { o[0] = 2; o[1] = 0; o[2] = 1; }
        if (fused) {
            int r[2], n = 0;
            for (int i = 0; i < 3; i++)
                if (o[i] != 0) r[n++] = o[i];
            o[0] = 0;
            o[1] = r[0];
            o[2] = r[1];
        }
    }

    // the transposed filter of filter count m: the three sweeps in reverse
    // order (each periodic line filter is symmetric)
    void filter_transposed(const QSet<real> &in, const QSet<real> &out, const QSet<real> &scratch,
                           int m)
    {
        int o[3];
        filter_order(m, o, fx_ok);
        if (adj_general) {
            filter_set_T(o[2], in, out);
            filter_set_T(o[1], out, scratch);
            filter_set_T(o[0], scratch, out);
            return;
        }
        filter_sweep(o[2], in, out);
        filter_sweep(o[1], out, scratch);
        filter_sweep(o[0], scratch, out);
    }

    // stages per step of the configured scheme
    int nstages() const { return cfg.rk == CS3D_TVDRK3 ? 3 : 4; }

    // one step on the fused forward path from F.q; stg (when given) receives
    // the stage states q, Q2, Q3 (and Q4 for RK4).  Replays (dt > 0)
    // restore the step counters the filter round robin reads.
    struct StepRec {
        double dt;
        int timeStep, filterTimeStep, order;  // counters before the step; order -1: no filter
        int stale;  // the first stage read the stale sponge planes
    };

    // srec (sponges): receives the stale planes and the averages at the
    // step's start and the filtered state before its sponge update
    StepRec fused_step(const StepRec *replay, const QSet<real> *stg,
                       const QSet<real> *srec = nullptr)
    {
        StepRec r;
        if (replay) {
            r = *replay;
            timeStep = r.timeStep + 1;
            filterTimeStep = r.filterTimeStep;
            stale_pending = r.stale != 0;
            hc(launch_dt(dsc, cfg.cfl, r.dt, st), "dt");
        } else {
            r.timeStep = timeStep;
            r.filterTimeStep = filterTimeStep;
            r.stale = stale_pending ? 1 : 0;
            pre_step();
            StepScalars hs;
            hc(hipMemcpyAsync(&hs, dsc, sizeof(hs), hipMemcpyDeviceToHost, st), "scalars");
            hc(hipStreamSynchronize(st), "sync");
            r.dt = hs.dt;
            r.order = timeStep % cfg.filter_step == 0 ? (filterTimeStep + 1) % 3 : -1;
        }
        if (stg) copy_set(stg[0], F.q);
        if (srec) {
            copy_set(srec[0], F.qstale);
            copy_set(srec[1], F.avg);
        }
        // the last sweep of stage s < ns writes stage state s + 1 straight
        // into the store, and stage s + 1 reads it there
        const QSet<real> own = F.qk;
        const int ns = nstages();
        for (int s = 1; s <= ns; s++) {
            if (stg && s < ns) F.qkw = stg[s];
            sub_step(s, s == ns);
            if (stg && s < ns) F.qk = stg[s];
        }
        F.qk = F.qkw = own;
        if (srec) copy_set(srec[2], F.q);
        post_step();
        return r;
    }

    // One step back with walls (the general path): every stage state was
    // wall-corrected before its right-hand side, the first one in place
    // (the RK combination builds on it).  RK4: with la the cotangent of
    // the unfiltered result, r4 = W4^T J4^T (la/6), r3 = W3^T J3^T (la/3 +
    // r4), r2 = W2^T J2^T (la/3 + r3/2), and lam = W1^T (la + r4 + r3 + r2 +
    // J1^T (la/6 + r2/2)).  TVDRK3: Q3b = W3^T (2/3 la + J3^T (2/3 la)),
    // Q2b = W2^T (Q3b/4 + J2^T (Q3b/4)), lam = W1^T (la/3 + 3/4 Q3b + Q2b +
    // J1^T Q2b).  (J_s: the derivative of dt R at the corrected state.)
    void adjoint_step_walls(const QSet<real> *stg, const StepRec &r, const QSet<real> &lam)
    {
        const real dt = (real)r.dt;
        QSet<real> la = lam;
        if (r.order >= 0) {
            filter_transposed(lam, la5, sc5, r.order);
            la = la5;
        }
        const QSet<real> X = vb5, Y = vb5b, acc = qb5;
        const QSet<real> none{};
        const long n = N;
        // the first stage after a sponge update corrects the pre-reset set
        // with the stage state (its planes: gcur_S)
        const QSet<real> *S = r.stale ? gcur_S : nullptr;
        if (r.stale && !S) throw std::logic_error("adjoint: stale stage without its planes");
        auto lc = [&](const QSet<real> &o, real ca, const QSet<real> &a, real cb, const QSet<real> &b) {
            hc(launch_gadj_lincomb<real>(o, ca, a, cb, b, n, st), "adjoint combination");
        };
        if (cfg.rk == CS3D_TVDRK3) {
            lc(X, real(2) / 3, la, 0, none);                  // 2/3 la
            wall_apply(stg[2], gw_q);
            { AdjVjp<real> f{}; f.cv = dt; f.out = Y; vjp_corrected(X, f, false); }
            lc(Y, 1, X, 1, Y);                                // 2/3 la + J3^T
            wall_vjp(Y);                                      // Q3b
            lc(acc, real(1) / 3, la, real(3) / 4, Y);         // la/3 + 3/4 Q3b
            lc(X, real(1) / 4, Y, 0, none);                   // Q3b/4
            wall_apply(stg[1], gw_q);
            { AdjVjp<real> f{}; f.cv = dt; f.out = Y; vjp_corrected(X, f, false); }
            lc(Y, 1, X, 1, Y);
            wall_vjp(Y);                                      // Q2b
            lc(acc, 1, acc, 1, Y);
            wall_apply(stg[0], gw_q, S);
            { AdjVjp<real> f{}; f.cv = dt; f.out = X; vjp_corrected(Y, f, S != nullptr); }
            lc(lam, 1, acc, 1, X);
            wall_vjp(lam, S ? &gl_stale : nullptr);
            return;
        }
        copy_set(acc, la);
        vjp_walls(stg[3], la, dt / 6, Y);                     // r4
        lc(acc, 1, acc, 1, Y);
        lc(X, real(1) / 3, la, 1, Y);                         // la/3 + r4
        vjp_walls(stg[2], X, dt, Y);                          // r3
        lc(acc, 1, acc, 1, Y);
        lc(X, real(1) / 3, la, real(1) / 2, Y);               // la/3 + r3/2
        vjp_walls(stg[1], X, dt, Y);                          // r2
        lc(acc, 1, acc, 1, Y);
        lc(X, real(1) / 6, la, real(1) / 2, Y);               // la/6 + r2/2
        wall_apply(stg[0], gw_q, S);
        { AdjVjp<real> f{}; f.cv = dt; f.out = Y; vjp_corrected(X, f, S != nullptr); }
        lc(lam, 1, acc, 1, Y);
        wall_vjp(lam, S ? &gl_stale : nullptr);
    }

    // lam (cotangent of q_{n+1}) -> cotangent of q_n for the step whose
    // stage states are stg
    void adjoint_step(const QSet<real> *stg, const StepRec &r, const QSet<real> &lam)
    {
        if (adj_general && any_wall()) {
            adjoint_step_walls(stg, r, lam);
            return;
        }
        const real rdt = (real)r.dt;
        QSet<real> la = lam;
        if (r.order >= 0) {
            filter_transposed(lam, la5, sc5, r.order);
            la = la5;
        }
        // acc = q + k1/6 + k2/3 + k3/3 + k4/6 (filtered into q_{n+1});
        // Q2 = q + k1/2, Q3 = q + k2/2, Q4 = q + k3; k_s = dt R(Q_s).  The
        // first product reads dt/6 la as la scaled on load; the stage
        // cotangents after it alternate between the sets A and B (the y pass
        // reads v's windows while it writes the next one)
        const QSet<real> A = vb5, B = vb5b;
        AdjVjp<real> f{};
        f.acc = qb5;
        f.la = la;
        if (cfg.rk == CS3D_TVDRK3) {
            // Q2 = q + k1, Q3 = 3/4 q + 1/4 (Q2 + k2), acc = 1/3 q + 2/3 (Q3 +
            // k3): dt k3b = 2 dt/3 la; dt k2b = dt/4 Q3b = dt/6 la + dt/4 res3;
            // dt k1b = dt Q2b = dt k2b + dt res2; qb = la + res3 + res2 + res1
            f.cv = 2 * rdt / 3;  // stage 3: v = la, next A
            f.acc_base = la;
            f.vn = A;
            f.ca = rdt / 6;
            f.cb = rdt / 4;
            vjp(stg[2], la, f);
            f.cv = 1;  // stage 2: v = A, next B = A + dt res
            f.acc_base = qb5;
            f.la = A;
            f.vn = B;
            f.ca = 1;
            f.cb = rdt;
            vjp(stg[1], A, f);
            AdjVjp<real> l{};  // stage 1: v = B; lam = acc + res
            l.cv = 1;
            l.acc = lam;
            l.acc_base = qb5;
            vjp(stg[0], B, l, r.stale != 0);
            return;
        }
        f.cv = rdt / 6;  // stage 4: v = la, next A
        f.acc_base = la;
        f.vn = A;
        f.ca = rdt / 3;
        f.cb = rdt;
        vjp(stg[3], la, f);
        f.cv = 1;  // stage 3: v = A, next B
        f.acc_base = qb5;
        f.vn = B;
        f.ca = rdt / 3;
        f.cb = rdt / 2;
        vjp(stg[2], A, f);
        f.vn = A;  // stage 2: v = B, next A
        f.ca = rdt / 6;
        f.cb = rdt / 2;
        vjp(stg[1], B, f);
        AdjVjp<real> l{};  // stage 1: v = A; lam = acc + res
        l.cv = 1;
        l.acc = lam;
        l.acc_base = qb5;
        vjp(stg[0], A, l, r.stale != 0);
    }

    // five host fields (FP64) <-> a device set, no reference offsets
    void upload_set(const double *const *in, const QSet<real> &dst)
    {
        std::vector<real> m(4 * (size_t)N), e((size_t)N);
        for (long i = 0; i < N; i++) {
            for (int v = 0; v < 4; v++) m[4 * i + v] = (real)in[v][i];
            e[i] = (real)in[4][i];
        }
        hc(hipMemcpy(dst.m, m.data(), m.size() * sizeof(real), hipMemcpyHostToDevice), "upload");
        hc(hipMemcpy(dst.e, e.data(), e.size() * sizeof(real), hipMemcpyHostToDevice), "upload");
    }
    void download_set(const QSet<real> &src, double *const *out)
    {
        std::vector<real> m(4 * (size_t)N), e((size_t)N);
        hc(hipStreamSynchronize(st), "sync");
        hc(hipMemcpy(m.data(), src.m, m.size() * sizeof(real), hipMemcpyDeviceToHost), "download");
        hc(hipMemcpy(e.data(), src.e, e.size() * sizeof(real), hipMemcpyDeviceToHost), "download");
        for (int v = 0; v < 5; v++)
            if (out[v])
                for (long i = 0; i < N; i++) out[v][i] = (double)(v < 4 ? m[4 * i + v] : e[i]);
    }

    void download(const real *src, double *dst)
    {
        std::vector<real> buf((size_t)N);
        hc(hipStreamSynchronize(st), "sync");
        hc(hipMemcpy(buf.data(), src, (size_t)N * sizeof(real), hipMemcpyDeviceToHost), "download");
        for (long i = 0; i < N; i++) dst[i] = (double)buf[i];
    }

    void rhs_vjp(int set, const double *const *lam, double *const *out) override
    {
        adj_require();
        for (int c = 0; c < 5; c++)
            if (!lam[c]) throw std::invalid_argument("rhs_vjp needs all five cotangent fields");
        StepScalars h;
        hc(hipMemcpyAsync(&h, dsc, sizeof(h), hipMemcpyDeviceToHost, st), "scalars");
        hc(hipStreamSynchronize(st), "sync");
        upload_set(lam, la5);
        AdjVjp<real> f{};
        f.cv = (real)h.dt;
        f.out = qb5;
        // with sponges: the averages and the stale planes held fixed
        const bool stale = sponge && set == 1 && stale_pending;
        gcur_S = stale ? &F.qstale : nullptr;
        if (sponge) zero_set(gl_avg);
        // with walls the map cs3d_rhs evaluates is Q -> dt R(W(Q))
        if (any_wall())
            vjp_walls(set == 1 ? F.q : F.qk, la5, f.cv, qb5, stale ? &F.qstale : nullptr);
        else
            vjp(set == 1 ? F.q : F.qk, la5, f, stale);
        download_set(qb5, out);
    }

    void zero_set(const QSet<real> &a)
    {
        hc(hipMemsetAsync(a.m, 0, 4 * (size_t)N * sizeof(real), st), "zero");
        hc(hipMemsetAsync(a.e, 0, (size_t)N * sizeof(real), st), "zero");
    }

    void ke_gradient(int nsteps, int K, double *J, double *gU, double *gV, double *gW) override
    {
        if (nsteps < 1 || K < 1) throw std::invalid_argument("nsteps and checkpoint must be >= 1");
        adj_require();
        // with sponges the averages at the start are the initial state
        // itself (set_initial_conditions), so the gradient counts their
        // cotangent as the state's
        if (sponge && timeStep != 0)
            throw std::invalid_argument("with sponges the gradient starts from set_initial_conditions");
        const int nseg = (nsteps + K - 1) / K, last0 = (nseg - 1) * K;
        // checkpoints, the stage states of one segment and the final state
        // come from a pool that persists across calls
        size_t used = 0;
        auto state = [&]() {
            if (used == adjpool.size()) adjpool.push_back(adj_alloc_set());
            return adjpool[used++];
        };
        const int ns = nstages();
        std::vector<QSet<real>> ck(nseg), stg(ns * (size_t)K);
        for (auto &c : stg) c = state();
        const QSet<real> fin = state();
        for (auto &c : ck) c = state();
        auto stages = [&](int i) { return (const QSet<real> *)&stg[ns * (size_t)i]; };
        // sponges: per checkpoint the averages and the stale planes, per step
        // of a segment the planes, the averages and the filtered state
        // (fused_step), and the end state's averages and planes
        std::vector<QSet<real>> ckA, ckS, srec, ckL;
        QSet<real> finA{}, finS{}, finL{};
        // the compensated mode: the state's low parts at each checkpoint and
        // at the end, so that the replays retrace the forward pass exactly
        // (the products themselves take the FP32 states)
        if (cfg.compensated) {
            ckL.resize(nseg);
            for (auto &c : ckL) c = state();
            finL = state();
        }
        if (sponge) {
            ckA.resize(nseg);
            ckS.resize(nseg);
            srec.resize(3 * (size_t)K);
            for (auto &c : ckA) c = state();
            for (auto &c : ckS) c = state();
            for (auto &c : srec) c = state();
            finA = state();
            finS = state();
        }
        auto sponge_rec = [&](int i) { return sponge ? (const QSet<real> *)&srec[3 * (size_t)i] : nullptr; };
        auto save_ck = [&](int c) {
            copy_set(ck[c], F.q);
            if (cfg.compensated) copy_set(ckL[c], F.qlo);
            if (sponge) {
                copy_set(ckA[c], F.avg);
                copy_set(ckS[c], F.qstale);
            }
        };
        std::vector<StepRec> rec(nsteps);
        // forward pass on the fused path: checkpoints, and the stage states
        // of the last segment
        for (int n = 0; n < nsteps; n++) {
            if (n % K == 0) save_ck(n / K);
            const bool keep = n >= last0;
            rec[n] = fused_step(nullptr, keep ? stages(n - last0) : nullptr,
                                keep ? sponge_rec(n - last0) : nullptr);
        }
        StepScalars after;
        hc(hipMemcpyAsync(&after, dsc, sizeof(after), hipMemcpyDeviceToHost, st), "scalars");
        hc(hipStreamSynchronize(st), "sync");
        const int ts_end = timeStep, fts_end = filterTimeStep;
        const bool stale_end = stale_pending;
        if (J) hc(kinetic_energy(J) == CS3D_OK ? hipSuccess : hipErrorUnknown, "kinetic energy");
        copy_set(fin, F.q);
        if (cfg.compensated) copy_set(finL, F.qlo);
        if (sponge) {
            copy_set(finA, F.avg);
            copy_set(finS, F.qstale);
            zero_set(gl_avg);
        }
        // cotangent of the final state
        hc(launch_adj_ke_seed<real>(F.q, ph, lamb, N, st), "adjoint seed");
        // reverse, segment by segment; with sponges each step first reverses
        // its sponge update (taking in the stale cotangent the next step's
        // first stage left), then its stages
        bool have_stale = false;
        for (int sgi = nseg - 1; sgi >= 0; sgi--) {
            const int n0 = sgi * K, n1 = std::min(nsteps, n0 + K);
            if (sgi != nseg - 1) {
                copy_set(F.q, ck[sgi]);
                if (cfg.compensated) copy_set(F.qlo, ckL[sgi]);
                if (sponge) {
                    copy_set(F.avg, ckA[sgi]);
                    copy_set(F.qstale, ckS[sgi]);
                }
                for (int n = n0; n < n1; n++) fused_step(&rec[n], stages(n - n0), sponge_rec(n - n0));
            }
            for (int n = n1 - 1; n >= n0; n--) {
                if (sponge) {
                    const QSet<real> *sr = sponge_rec(n - n0);
                    const double eps = 1.0 / (cfg.sponge_avgT / rec[n].dt + 1.0);
                    hc(launch_gadj_sponge_post<real>(sr[2], sr[1], lamb, gl_avg,
                                                     have_stale ? gl_stale : QSet<real>{}, base,
                                                     ph, eps, cfg.sponge_epsP, N, st),
                       "sponge update reverse");
                    gcur_S = &sr[0];
                }
                adjoint_step(stages(n - n0), rec[n], lamb);
                have_stale = rec[n].stale != 0;
            }
        }
        if (have_stale) throw std::logic_error("adjoint: stale first stage at the start");
        if (sponge) hc(launch_gadj_add<real>(lamb, gl_avg, N, st), "adjoint averages");
        gcur_S = nullptr;
        // to the initial primitive velocities at fixed rho and p
        hc(launch_adj_to_velocity<real>(ck[0], lamb, ph, gvel[0], gvel[1], gvel[2], N, st),
           "velocity gradient");
        if (gU) download(gvel[0], gU);
        if (gV) download(gvel[1], gV);
        if (gW) download(gvel[2], gW);
        // back to the final state of the forward pass
        copy_set(F.q, fin);
        if (cfg.compensated) copy_set(F.qlo, finL);
        if (sponge) {
            copy_set(F.avg, finA);
            copy_set(F.qstale, finS);
        }
        timeStep = ts_end;
        filterTimeStep = fts_end;
        stale_pending = stale_end;
        hc(hipMemcpyAsync(dsc, &after, sizeof(after), hipMemcpyHostToDevice, st), "scalars");
        hc(hipStreamSynchronize(st), "sync");
        diag_valid = false;  // the replays overwrote the step diagnostics
    }

    void adjoint_release() override
    {
        hc(hipStreamSynchronize(st), "sync");
        for (void *p : adjalloc) (void)hipFree(p);
        adjalloc.clear();
        adjpool.clear();
        for (int d = 0; d < 3; d++) {
            D1T[d].release();
            D2T[d].release();
            FLT[d].release();
        }
        adj_ready = false;
    }

    void synchronize() override { hc(hipStreamSynchronize(st), "sync"); }
    void *stream() override { return (void *)st; }
};

SolverBase *make_solver(const cs3d_config &c)
{
    if (c.precision == CS3D_FP32) return new SolverT<float>(c);
    return new SolverT<double>(c);
}

}  // namespace cs3d
