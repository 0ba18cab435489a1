// aero_optics.hip -- optical-path integrals of the density along slanted
// rays: AeroOptics::computeAO (AeroOptics.cpp:45-93) with
// generateBaseBeamDiff (:3-24) and Utils.cpp's getBaseNodeIndex and
// linearInterpolation (:242-332).
//
// For angle p of nA (a_p = p/(nA-1) * maxAngle degrees) and every (x_i,
// z_k): Nline points of the ray from (x_i, 0.2 Ly, z_k) to y = 0.8 Ly,
// leaning by a_p in x with x wrapped by fmod(., Lx); the density trilinearly
// interpolated at each, integrated by the trapezoid rule in FP64.  One
// thread per (x_i, z_k, angle); consecutive threads take consecutive x_i, so
// the gathers of one ray point fall on neighbouring cells.
//
// Last-plane rule: where the base index is the last point of an axis the
// reference reads one point past its arrays with weight zero (z_k at
// k = Nz - 1) or of rounding size; here the neighbour's coordinate is the
// formula's next point and its value the last point's own.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/cs3d.h"
#include "aero_optics.hpp"
#include "capi_common.hpp"

namespace cs3d {

// Domain.hpp:30-45: x_i = i / (N - 1) L
__host__ __device__ inline double ao_coord(int i, int n, double L)
{
    return (((double)i) / ((double)n - 1.0)) * L;
}

void ao_geometry(int nA, double maxAngle, double Ly, double ds, AoAngles &g)
{
    const double totalDY = 0.8 * Ly - 0.2 * Ly;
    g.n = nA;
    for (int ap = 0; ap < nA; ap++) {
        const double ang = (((double)ap / ((double)(nA - 1))) * maxAngle) * M_PI / 180.0;
        const double Ldiag = totalDY / std::cos(ang), totalDX = Ldiag * std::sin(ang);
        const int Nline = (int)std::ceil(Ldiag / ds) + 1;
        g.nline[ap] = Nline;
        g.dsLine[ap] = Ldiag / (double)(Nline - 1);
        g.dxLine[ap] = totalDX / (double)(Nline - 1);
        g.dyLine[ap] = totalDY / (double)(Nline - 1);
    }
}

// density sources: a conserved set's rho' plus rho_ref, or an FP64 field
template <typename real>
struct DenSet {
    const real *m;
    double rho_ref;
    __device__ double operator()(long i) const { return rho_ref + (double)m[4 * i]; }
};
struct DenPtr {
    const double *f;
    __device__ double operator()(long i) const { return f[i]; }
};

template <typename Den>
__global__ __launch_bounds__(256) void k_opl(Den den, AoDomain d, AoAngles g, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y, ap = blockIdx.z;
    if (i >= d.Nx) return;
    const int n[3] = {d.Nx, d.Ny, d.Nz};
    const double L[3] = {d.Lx, d.Ly, d.Lz};
    const double h[3] = {ao_coord(1, d.Nx, d.Lx), ao_coord(1, d.Ny, d.Ly), ao_coord(1, d.Nz, d.Lz)};
    const int Nline = g.nline[ap];
    const double dsLine = g.dsLine[ap], dxLine = g.dxLine[ap], dyLine = g.dyLine[ap];
    const double lower = 0.2 * d.Ly, xi = ao_coord(i, d.Nx, d.Lx), zk = ao_coord(k, d.Nz, d.Lz);
    const long sxy = (long)d.Nx * d.Ny;
    auto at = [&](int sp) -> double {
        const double xp[3] = {fmod(xi + (double)(sp * dxLine), d.Lx), lower + (double)(sp * dyLine),
                              zk};
        int i0[3], i1[3];
        double w[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const int b = (int)(xp[a] / h[a]);
            const double c0 = ao_coord(b, n[a], L[a]), c1 = ao_coord(b + 1, n[a], L[a]);
            w[a] = (xp[a] - c0) / (c1 - c0);
            i0[a] = b;
            i1[a] = b + 1 < n[a] ? b + 1 : n[a] - 1;
        }
        auto F = [&](int ii, int jj, int kk) { return den((long)kk * sxy + (long)jj * d.Nx + ii); };
        const double c00 = F(i0[0], i0[1], i0[2]) * (1 - w[0]) + F(i1[0], i0[1], i0[2]) * w[0];
        const double c01 = F(i0[0], i0[1], i1[2]) * (1 - w[0]) + F(i1[0], i0[1], i1[2]) * w[0];
        const double c10 = F(i0[0], i1[1], i0[2]) * (1 - w[0]) + F(i1[0], i1[1], i0[2]) * w[0];
        const double c11 = F(i0[0], i1[1], i1[2]) * (1 - w[0]) + F(i1[0], i1[1], i1[2]) * w[0];
        const double c0 = c00 * (1 - w[1]) + c10 * w[1];
        const double c1 = c01 * (1 - w[1]) + c11 * w[1];
        return c0 * (1 - w[2]) + c1 * w[2];
    };
    // trapezoid rule in the reference's order: both ends, then the interior
    double acc = at(0) * dsLine / 2.0 + at(Nline - 1) * dsLine / 2.0;
    for (int sp = 1; sp < Nline - 1; sp++) acc += at(sp) * dsLine;
    out[(long)ap * d.Nx * d.Nz + (long)k * d.Nx + i] = acc;
}

template <typename Den>
static hipError_t launch_opl(const Den &den, const AoDomain &d, const AoAngles &g, double *out,
                             hipStream_t st)
{
    if (g.n < 2 || g.n > AO_MAX_ANGLES || d.Nz > 65535) return hipErrorInvalidValue;
    dim3 grid((d.Nx + 255) / 256, d.Nz, g.n), block(256);
    hipLaunchKernelGGL((k_opl<Den>), grid, block, 0, st, den, d, g, out);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_opl_set(const real *quads, double rho_ref, const AoDomain &d, const AoAngles &g,
                          double *out, hipStream_t st)
{
    return launch_opl(DenSet<real>{quads, rho_ref}, d, g, out, st);
}
template hipError_t launch_opl_set<float>(const float *, double, const AoDomain &,
                                          const AoAngles &, double *, hipStream_t);
template hipError_t launch_opl_set<double>(const double *, double, const AoDomain &,
                                           const AoAngles &, double *, hipStream_t);

std::string ao_check(int Nx, int Ny, int Nz, double Lx, double Ly, double Lz, int nA,
                     double maxAngle)
{
    if (Nx < 2 || Ny < 2 || Nz < 2 || Nz > 65535) return "grid dimensions must be 2 .. 65535";
    if (!(Lx > 0.0 && Ly > 0.0 && Lz > 0.0)) return "domain lengths must be positive";
    if (nA < 2 || nA > AO_MAX_ANGLES)
        return "num_angles must be 2 .. " + std::to_string(AO_MAX_ANGLES);
    if (!(maxAngle >= 0.0 && maxAngle < 90.0)) return "max_angle must be in [0, 90) degrees";
    return "";
}

}  // namespace cs3d

using namespace cs3d;

extern "C" int cs3d_aero_optics_field(int Nx, int Ny, int Nz, double Lx, double Ly, double Lz,
                                      const double *rho, int num_angles, double max_angle,
                                      double *opl)
{
    if (!rho || !opl) return set_error(CS3D_ERR_ARG, "cs3d_aero_optics_field: NULL argument");
    const std::string bad = ao_check(Nx, Ny, Nz, Lx, Ly, Lz, num_angles, max_angle);
    if (!bad.empty()) return set_error(CS3D_ERR_ARG, "cs3d_aero_optics_field: " + bad);
    AoDomain d{Nx, Ny, Nz, Lx, Ly, Lz};
    AoAngles g;
    ao_geometry(num_angles, max_angle, Ly,
                std::fmin(ao_coord(1, Nx, Lx), std::fmin(ao_coord(1, Ny, Ly), ao_coord(1, Nz, Lz))),
                g);
    const size_t N = (size_t)Nx * Ny * Nz, nout = (size_t)num_angles * Nx * Nz;
    double *drho = nullptr, *dout = nullptr;
    hipError_t e = hipMalloc(&drho, N * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dout, nout * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(drho, rho, N * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_opl(DenPtr{drho}, d, g, dout, nullptr);
    if (e == hipSuccess) e = hipMemcpy(opl, dout, nout * sizeof(double), hipMemcpyDeviceToHost);
    if (drho) (void)hipFree(drho);
    if (dout) (void)hipFree(dout);
    if (e != hipSuccess)
        return set_error(CS3D_ERR_HIP, std::string("cs3d_aero_optics_field: ") + hipGetErrorString(e));
    return CS3D_OK;
}
