// aero_optics.hpp -- the optical-path integrals (aero_optics.hip) shared
// with the solver.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace cs3d {

constexpr int AO_MAX_ANGLES = 64;

struct AoDomain {
    int Nx, Ny, Nz;
    double Lx, Ly, Lz;
};

// the ray geometry of each angle (generateBaseBeamDiff, AeroOptics.cpp:3-24)
struct AoAngles {
    int n;
    int nline[AO_MAX_ANGLES];
    double dsLine[AO_MAX_ANGLES], dxLine[AO_MAX_ANGLES], dyLine[AO_MAX_ANGLES];
};

// ds = min(dx, dy, dz) (AeroOptics.hpp:52)
void ao_geometry(int nA, double maxAngle, double Ly, double ds, AoAngles &g);
// empty when the arguments are usable, else the reason
std::string ao_check(int Nx, int Ny, int Nz, double Lx, double Ly, double Lz, int nA,
                     double maxAngle);
// the integrals of rho_ref + rho' with rho' the first value of each quad
template <typename real>
hipError_t launch_opl_set(const real *quads, double rho_ref, const AoDomain &d, const AoAngles &g,
                          double *out, hipStream_t st);

}  // namespace cs3d
