// profiles.hip -- x-z plane statistics of the state per y index, the
// shear-layer profiles of shearLayerInfoCalc (CSolver_AWS.cpp:2259-2456).
//
// One workgroup per (j, k) row of Nx points: FP64 sums of the row in a
// fixed order; the host adds the rows of each plane in k order.  Pass 0:
// rho, U, V, W, p, T.  Pass 1, with the plane means: the second moments
// (rho', u', v', w', p', T' squared, u'v', u'w', v'w').
#include "solver_kernels.hpp"
#include "stage_common.hpp"

namespace cs3d {

constexpr int PROF_THREADS = 256;
constexpr int PROF_Q = 9;

template <typename real>
__global__ __launch_bounds__(PROF_THREADS) void k_plane_rows(Fields<real> F, Phys<real> ph, int Nx,
                                                             int Ny, int pass,
                                                             const double *__restrict__ mean,
                                                             double *__restrict__ rows)
{
    const int j = blockIdx.x, k = blockIdx.y;
    const long base = ((long)k * Ny + j) * Nx;
    double acc[PROF_Q] = {};
    double m[6] = {};
    if (pass == 1)
        for (int q = 0; q < 6; q++) m[q] = mean[(long)j * 6 + q];
    for (int i = threadIdx.x; i < Nx; i += PROF_THREADS) {
        const long ii = base + i;
        real irho, U, V, W, pp, Tp;
        prim(ph, F.q[0][ii], F.q[1][ii], F.q[2][ii], F.q[3][ii], F.q[4][ii], irho, U, V, W, pp, Tp);
        const double v[6] = {(double)ph.rho_ref + (double)F.q[0][ii], (double)U, (double)V,
                             (double)W, (double)ph.p_ref + (double)pp,
                             (double)ph.T0s + (double)Tp};
        if (pass == 0) {
            for (int q = 0; q < 6; q++) acc[q] += v[q];
        } else {
            double d[6];
            for (int q = 0; q < 6; q++) {
                d[q] = v[q] - m[q];
                acc[q] += d[q] * d[q];
            }
            acc[6] += d[1] * d[2];
            acc[7] += d[1] * d[3];
            acc[8] += d[2] * d[3];
        }
    }
    __shared__ double sh[PROF_Q][PROF_THREADS];
    for (int q = 0; q < PROF_Q; q++) sh[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int w = PROF_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < PROF_Q; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
        __syncthreads();
    }
    if ((int)threadIdx.x < PROF_Q)
        rows[((long)k * Ny + j) * PROF_Q + threadIdx.x] = sh[threadIdx.x][0];
}

template <typename real>
hipError_t launch_plane_rows(const Fields<real> &F, const Phys<real> &ph, int Nx, int Ny, int Nz,
                             int pass, const double *mean, double *rows, hipStream_t st)
{
    if (Ny > 65535 || Nz > 65535) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((k_plane_rows<real>), dim3(Ny, Nz), dim3(PROF_THREADS), 0, st, F, ph, Nx,
                       Ny, pass, mean, rows);
    return hipGetLastError();
}

template hipError_t launch_plane_rows<float>(const Fields<float> &, const Phys<float> &, int, int,
                                             int, int, const double *, double *, hipStream_t);
template hipError_t launch_plane_rows<double>(const Fields<double> &, const Phys<double> &, int,
                                              int, int, int, const double *, double *,
                                              hipStream_t);

}  // namespace cs3d
