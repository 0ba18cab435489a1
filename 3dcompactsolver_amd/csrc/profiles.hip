// profiles.hip -- x-z plane statistics of the state per y index, the
// shear-layer profiles of shearLayerInfoCalc (CSolver_AWS.cpp:2259-2456).
//
// One workgroup per (j, k) row of Nx points: FP64 sums of the row in a
// fixed order; the host adds the rows of each plane in k order.  Pass 0:
// rho, U, V, W, p, T.  Pass 1, with the plane means: the second moments
// (rho', u', v', w', p', T' squared, u'v', u'w', v'w').
#include <mutex>

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
        real irho, U, V, W, pp, Tp, q5[5];
        qset_load(F.q, ii, q5);
        prim(ph, q5[0], q5[1], q5[2], q5[3], q5[4], irho, U, V, W, pp, Tp);
        const double v[6] = {(double)ph.rho_ref + (double)q5[0], (double)U, (double)V,
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

// ------------------------------------------------ profiles without allocation
//
// For callers that cannot allocate device memory (code under graph capture,
// serving loops): the row sums, the plane means and moments and the 11
// output rows live in one file-scope device array, each GPU holding its own
// copy until the process ends.  Layout in doubles: rows (SHEAR_MAX_ROWS x
// PROF_Q) | means (SHEAR_MAX_NY x 6) | moments (SHEAR_MAX_NY x PROF_Q) |
// output (SHEAR_MAX_NY x 11): 10.6 MiB.  Every part is written in full
// before it is read, so nothing needs clearing between calls.
constexpr long SHEAR_OFF_MEAN = SHEAR_MAX_ROWS * PROF_Q;
constexpr long SHEAR_OFF_MOM = SHEAR_OFF_MEAN + (long)SHEAR_MAX_NY * 6;
constexpr long SHEAR_OFF_OUT = SHEAR_OFF_MOM + (long)SHEAR_MAX_NY * PROF_Q;
constexpr long SHEAR_SCRATCH = SHEAR_OFF_OUT + (long)SHEAR_MAX_NY * 11;

static __device__ __align__(256) double scratch[SHEAR_SCRATCH];

// the array's device address per GPU, taken once with that GPU current
// (the name `scratch` is only a host-side stand-in in host code)
constexpr int SHEAR_MAX_DEVICES = 64;
static double *g_scratch[SHEAR_MAX_DEVICES];
static std::mutex g_scratch_mu;

static hipError_t scratch_address(int dev, double **out)
{
    if (dev < 0 || dev >= SHEAR_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return e;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        g_scratch[dev] = (double *)p;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

// dst[j * nq + q] = (sum over k in ascending order of the rows' entry q)
// * inv: the order and arithmetic of the host reduction of
// cs3d_diag_shear_layer
__global__ void k_plane_reduce(const double *__restrict__ rows, int Ny, int Nz, int nq,
                               double inv, double *__restrict__ dst)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= Ny * nq) return;
    const int j = idx / nq, q = idx - j * nq;
    double s = 0.0;
    for (int k = 0; k < Nz; k++) s += rows[((long)k * Ny + j) * PROF_Q + q];
    dst[idx] = s * inv;
}

// the 11 output rows (Q12: v'w' into uwprime and vwprime zero unless fixed)
__global__ void k_shear_rows_out(const double *__restrict__ mean, const double *__restrict__ mom,
                                 int Ny, int fix_q12, double *__restrict__ o)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= Ny) return;
    const double *m = mom + (long)j * PROF_Q;
    for (int q = 0; q < 6; q++) o[(long)q * Ny + j] = m[q];
    o[6L * Ny + j] = m[6];
    o[7L * Ny + j] = fix_q12 ? m[7] : m[7] + m[8];
    o[8L * Ny + j] = fix_q12 ? m[8] : 0.0;
    o[9L * Ny + j] = mean[(long)j * 6 + 1];
    o[10L * Ny + j] = mean[(long)j * 6 + 0];
}

template <typename real>
hipError_t shear_profiles_static(const Fields<real> &F, const Phys<real> &ph, int Nx, int Ny,
                                 int Nz, int fix_q12, double *host_out, hipStream_t st)
{
    if ((long)Ny * Nz > SHEAR_MAX_ROWS || Ny > SHEAR_MAX_NY) return hipErrorInvalidValue;
    int dev = 0;
    hipError_t e = hipStreamGetDevice(st, &dev);
    double *base = nullptr;
    if (e == hipSuccess) e = scratch_address(dev, &base);
    if (e != hipSuccess) return e;
    double *rows = base, *mean = base + SHEAR_OFF_MEAN, *mom = base + SHEAR_OFF_MOM,
           *out = base + SHEAR_OFF_OUT;
    const double inv = 1.0 / ((double)Nx * Nz);
    const int T = 256;
    if ((e = launch_plane_rows<real>(F, ph, Nx, Ny, Nz, 0, nullptr, rows, st)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_plane_reduce, dim3((Ny * 6 + T - 1) / T), dim3(T), 0, st, rows, Ny, Nz, 6,
                       inv, mean);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = launch_plane_rows<real>(F, ph, Nx, Ny, Nz, 1, mean, rows, st)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_plane_reduce, dim3((Ny * PROF_Q + T - 1) / T), dim3(T), 0, st, rows, Ny,
                       Nz, PROF_Q, inv, mom);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_shear_rows_out, dim3((Ny + T - 1) / T), dim3(T), 0, st, mean, mom, Ny,
                       fix_q12, out);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(host_out, out, (size_t)Ny * 11 * sizeof(double),
                            hipMemcpyDeviceToHost, st)) != hipSuccess)
        return e;
    return hipStreamSynchronize(st);
}

template hipError_t shear_profiles_static<float>(const Fields<float> &, const Phys<float> &, int,
                                                 int, int, int, double *, hipStream_t);
template hipError_t shear_profiles_static<double>(const Fields<double> &, const Phys<double> &,
                                                  int, int, int, int, double *, hipStream_t);

}  // namespace cs3d
