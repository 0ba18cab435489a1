// stage_common.hpp -- device helpers shared by the stage sweeps: ideal-gas
// relations on the perturbation-form state, and the per-lane set-up of a
// sweep workgroup.
#pragma once

#include "plan_dev.hpp"
#include "solver_kernels.hpp"

namespace cs3d {

// Working-precision reciprocal and the viscosity power law.  FP32 uses the
// hardware reciprocal and log2/exp2 (about 1 ulp each, far inside the FP32
// error budget of the scheme); FP64 keeps correctly rounded division and pow.
template <typename real>
__device__ __forceinline__ real rcp(real x);
template <>
__device__ __forceinline__ float rcp<float>(float x)
{
    return __builtin_amdgcn_rcpf(x);
}
template <>
__device__ __forceinline__ double rcp<double>(double x)
{
    return 1.0 / x;
}

template <typename real>
__device__ __forceinline__ real pow076(real x);
template <>
__device__ __forceinline__ float pow076<float>(float x)
{
    return __builtin_amdgcn_exp2f(0.76f * __builtin_amdgcn_logf(x));
}
template <>
__device__ __forceinline__ double pow076<double>(double x)
{
    return pow(x, 0.76);
}

// IdealGas.hpp:49-55 on the temperature: mu and Amu = dmu/dT
template <typename real>
__device__ __forceinline__ void visc(const Phys<real> &ph, real T, real &mu, real &amu)
{
    mu = ph.mu_ref * pow076(T * ph.iTref_ig);
    amu = real(0.76) * mu * rcp(T);
}

// conserved (perturbation form) -> 1/rho, velocity, p' = p - p_ref,
// T' = T - T0s  (IdealGas.hpp:57-71)
template <typename real>
__device__ __forceinline__ void prim(const Phys<real> &ph, real rp, real ru, real rv, real rw,
                                     real re, real &irho, real &U, real &V, real &W, real &pp,
                                     real &Tp)
{
    real rho = ph.rho_ref + rp;
    irho = rcp(rho);
    U = ru * irho;
    V = rv * irho;
    W = rw * irho;
    real ke = real(0.5) * (ru * U + rv * V + rw * W);
    pp = ph.gm1 * (re - ke);
    Tp = (pp - rp * ph.RT0s) * irho * ph.iR;
}

// Per-lane set-up of a sweep kernel: lane geometry, exchange ring (margin KRV,
// PLANES planes per parity), the line this lane works on (inactive lanes of a
// partial last workgroup take the last line and only skip their stores) and
// the lane's window `win`.  Needs M, HW, W and UNI in scope.
#define CS3D_LANE_SETUP(MAPV, OPP, KRV, PLANES)                                         \
    extern __shared__ char smem_raw[];                                                  \
    real *xb = (real *)smem_raw;                                                        \
    const int P = (OPP)->P;                                                             \
    const int LX = blockDim.x / P;                                                      \
    Lane<real, MAPV> c = make_lane<real, MAPV>((OPP), P, LX, (KRV), (PLANES), xb);      \
    const int L = xcd_block(blockIdx.x, gridDim.x) * LX + c.line;                       \
    const bool active = L < lm.nlines;                                                  \
    const int Lc = active ? L : lm.nlines - 1;                                          \
    const int n = (OPP)->n, per = (OPP)->cyclic;                                        \
    Win<real, M, HW, vec_windows<real, M, MAPV, UNI>()> win;                            \
    win.init((unsigned)lm.base(Lc), (unsigned)lm.sp, c.s, n, per);

// N1 first-derivative and N2 second-derivative solves
template <int N1, int N2, typename real, int M, bool UNI, int MAP>
__device__ __forceinline__ void solve_d12(const OpDev<real> *__restrict__ d1,
                                          const OpDev<real> *__restrict__ d2,
                                          Lane<real, MAP> &c, real (&r)[N1 + N2][M])
{
    solve_batch2<N1, N2, kc_deriv<real>(), kc_d2<real>(), real, M, UNI, MAP>(d1, d2, c, r);
}

// ring margin of a sweep that uses both derivative operators
template <typename real>
__device__ __forceinline__ int sweep_kr(const OpDev<real> *d1, const OpDev<real> *d2)
{
    int k = d1->K > d2->K ? d1->K : d2->K;
    return k > kc_deriv<real>() ? k : kc_deriv<real>();
}

// ring margin of the x sweep that also runs the fused x filter: on lines
// of 64 segments the filter's carries come from the wavefront scan and do
// not use the ring (the host sizes the ring the same way)
template <typename real>
__device__ __forceinline__ int xfilt_kr(const OpDev<real> *d1, const OpDev<real> *d2,
                                        const OpDev<real> *fx)
{
    int k = sweep_kr(d1, d2);
    if (fx->P == 64) return k;
    k = fx->K > k ? fx->K : k;
    return kc_filter_x<real>() > k ? kc_filter_x<real>() : k;
}

// q windows a0..a4 (perturbation form) and ir = 1/rho at window index t
#define CS3D_U(t) (a1[t] * ir[t])
#define CS3D_V(t) (a2[t] * ir[t])
#define CS3D_W(t) (a3[t] * ir[t])

template <typename real, int W>
__device__ __forceinline__ real pprime(const Phys<real> &ph, const real (&a1)[W],
                                       const real (&a2)[W], const real (&a3)[W],
                                       const real (&a4)[W], const real (&ir)[W], int t)
{
    const real U = a1[t] * ir[t], V = a2[t] * ir[t], Wv = a3[t] * ir[t];
    return ph.gm1 * (a4[t] - real(0.5) * (a1[t] * U + a2[t] * V + a3[t] * Wv));
}

// the last (x) sweep: lines are contiguous, so small workgroups lose no
// coalescing and leave each lane the larger register budget its 17 solves
// and the pointwise assembly need
constexpr int X_SWEEP_THREADS = 256;

template <typename real, bool UNI>
hipError_t launch_ypass(const OpDev<real> *d1, const OpDev<real> *d2, const LaunchGeom &lg,
                        const LineMap &lm, const Fields<real> &F, const Phys<real> &ph,
                        int use_k, const StageArgs &sa, hipStream_t st);
// M: SEG_DERIV, or SEG_DERIV3 on uniform plans
template <typename real, bool UNI>
hipError_t launch_zpass(const OpDev<real> *d1, const OpDev<real> *d2, int M, const LaunchGeom &lg,
                        const LineMap &lm, const Fields<real> &F, const Phys<real> &ph,
                        int use_k, const StageArgs &sa, hipStream_t st);
// fx non-null: the fused x filter of the step's last stage (stage_x.hip)
template <typename real, bool UNI>
hipError_t launch_xpass(const OpDev<real> *d1, const OpDev<real> *d2, const OpDev<real> *fx,
                        const LaunchGeom &lg, const LineMap &lm, const Fields<real> &F,
                        const Phys<real> &ph, const StageArgs &sa, const StepScalars *sc,
                        hipStream_t st);

}  // namespace cs3d
