// stage_y.hip -- Y sweep of one RK stage (lines along y, MAP_COL lanes:
// every load instruction reads consecutive x).
//
// y Euler fluxes (CSolver_AWS.cpp:949-963) and their derivatives
// (:1051-1073), first and second y-derivatives of U, V, W, T (:969-1049),
// the mixed Uxy, Vxy from the stored Ux, Vx (:991-1003); stores Uy Vy Wy Ty
// and adds every term linear in a y-derivative to the right-hand side.
#include "stage_common.hpp"

namespace cs3d {

template <typename real, int M, bool UNI>
__global__ __launch_bounds__(sweep_threads<real>()) void k_ypass(
    const OpDev<real> *__restrict__ d1, const OpDev<real> *__restrict__ d2, LineMap lm,
    Fields<real> F, Phys<real> ph, int use_k)
{
    CS3D_LANE_SETUP(MAP_COL, d1)
    constexpr int HW = 2, W = M + 2 * HW;
    const real *q0 = use_k ? F.qk[0] : F.q[0];
    const real *q1 = use_k ? F.qk[1] : F.q[1];
    const real *q2 = use_k ? F.qk[2] : F.q[2];
    const real *q3 = use_k ? F.qk[3] : F.q[3];
    const real *q4 = use_k ? F.qk[4] : F.q[4];
    real a0[W], a1[W], a2[W], a3[W], a4[W], ir[W];
    load_window<real, M, HW>([&](int p) { return q0[base + p * sp]; }, c.s, n, per, a0);
    load_window<real, M, HW>([&](int p) { return q1[base + p * sp]; }, c.s, n, per, a1);
    load_window<real, M, HW>([&](int p) { return q2[base + p * sp]; }, c.s, n, per, a2);
    load_window<real, M, HW>([&](int p) { return q3[base + p * sp]; }, c.s, n, per, a3);
    load_window<real, M, HW>([&](int p) { return q4[base + p * sp]; }, c.s, n, per, a4);
#pragma unroll
    for (int t = 0; t < W; t++) ir[t] = real(1) / (ph.rho_ref + a0[t]);
    real pp[W], w[W], r[M];
#pragma unroll
    for (int t = 0; t < W; t++) pp[t] = pprime(ph, a1, a2, a3, a4, ir, t);
    real R0[M], Ru[M], Rv[M], Rw[M], RE[M];
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = (a4[t] + pp[t]) * CS3D_V(t);
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) RE[t] = -r[t];
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = a2[t] * CS3D_V(t) + pp[t];
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) Rv[t] = -r[t];
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, a2, r);
#pragma unroll
    for (int t = 0; t < M; t++) R0[t] = -r[t];
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = a1[t] * CS3D_V(t);
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) Ru[t] = -r[t];
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = a3[t] * CS3D_V(t);
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) Rw[t] = -r[t];

    auto store = [&](real *dst, const real(&v)[M]) {
        if (active) {
#pragma unroll
            for (int t = 0; t < M; t++)
                if (t < c.len) dst[base + (long)(c.s + t) * sp] = v[t];
        }
    };
    real mu[M], Uc[M], Vc[M];
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = (pp[t] - a0[t] * ph.RT0s) * ir[t] * ph.iR;
#pragma unroll
    for (int t = 0; t < M; t++) {
        real amu;
        visc(ph, ph.T0s + w[HW + t], mu[t], amu);
        Uc[t] = CS3D_U(HW + t);
        Vc[t] = CS3D_V(HW + t);
    }
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, r);
    store(F.dY[3], r);
    compact_op<OP_D2, real, M, HW, UNI, MAP_COL>(d2, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) RE[t] += ph.cpPr * (mu[t] * r[t]);
    const real c43 = real(4) / real(3), c13 = real(1) / real(3);
    // U: Uy, Uyy
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = CS3D_U(t);
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, r);
    store(F.dY[0], r);
    compact_op<OP_D2, real, M, HW, UNI, MAP_COL>(d2, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) {
        Ru[t] += mu[t] * r[t];
        RE[t] += mu[t] * (Uc[t] * r[t]);
    }
    // V: Vy, Vyy
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = CS3D_V(t);
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, r);
    store(F.dY[1], r);
#pragma unroll
    for (int t = 0; t < M; t++) RE[t] -= ph.h_ref * r[t];
    compact_op<OP_D2, real, M, HW, UNI, MAP_COL>(d2, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) {
        Rv[t] += mu[t] * (c43 * r[t]);
        RE[t] += mu[t] * (Vc[t] * (c43 * r[t]));
    }
    // W: Wy, Wyy
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = CS3D_W(t);
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, r);
    store(F.dY[2], r);
    compact_op<OP_D2, real, M, HW, UNI, MAP_COL>(d2, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) {
        Rw[t] += mu[t] * r[t];
        RE[t] += mu[t] * (w[HW + t] * r[t]);
    }
    // mixed Uxy = d/dy(Ux), Vxy = d/dy(Vx)
    load_window<real, M, HW>([&](int p) { return F.dX[0][base + p * sp]; }, c.s, n, per, w);
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) {
        Rv[t] += mu[t] * (c13 * r[t]);
        RE[t] += mu[t] * (c13 * (Vc[t] * r[t]));
    }
    load_window<real, M, HW>([&](int p) { return F.dX[1][base + p * sp]; }, c.s, n, per, w);
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) {
        Ru[t] += mu[t] * (c13 * r[t]);
        RE[t] += mu[t] * (c13 * (Uc[t] * r[t]));
    }
    if (active) {
#pragma unroll
        for (int t = 0; t < M; t++) {
            if (t < c.len) {
                long i = base + (long)(c.s + t) * sp;
                F.R[0][i] = F.R[0][i] + R0[t];
                F.R[1][i] = F.R[1][i] + Ru[t];
                F.R[2][i] = F.R[2][i] + Rv[t];
                F.R[3][i] = F.R[3][i] + Rw[t];
                F.R[4][i] = F.R[4][i] + RE[t];
            }
        }
    }
}

template <typename real, bool UNI>
hipError_t launch_ypass(const OpDev<real> *d1, const OpDev<real> *d2, int threads, int LX,
                        const LineMap &lm, const Fields<real> &F, const Phys<real> &ph,
                        int use_k, hipStream_t st)
{
    dim3 grid((lm.nlines + LX - 1) / LX), block(threads);
    size_t shm = 2 * (size_t)threads * sizeof(real);
    hipLaunchKernelGGL((k_ypass<real, SEG_DERIV, UNI>), grid, block, shm, st, d1, d2, lm, F, ph,
                       use_k);
    return hipGetLastError();
}

#define CS3D_INST_Y(real, UNI)                                                              \
    template hipError_t launch_ypass<real, UNI>(const OpDev<real> *, const OpDev<real> *, int, \
                                                int, const LineMap &, const Fields<real> &,    \
                                                const Phys<real> &, int, hipStream_t);
CS3D_INST_Y(float, true)
CS3D_INST_Y(float, false)
CS3D_INST_Y(double, true)
CS3D_INST_Y(double, false)

}  // namespace cs3d
