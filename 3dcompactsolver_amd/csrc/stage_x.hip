// stage_x.hip -- X sweep of one RK stage (lines along x, contiguous in
// memory, MAP_ROW lanes).
//
// From the stage state: the x Euler fluxes (CSolver_AWS.cpp:861-867) and
// their first derivatives (:915-922), the first and second x-derivatives of
// U, V, W, T (:877-912); stores Ux Vx Wx Tx and starts the right-hand side
// with every term that is linear in an x-derivative.  The 13 line solves run
// as two batches (5 flux derivatives; 4 first + 4 second derivatives), each
// batch sharing one pair of exchange barriers.
//
// The reference's X block forms its fluxes and derivatives from the
// primitive arrays as they stand (CSolver_AWS.cpp:861-912).  At the first
// stage those were computed (updateNonConservedData, :1913-1925) before
// updateSponge reset the sponge planes of the state (:1955-2021), so on
// those planes the primitives belong to the state before the reset while the
// conserved fluxes use the state after it.  The STALE variant reproduces
// this from the copy of the pre-reset plane values that the sponge update
// keeps.
#include "stage_common.hpp"

namespace cs3d {

template <typename real, int M, bool UNI, bool STALE>
__global__ __launch_bounds__(sweep_threads<real>()) void k_xpass(
    const OpDev<real> *__restrict__ d1, const OpDev<real> *__restrict__ d2, LineMap lm,
    Fields<real> F, Phys<real> ph, int use_k, StageArgs sa)
{
    constexpr int HW = 2, W = M + 2 * HW;
    CS3D_LANE_SETUP(MAP_ROW, d1, sweep_kr(d1, d2), SWEEP_PLANES)
    const real *q0 = use_k ? F.qk[0] : F.q[0];
    const real *q1 = use_k ? F.qk[1] : F.q[1];
    const real *q2 = use_k ? F.qk[2] : F.q[2];
    const real *q3 = use_k ? F.qk[3] : F.q[3];
    const real *q4 = use_k ? F.qk[4] : F.q[4];
    real a0[W], a1[W], a2[W], a3[W], a4[W];
    real ir[W], U[W], V[W], Wv[W], pp[W], Tp[W];
    win.load(q0, a0);
    win.load(q1, a1);
    win.load(q2, a2);
    win.load(q3, a3);
    win.load(q4, a4);
#pragma unroll
    for (int t = 0; t < W; t++)
        prim(ph, a0[t], a1[t], a2[t], a3[t], a4[t], ir[t], U[t], V[t], Wv[t], pp[t], Tp[t]);
    if (STALE) {
        // primitives on sponge planes from the pre-reset copy
        const int jj = Lc % sa.Ny, kk = Lc / sa.Ny;
        const bool line_plane = (sa.sponge_face[2] && jj == 0) ||
                                (sa.sponge_face[3] && jj == sa.Ny - 1) ||
                                (sa.sponge_face[4] && kk == 0) ||
                                (sa.sponge_face[5] && kk == sa.Nz - 1);
#pragma unroll
        for (int t = 0; t < W; t++) {
            int p = c.s - HW + t;
            p += p < 0 ? n : 0;
            p -= p >= n ? n : 0;
            const bool plane = line_plane || (sa.sponge_face[0] && p == 0) ||
                               (sa.sponge_face[1] && p == n - 1);
            if (plane)
                prim(ph, ldo(F.qstale[0], win.at(t)), ldo(F.qstale[1], win.at(t)),
                     ldo(F.qstale[2], win.at(t)), ldo(F.qstale[3], win.at(t)),
                     ldo(F.qstale[4], win.at(t)), ir[t], U[t], V[t], Wv[t], pp[t], Tp[t]);
        }
    }
    real R0[M], Ru[M], Rv[M], Rw[M], RE[M];
    {
        // x fluxes; the energy flux (rhoE + p)U = h_ref U + (rhoE' + p')U has
        // its h_ref part added below as h_ref * Ux
        real r[5][M], f[W];
#pragma unroll
        for (int t = 0; t < W; t++) f[t] = (a4[t] + pp[t]) * U[t];
        rhs<OP_D1, real, M, HW, UNI>(d1, c, f, r[0]);
#pragma unroll
        for (int t = 0; t < W; t++) f[t] = a1[t] * U[t] + pp[t];
        rhs<OP_D1, real, M, HW, UNI>(d1, c, f, r[1]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, a1, r[2]);
#pragma unroll
        for (int t = 0; t < W; t++) f[t] = a2[t] * U[t];
        rhs<OP_D1, real, M, HW, UNI>(d1, c, f, r[3]);
#pragma unroll
        for (int t = 0; t < W; t++) f[t] = a3[t] * U[t];
        rhs<OP_D1, real, M, HW, UNI>(d1, c, f, r[4]);
        solve_batch<5, kc_deriv<real>(), real, M, UNI>(d1, c, r);
#pragma unroll
        for (int t = 0; t < M; t++) {
            RE[t] = -r[0][t];
            Ru[t] = -r[1][t];
            R0[t] = -r[2][t];
            Rv[t] = -r[3][t];
            Rw[t] = -r[4][t];
        }
    }
    // first and second derivatives of T', U, V, W
    real r[8][M];
    rhs<OP_D1, real, M, HW, UNI>(d1, c, Tp, r[0]);
    rhs<OP_D1, real, M, HW, UNI>(d1, c, U, r[1]);
    rhs<OP_D1, real, M, HW, UNI>(d1, c, V, r[2]);
    rhs<OP_D1, real, M, HW, UNI>(d1, c, Wv, r[3]);
    rhs<OP_D2, real, M, HW, UNI>(d2, c, Tp, r[4]);
    rhs<OP_D2, real, M, HW, UNI>(d2, c, U, r[5]);
    rhs<OP_D2, real, M, HW, UNI>(d2, c, V, r[6]);
    rhs<OP_D2, real, M, HW, UNI>(d2, c, Wv, r[7]);
    solve_batch2<4, 4, kc_deriv<real>(), real, M, UNI>(d1, d2, c, r);

    const real c43 = real(4) / real(3);
#pragma unroll
    for (int t = 0; t < M; t++) {
        real mu, amu;
        visc(ph, ph.T0s + Tp[HW + t], mu, amu);
        RE[t] += ph.cpPr * (mu * r[4][t]);
        RE[t] -= ph.h_ref * r[1][t];
        Ru[t] += mu * (c43 * r[5][t]);
        RE[t] += mu * (U[HW + t] * (c43 * r[5][t]));
        Rv[t] += mu * r[6][t];
        RE[t] += mu * (V[HW + t] * r[6][t]);
        Rw[t] += mu * r[7][t];
        RE[t] += mu * (Wv[HW + t] * r[7][t]);
    }
    if (active) {
        real Ux[M], Vx[M], Wx[M], Tx[M];
#pragma unroll
        for (int t = 0; t < M; t++) {
            Tx[t] = r[0][t];
            Ux[t] = r[1][t];
            Vx[t] = r[2][t];
            Wx[t] = r[3][t];
        }
        win.store(F.dX[0], Ux, c.len);
        win.store(F.dX[1], Vx, c.len);
        win.store(F.dX[2], Wx, c.len);
        win.store(F.dX[3], Tx, c.len);
        win.store(F.R[0], R0, c.len);
        win.store(F.R[1], Ru, c.len);
        win.store(F.R[2], Rv, c.len);
        win.store(F.R[3], Rw, c.len);
        win.store(F.R[4], RE, c.len);
    }
}

template <typename real, bool UNI>
hipError_t launch_xpass(const OpDev<real> *d1, const OpDev<real> *d2, const LaunchGeom &lg,
                        const LineMap &lm, const Fields<real> &F, const Phys<real> &ph,
                        int use_k, const StageArgs &sa, hipStream_t st)
{
    dim3 grid((lm.nlines + lg.LX - 1) / lg.LX), block(lg.threads);
    if (sa.stale)
        hipLaunchKernelGGL((k_xpass<real, SEG_DERIV, UNI, true>), grid, block, lg.shm, st, d1, d2,
                           lm, F, ph, use_k, sa);
    else
        hipLaunchKernelGGL((k_xpass<real, SEG_DERIV, UNI, false>), grid, block, lg.shm, st, d1,
                           d2, lm, F, ph, use_k, sa);
    return hipGetLastError();
}

#define CS3D_INST_X(real, UNI)                                                                  \
    template hipError_t launch_xpass<real, UNI>(const OpDev<real> *, const OpDev<real> *,          \
                                                const LaunchGeom &, const LineMap &,               \
                                                const Fields<real> &, const Phys<real> &, int,     \
                                                const StageArgs &, hipStream_t);
CS3D_INST_X(float, true)
CS3D_INST_X(float, false)
CS3D_INST_X(double, true)
CS3D_INST_X(double, false)

}  // namespace cs3d
