// stage_y.hip -- first sweep of one RK stage: lines along y (MAP_COL lanes,
// every load instruction reads consecutive x).
//
// From the stage state: the y Euler fluxes (CSolver_AWS.cpp:949-963) and
// their derivatives (:1051-1073), the first and second y-derivatives of U,
// V, W, T (:969-1049).  Stores Uy Vy Wy Ty (as the pairs (Uy, Ty),
// (Vy, Wy)) and starts the right-hand side with every term linear in one of
// these (continuity and momenta as one quad per point, energy apart).  The mixed derivatives the
// reference takes along y of x-derivatives (Uxy, Vxy, :991-1003) are taken
// along x of the stored Uy, Vy in the last sweep: compact operators along
// different axes commute.  The 13 line solves run as one batch.
//
// STALE: at the first stage after updateSponge the reference's x block uses
// primitives that predate the sponge reset on the sponge planes
// (CSolver_AWS.cpp:861-912 with :1913-1925 and :1955-2021), so its Uxy, Vxy
// are y-derivatives of x-derivatives of those.  The STALE variant also
// solves D_y of the stale U and V (into F.mix[0], F.mix[1]) for the last
// sweep to differentiate along x.
#include "stage_common.hpp"

namespace cs3d {

template <typename real, int M, bool UNI, bool STALE>
__global__ __launch_bounds__((col_threads<real, 1, UNI, STALE>())) void k_ypass(
    const OpDev<real> *__restrict__ d1, const OpDev<real> *__restrict__ d2, LineMap lm,
    Fields<real> F, Phys<real> ph, int use_k, StageArgs sa)
{
    constexpr int HW = 2, W = M + 2 * HW;
    CS3D_LANE_SETUP(MAP_COL, d1, sweep_kr(d1, d2), SWEEP_PLANES)
    const QSet<real> qs = use_k ? F.qk : F.q;
    real a[5][W];
    real(&a0)[W] = a[0];
    real(&a1)[W] = a[1];
    real(&a2)[W] = a[2];
    real(&a3)[W] = a[3];
    real(&a4)[W] = a[4];
    real ir[W], U[W], V[W], Wv[W], pp[W], Tp[W];
    // (staging the halos through LDS measured 3% slower than these direct
    // window loads: its barriers cost more than the halo reads)
    win.load_quad(qs.m, a0, a1, a2, a3);
    win.load(qs.e, a4);
#pragma unroll
    for (int t = 0; t < W; t++)
        prim(ph, a0[t], a1[t], a2[t], a3[t], a4[t], ir[t], U[t], V[t], Wv[t], pp[t], Tp[t]);
    const real c43 = real(4) / real(3);
    real R0[M], Ru[M], Rv[M], Rw[M], RE[M];
    // the 5 y fluxes (the h_ref V part of the energy flux is added as
    // h_ref Vy), STALE: the first derivatives of the stale U, V, then the
    // first and second derivatives of T', U, V, W.  One batch of 13 solves;
    // the STALE variant's 15 exceed the exchange ring's planes and run as
    // 7 + 8.
    constexpr int NS = STALE ? 2 : 0, I1 = 5 + NS, N1 = I1 + 4;
    static_assert((STALE ? 8 : N1 + 4) <= SWEEP_PLANES, "batch wider than the exchange ring");
    real r[N1 + 4][M];
    {
        real f[W];
#pragma unroll
        for (int t = 0; t < W; t++) f[t] = (a4[t] + pp[t]) * V[t];
        rhs<OP_D1, real, M, HW, UNI>(d1, c, f, r[0]);
#pragma unroll
        for (int t = 0; t < W; t++) f[t] = a2[t] * V[t] + pp[t];
        rhs<OP_D1, real, M, HW, UNI>(d1, c, f, r[1]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, a2, r[2]);
#pragma unroll
        for (int t = 0; t < W; t++) f[t] = a1[t] * V[t];
        rhs<OP_D1, real, M, HW, UNI>(d1, c, f, r[3]);
#pragma unroll
        for (int t = 0; t < W; t++) f[t] = a3[t] * V[t];
        rhs<OP_D1, real, M, HW, UNI>(d1, c, f, r[4]);
    }
    if constexpr (STALE) {
        // U and V with the sponge planes' primitives from the pre-reset
        // copy; lines along y: L = i + Nx k, the point index is j
        const int li = Lc % sa.Nx, lk = Lc / sa.Nx;
        const bool line_plane = (sa.sponge_face[0] && li == 0) ||
                                (sa.sponge_face[1] && li == sa.Nx - 1) ||
                                (sa.sponge_face[4] && lk == 0) ||
                                (sa.sponge_face[5] && lk == sa.Nz - 1);
        real us[W], vs[W];
#pragma unroll
        for (int t = 0; t < W; t++) {
            int p = c.s - HW + t;
            p += p < 0 ? n : 0;
            p -= p >= n ? n : 0;
            const bool plane = line_plane || (sa.sponge_face[2] && p == 0) ||
                               (sa.sponge_face[3] && p == n - 1);
            us[t] = U[t];
            vs[t] = V[t];
            if (plane) {
                real i_, u_, v_, w_, p_, t_, qo[5];
                qset_load_at(F.qstale, win.at(t), qo);
                prim(ph, qo[0], qo[1], qo[2], qo[3], qo[4], i_, u_, v_, w_, p_, t_);
                us[t] = u_;
                vs[t] = v_;
            }
        }
        rhs<OP_D1, real, M, HW, UNI>(d1, c, us, r[5]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, vs, r[6]);
    }
    rhs<OP_D1, real, M, HW, UNI>(d1, c, Tp, r[I1 + 0]);
    rhs<OP_D1, real, M, HW, UNI>(d1, c, U, r[I1 + 1]);
    rhs<OP_D1, real, M, HW, UNI>(d1, c, V, r[I1 + 2]);
    rhs<OP_D1, real, M, HW, UNI>(d1, c, Wv, r[I1 + 3]);
    rhs<OP_D2, real, M, HW, UNI>(d2, c, Tp, r[N1 + 0]);
    rhs<OP_D2, real, M, HW, UNI>(d2, c, U, r[N1 + 1]);
    rhs<OP_D2, real, M, HW, UNI>(d2, c, V, r[N1 + 2]);
    rhs<OP_D2, real, M, HW, UNI>(d2, c, Wv, r[N1 + 3]);
    if constexpr (STALE) {
        solve_batch<I1, kc_deriv<real>(), real, M, UNI>(d1, c, *(real(*)[I1][M])r);
        solve_d12<4, 4, real, M, UNI>(d1, d2, c, *(real(*)[8][M])(r + I1));
        if (active) {
            win.template store<true>(F.mix[0], r[5], c.len);
            win.template store<true>(F.mix[1], r[6], c.len);
        }
    } else {
        solve_d12<N1, 4, real, M, UNI>(d1, d2, c, r);
    }
#pragma unroll
    for (int t = 0; t < M; t++) {
        RE[t] = -r[0][t];
        Rv[t] = -r[1][t];
        R0[t] = -r[2][t];
        Ru[t] = -r[3][t];
        Rw[t] = -r[4][t];
        real mu, amu;
        visc(ph, ph.T0s + Tp[HW + t], mu, amu);
        const real Uc = U[HW + t], Vc = V[HW + t], Wc = Wv[HW + t];
        RE[t] += ph.cpPr * (mu * r[N1][t]);
        Ru[t] += mu * r[N1 + 1][t];
        RE[t] += mu * (Uc * r[N1 + 1][t]);
        RE[t] -= ph.h_ref * r[I1 + 2][t];
        Rv[t] += mu * (c43 * r[N1 + 2][t]);
        RE[t] += mu * (Vc * (c43 * r[N1 + 2][t]));
        Rw[t] += mu * r[N1 + 3][t];
        RE[t] += mu * (Wc * r[N1 + 3][t]);
    }
    if (active) {
        win.template store_pair<true>(F.dYut, r[I1 + 1], r[I1 + 0], c.len);
        win.template store_pair<true>(F.dYvw, r[I1 + 2], r[I1 + 3], c.len);
        win.template store_quad<true>(F.Ry4, R0, Ru, Rv, Rw, c.len);
        win.template store<true>(F.RyE, RE, c.len);
    }
}

template <typename real, bool UNI>
hipError_t launch_ypass(const OpDev<real> *d1, const OpDev<real> *d2, const LaunchGeom &lg,
                        const LineMap &lm, const Fields<real> &F, const Phys<real> &ph,
                        int use_k, const StageArgs &sa, hipStream_t st)
{
    dim3 grid((lm.nlines + lg.LX - 1) / lg.LX), block(lg.threads);
    const int cap = sa.stale ? col_threads<real, 1, UNI, true>() : col_threads<real, 1, UNI, false>();
    if (lg.threads > cap) return hipErrorInvalidConfiguration;
    if (sa.stale)
        hipLaunchKernelGGL((k_ypass<real, SEG_DERIV, UNI, true>), grid, block, lg.shm, st, d1, d2,
                           lm, F, ph, use_k, sa);
    else
        hipLaunchKernelGGL((k_ypass<real, SEG_DERIV, UNI, false>), grid, block, lg.shm, st, d1,
                           d2, lm, F, ph, use_k, sa);
    return hipGetLastError();
}

#define CS3D_INST_Y(real, UNI)                                                                  \
    template hipError_t launch_ypass<real, UNI>(const OpDev<real> *, const OpDev<real> *,          \
                                                const LaunchGeom &, const LineMap &,               \
                                                const Fields<real> &, const Phys<real> &, int,     \
                                                const StageArgs &, hipStream_t);
CS3D_INST_Y(float, true)
CS3D_INST_Y(float, false)
CS3D_INST_Y(double, true)
CS3D_INST_Y(double, false)

}  // namespace cs3d
