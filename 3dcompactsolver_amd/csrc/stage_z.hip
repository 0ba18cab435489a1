// stage_z.hip -- middle sweep of one RK stage: lines along z (MAP_COL
// lanes).
//
// From the stage state: the z Euler fluxes (CSolver_AWS.cpp:1095-1108) and
// their derivatives (:1195-1221), the first and second z-derivatives of U,
// V, W, T (:1113-1159), and the mixed Vyz, Wyz as z-derivatives of the
// stored Vy, Wy (:1161-1192).  Stores Uz Vz Wz Tz (as the pairs (Uz, Wz),
// (Vz, Tz)) and its own partial right-hand side (continuity and momenta as
// one quad per point, energy apart): every term linear in one of these.
// Uxz and Wxz are taken
// along x of the stored Uz, Wz in the last sweep (operators along different
// axes commute).  Solves: one batch of 13 (5 flux + 4 first + 4 second
// derivatives), then the 2 mixed ones; FP64 on uniform lines all 15 in one
// batch (its 8-line ring has the room; the second batch's barriers and the
// staging of (Vy, Wy) cost 5% of that sweep).
//
// STALE (first stage after updateSponge, see stage_y.hip): the reference's
// Uxz, Wxz are z-derivatives of x-derivatives of the stale primitives; the
// STALE variant also solves D_z of the stale U and W (into F.mix[2],
// F.mix[3]).
#include "stage_common.hpp"

namespace cs3d {

template <typename real, int M, bool UNI, bool STALE>
__global__ __launch_bounds__((col_threads<real, 2, UNI, STALE>())) void k_zpass(
    const OpDev<real> *__restrict__ d1, const OpDev<real> *__restrict__ d2, LineMap lm,
    Fields<real> F, Phys<real> ph, int use_k, StageArgs sa)
{
    constexpr int HW = 2, W = M + 2 * HW;
    // MRG: the mixed Vyz, Wyz join the first batch (15 solves, one batch
    // and no staging of (Vy, Wy) fewer); the FP64 ring of 8 lines has room
    constexpr bool MRG = z_merged<real>(UNI, STALE);
    CS3D_LANE_SETUP(MAP_COL, d1, sweep_kr(d1, d2), MRG ? SWEEP_PLANES + 2 : SWEEP_PLANES)
    const QSet<real> qs = use_k ? F.qk : F.q;
    // the solves of the first batch: 0-4 fluxes, 5-8 first derivatives of
    // T', U, V, W, (MRG: 9-10 Vyz, Wyz,) then their second derivatives
    constexpr int NB = MRG ? 15 : 13, I2 = NB - 4;
    real r[NB][M];
    if constexpr (MRG) {
        real mw[2][W];
        win.load_pair(F.dYvw, mw[0], mw[1]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, mw[0], r[9]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, mw[1], r[10]);
    }
    // the conserved windows are read directly: staging them through LDS
    // (as the y sweep does) measured slower here
    real a0[W], a1[W], a2[W], a3[W], a4[W];
    real ir[W], U[W], V[W], Wv[W], pp[W], Tp[W];
    win.load_quad(qs.m, a0, a1, a2, a3);
    win.load(qs.e, a4);
    // the own points of (Vy, Wy) for the mixed derivatives, issued now so
    // they arrive during the first solve batch (not in the STALE variant or
    // in FP64, whose registers are full)
    constexpr bool PRE = UNI && !STALE && sizeof(real) == 4;
    real pv[2][M];
    if constexpr (PRE) {
#pragma unroll
        for (int t = 0; t < M; t++) {
            pv[0][t] = ldo_part<2>(F.dYvw, win.mid(t), 0);
            pv[1][t] = ldo_part<2>(F.dYvw, win.mid(t), 1);
        }
    }
#pragma unroll
    for (int t = 0; t < W; t++)
        prim(ph, a0[t], a1[t], a2[t], a3[t], a4[t], ir[t], U[t], V[t], Wv[t], pp[t], Tp[t]);
    const real c43 = real(4) / real(3), c13 = real(1) / real(3);
    // the stale U and W of the sponge planes (lines along z: L = i + Nx j,
    // the point index is k)
    constexpr int NS = STALE ? 2 : 0;
    real mst[NS > 0 ? NS : 1][M];
    if constexpr (STALE) {
        const int li = Lc % sa.Nx, lj = Lc / sa.Nx;
        const bool line_plane = (sa.sponge_face[0] && li == 0) ||
                                (sa.sponge_face[1] && li == sa.Nx - 1) ||
                                (sa.sponge_face[2] && lj == 0) ||
                                (sa.sponge_face[3] && lj == sa.Ny - 1);
        real us[W], ws[W];
#pragma unroll
        for (int t = 0; t < W; t++) {
            int p = c.s - HW + t;
            p += p < 0 ? n : 0;
            p -= p >= n ? n : 0;
            const bool plane = line_plane || (sa.sponge_face[4] && p == 0) ||
                               (sa.sponge_face[5] && p == n - 1);
            us[t] = U[t];
            ws[t] = Wv[t];
            if (plane) {
                real i_, u_, v_, w_, p_, t_, qo[5];
                qset_load_at(F.qstale, win.at(t), qo);
                prim(ph, qo[0], qo[1], qo[2], qo[3], qo[4], i_, u_, v_, w_, p_, t_);
                us[t] = u_;
                ws[t] = w_;
            }
        }
        rhs<OP_D1, real, M, HW, UNI>(d1, c, us, mst[0]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, ws, mst[1]);
    }
    // 13 solves from the stage state: the z fluxes (the h_ref W part of the
    // energy flux is added as h_ref Wz), then the first and second
    // z-derivatives of T', U, V, W
    {
        real f[W];
#pragma unroll
        for (int t = 0; t < W; t++) f[t] = (a4[t] + pp[t]) * Wv[t];
        rhs<OP_D1, real, M, HW, UNI>(d1, c, f, r[0]);
#pragma unroll
        for (int t = 0; t < W; t++) f[t] = a3[t] * Wv[t] + pp[t];
        rhs<OP_D1, real, M, HW, UNI>(d1, c, f, r[1]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, a3, r[2]);
#pragma unroll
        for (int t = 0; t < W; t++) f[t] = a1[t] * Wv[t];
        rhs<OP_D1, real, M, HW, UNI>(d1, c, f, r[3]);
#pragma unroll
        for (int t = 0; t < W; t++) f[t] = a2[t] * Wv[t];
        rhs<OP_D1, real, M, HW, UNI>(d1, c, f, r[4]);
    }
    rhs<OP_D1, real, M, HW, UNI>(d1, c, Tp, r[5]);
    rhs<OP_D1, real, M, HW, UNI>(d1, c, U, r[6]);
    rhs<OP_D1, real, M, HW, UNI>(d1, c, V, r[7]);
    rhs<OP_D1, real, M, HW, UNI>(d1, c, Wv, r[8]);
    rhs<OP_D2, real, M, HW, UNI>(d2, c, Tp, r[I2]);
    rhs<OP_D2, real, M, HW, UNI>(d2, c, U, r[I2 + 1]);
    rhs<OP_D2, real, M, HW, UNI>(d2, c, V, r[I2 + 2]);
    rhs<OP_D2, real, M, HW, UNI>(d2, c, Wv, r[I2 + 3]);
    real mu[M], Uc[M], Vc[M], Wc[M];
#pragma unroll
    for (int t = 0; t < M; t++) {
        real amu;
        visc(ph, ph.T0s + Tp[HW + t], mu[t], amu);
        Uc[t] = U[HW + t];
        Vc[t] = V[HW + t];
        Wc[t] = Wv[HW + t];
    }
    static_assert(NB <= (MRG ? SWEEP_PLANES + 2 : SWEEP_PLANES) && 2 + NS <= SWEEP_PLANES,
                  "batch wider than the exchange ring");
    solve_d12<NB - 4, 4, real, M, UNI>(d1, d2, c, r);
    real R0[M], Ru[M], Rv[M], Rw[M], RE[M];
#pragma unroll
    for (int t = 0; t < M; t++) {
        RE[t] = -r[0][t];
        Rw[t] = -r[1][t];
        R0[t] = -r[2][t];
        Ru[t] = -r[3][t];
        Rv[t] = -r[4][t];
        RE[t] += ph.cpPr * (mu[t] * r[I2][t]);
        RE[t] -= ph.h_ref * r[8][t];
        Ru[t] += mu[t] * r[I2 + 1][t];
        RE[t] += mu[t] * (Uc[t] * r[I2 + 1][t]);
        Rv[t] += mu[t] * r[I2 + 2][t];
        RE[t] += mu[t] * (Vc[t] * r[I2 + 2][t]);
        Rw[t] += mu[t] * (c43 * r[I2 + 3][t]);
        RE[t] += mu[t] * (Wc[t] * (c43 * r[I2 + 3][t]));
    }
    if (active) {
        win.template store_pair<true>(F.dZuw, r[6], r[8], c.len);
        win.template store_pair<true>(F.dZvt, r[7], r[5], c.len);
    }
    if constexpr (MRG) {
#pragma unroll
        for (int t = 0; t < M; t++) {
            Rw[t] += mu[t] * (c13 * r[9][t]);
            RE[t] += mu[t] * (c13 * (Wc[t] * r[9][t]));
            Rv[t] += mu[t] * (c13 * r[10][t]);
            RE[t] += mu[t] * (c13 * (Vc[t] * r[10][t]));
        }
    } else {
        // Vyz, Wyz from the stored Vy, Wy (and the stale D_z U, D_z W)
        constexpr int NM = 2 + NS;
        real m[NM][M];
        real mw[2][W];
        if constexpr (PRE)
            load_windows_lds_pt<2, real, M, HW>(c, [&](int t, real(&v)[2]) {
                v[0] = pv[0][t];
                v[1] = pv[1][t];
            }, mw);
        else if constexpr (UNI)
            load_windows_lds_pair(c, win, F.dYvw, mw);
        else
            win.load_pair(F.dYvw, mw[0], mw[1]);
#pragma unroll
        for (int i = 0; i < 2; i++) rhs<OP_D1, real, M, HW, UNI>(d1, c, mw[i], m[i]);
        if constexpr (STALE) {
#pragma unroll
            for (int t = 0; t < M; t++) {
                m[2][t] = mst[0][t];
                m[3][t] = mst[1][t];
            }
        }
        solve_batch<NM, kc_deriv<real>(), real, M, UNI>(d1, c, m);
#pragma unroll
        for (int t = 0; t < M; t++) {
            Rw[t] += mu[t] * (c13 * m[0][t]);
            RE[t] += mu[t] * (c13 * (Wc[t] * m[0][t]));
            Rv[t] += mu[t] * (c13 * m[1][t]);
            RE[t] += mu[t] * (c13 * (Vc[t] * m[1][t]));
        }
        if constexpr (STALE) {
            if (active) {
                win.template store<true>(F.mix[2], m[2], c.len);
                win.template store<true>(F.mix[3], m[3], c.len);
            }
        }
    }
    if (!active) return;
    // a partial right-hand side of its own (the last sweep adds both): no
    // read-modify-write on the column path
    win.template store_quad<true>(F.Rz4, R0, Ru, Rv, Rw, c.len);
    win.template store<true>(F.RzE, RE, c.len);
}

template <typename real, bool UNI>
hipError_t launch_zpass(const OpDev<real> *d1, const OpDev<real> *d2, int M, const LaunchGeom &lg,
                        const LineMap &lm, const Fields<real> &F, const Phys<real> &ph,
                        int use_k, const StageArgs &sa, hipStream_t st)
{
    dim3 grid((lm.nlines + lg.LX - 1) / lg.LX), block(lg.threads);
    const int cap = sa.stale ? col_threads<real, 2, UNI, true>() : col_threads<real, 2, UNI, false>();
    if (lg.threads > cap) return hipErrorInvalidConfiguration;
    if constexpr (UNI) {
        if (M == SEG_DERIV3) {
            if (sa.stale)
                hipLaunchKernelGGL((k_zpass<real, SEG_DERIV3, true, true>), grid, block, lg.shm, st,
                                   d1, d2, lm, F, ph, use_k, sa);
            else
                hipLaunchKernelGGL((k_zpass<real, SEG_DERIV3, true, false>), grid, block, lg.shm,
                                   st, d1, d2, lm, F, ph, use_k, sa);
            return hipGetLastError();
        }
    }
    if (M != SEG_DERIV) return hipErrorInvalidValue;
    if (sa.stale)
        hipLaunchKernelGGL((k_zpass<real, SEG_DERIV, UNI, true>), grid, block, lg.shm, st, d1, d2,
                           lm, F, ph, use_k, sa);
    else
        hipLaunchKernelGGL((k_zpass<real, SEG_DERIV, UNI, false>), grid, block, lg.shm, st, d1,
                           d2, lm, F, ph, use_k, sa);
    return hipGetLastError();
}

#define CS3D_INST_Z(real, UNI)                                                                  \
    template hipError_t launch_zpass<real, UNI>(const OpDev<real> *, const OpDev<real> *, int,     \
                                                const LaunchGeom &, const LineMap &,               \
                                                const Fields<real> &, const Phys<real> &, int,     \
                                                const StageArgs &, hipStream_t);
CS3D_INST_Z(float, true)
CS3D_INST_Z(float, false)
CS3D_INST_Z(double, true)
CS3D_INST_Z(double, false)

}  // namespace cs3d
