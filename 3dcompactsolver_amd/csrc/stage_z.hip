// stage_z.hip -- Z sweep of one RK stage (lines along z, MAP_COL lanes),
// with the pointwise assembly and the RK update fused in.
//
// z Euler fluxes (CSolver_AWS.cpp:1095-1108) and their derivatives
// (:1195-1221), first and second z-derivatives of U, V, W, T (:1113-1159),
// the mixed Uxz, Wxz, Vyz, Wyz (:1161-1192); then, per point, the terms of
// solveContinuity .. solveEnergy (:1246-1504) that are products of first
// derivatives, the sponge source (twice in the energy, :1491/:1495), the
// zeroing of sponge boundary planes (postStepBCHandling, :1606-1670), and
// the stage update of RK4::updateConservedData (RK4.hpp:45-111) or
// TVDRK3::updateConservedData (TVDRK3.hpp:225-260).
#include "stage_common.hpp"

namespace cs3d {

// QS: the update needs the stage state at the point (sponge source, or the
// TVDRK3 combinations of stages 2 and 3); the RK4 path without sponge never
// carries it.
template <typename real, int M, bool UNI, bool QS>
__global__ __launch_bounds__(sweep_threads<real>()) void k_zpass(
    const OpDev<real> *__restrict__ d1, const OpDev<real> *__restrict__ d2, LineMap lm,
    Fields<real> F, Phys<real> ph, StageArgs sa, const StepScalars *__restrict__ sc)
{
    constexpr int HW = 2, W = M + 2 * HW;
    CS3D_LANE_SETUP(MAP_COL, d1, sweep_kr(d1, d2), SWEEP_PLANES)
    const bool use_k = sa.rkStep > 1;
    const real *q0 = use_k ? F.qk[0] : F.q[0];
    const real *q1 = use_k ? F.qk[1] : F.q[1];
    const real *q2 = use_k ? F.qk[2] : F.q[2];
    const real *q3 = use_k ? F.qk[3] : F.q[3];
    const real *q4 = use_k ? F.qk[4] : F.q[4];
    // the conserved windows are read directly: staging them through LDS as
    // well (like the mixed windows below) measured slower for this kernel
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
    const real c43 = real(4) / real(3), c13 = real(1) / real(3), c23 = real(2) / real(3);
    // One batch of 13 solves from the stage state: the z fluxes (the h_ref W
    // part of the energy flux is added as h_ref Wz in the assembly), then
    // the first and second z-derivatives of T', U, V, W.  Every window dies
    // here, before the mixed derivatives are loaded.
    real r[13][M];
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
    rhs<OP_D2, real, M, HW, UNI>(d2, c, Tp, r[9]);
    rhs<OP_D2, real, M, HW, UNI>(d2, c, U, r[10]);
    rhs<OP_D2, real, M, HW, UNI>(d2, c, V, r[11]);
    rhs<OP_D2, real, M, HW, UNI>(d2, c, Wv, r[12]);
    real mu[M], amu[M], Uc[M], Vc[M], Wc[M];
#pragma unroll
    for (int t = 0; t < M; t++) {
        visc(ph, ph.T0s + Tp[HW + t], mu[t], amu[t]);
        Uc[t] = U[HW + t];
        Vc[t] = V[HW + t];
        Wc[t] = Wv[HW + t];
    }
    solve_batch2<9, 4, kc_deriv<real>(), real, M, UNI>(d1, d2, c, r);
    real R0[M], Ru[M], Rv[M], Rw[M], RE[M];
#pragma unroll
    for (int t = 0; t < M; t++) {
        RE[t] = -r[0][t];
        Rw[t] = -r[1][t];
        R0[t] = -r[2][t];
        Ru[t] = -r[3][t];
        Rv[t] = -r[4][t];
        RE[t] += ph.cpPr * (mu[t] * r[9][t]);
        Ru[t] += mu[t] * r[10][t];
        RE[t] += mu[t] * (Uc[t] * r[10][t]);
        Rv[t] += mu[t] * r[11][t];
        RE[t] += mu[t] * (Vc[t] * r[11][t]);
        Rw[t] += mu[t] * (c43 * r[12][t]);
        RE[t] += mu[t] * (Wc[t] * (c43 * r[12][t]));
    }
    {
        // mixed derivatives of the stored x- and y-derivatives.  Uxz and Vyz
        // enter the right-hand side with the same coefficients
        // (CSolver_AWS.cpp:1161-1192), so their sum is taken as one
        // derivative of Ux + Vy.
        real m3[3][M];
        if constexpr (UNI) {
            real mw[3][W];
            load_windows_lds<3>(c, win, [&](int fi, unsigned o) {
                return fi == 0 ? ldo(F.dX[0], o) + ldo(F.dY[1], o)
                               : ldo(fi == 1 ? F.dX[2] : F.dY[2], o);
            }, mw);
#pragma unroll
            for (int i = 0; i < 3; i++) rhs<OP_D1, real, M, HW, UNI>(d1, c, mw[i], m3[i]);
        } else {
            real f[W];
#pragma unroll
            for (int t = 0; t < W; t++) f[t] = ldo(F.dX[0], win.at(t)) + ldo(F.dY[1], win.at(t));
            rhs<OP_D1, real, M, HW, UNI>(d1, c, f, m3[0]);
            win.load(F.dX[2], f);
            rhs<OP_D1, real, M, HW, UNI>(d1, c, f, m3[1]);
            win.load(F.dY[2], f);
            rhs<OP_D1, real, M, HW, UNI>(d1, c, f, m3[2]);
        }
        solve_batch<3, kc_deriv<real>(), real, M, UNI>(d1, c, m3);
#pragma unroll
        for (int t = 0; t < M; t++) {
            Rw[t] += mu[t] * (c13 * m3[0][t]);
            RE[t] += mu[t] * (c13 * (Wc[t] * m3[0][t]));
            Ru[t] += mu[t] * (c13 * m3[1][t]);
            RE[t] += mu[t] * (c13 * (Uc[t] * m3[1][t]));
            Rv[t] += mu[t] * (c13 * m3[2][t]);
            RE[t] += mu[t] * (c13 * (Vc[t] * m3[2][t]));
        }
    }
    const real *Tz = r[5], *Uz = r[6], *Vz = r[7], *Wz = r[8];
    if (!active) return;

    // The point values of Ux, Vy, Wx, Wy are read again here rather than
    // kept from the mixed-derivative windows: copies of the field pointers
    // the compiler cannot identify with the originals stop it from holding
    // 16 window centres in registers across the solves.
    const real *pdX0 = F.dX[0], *pdX2 = F.dX[2], *pdY1 = F.dY[1], *pdY2 = F.dY[2];
    asm volatile("" : "+s"(pdX0), "+s"(pdX2), "+s"(pdY1), "+s"(pdY2));
    const real dt = (real)sc->dt;
    // stage-update coefficients, uniform over the launch
    const bool first = sa.rkStep == 1, has_next = sa.rkStep < 4;
    const real ra = (sa.rkStep == 1 || sa.rkStep == 4) ? real(1) / real(6) : real(1) / real(3);
    const real rb = sa.rkStep == 3 ? real(1) : real(0.5);
    const bool tvd_to_k = sa.rkStep < 3;
    const real ta = sa.rkStep == 1 ? real(1) : sa.rkStep == 2 ? real(0.75) : real(1) / real(3);
    const real tb = sa.rkStep == 1 ? real(0) : sa.rkStep == 2 ? real(0.25) : real(2) / real(3);
    const real tc = sa.rkStep == 1 ? real(1) : sa.rkStep == 2 ? real(0.25) : real(2) / real(3);
    // lines along z: L = i + Nx j; the point index along the line is k
    const int li = Lc % sa.Nx, lj = Lc / sa.Nx;
    const bool line_plane = (sa.sponge_face[0] && li == 0) ||
                            (sa.sponge_face[1] && li == sa.Nx - 1) ||
                            (sa.sponge_face[2] && lj == 0) ||
                            (sa.sponge_face[3] && lj == sa.Ny - 1);
#pragma unroll
    for (int t = 0; t < M; t++) {
        if (!UNI && t >= c.len) continue;
        const unsigned i = win.mid(t);
        const int kk = c.s + t;
        const real Ux = ldo(pdX0, i), Vx = ldo(F.dX[1], i), Wx = ldo(pdX2, i),
                   Tx = ldo(F.dX[3], i);
        const real Uy = ldo(F.dY[0], i), Vy = ldo(pdY1, i), Wy = ldo(pdY2, i),
                   Ty = ldo(F.dY[3], i);
        const real uz = Uz[t], vz = Vz[t], wz = Wz[t];
        const real MuX = amu[t] * Tx, MuY = amu[t] * Ty, MuZ = amu[t] * Tz[t];
        const real Uu = Uc[t], Vv = Vc[t], Ww = Wc[t], m_ = mu[t];
        real su = c43 * MuX * (Ux - real(0.5) * Vy - real(0.5) * wz) + MuY * (Uy + Vx) +
                  MuZ * (Wx + uz);
        real sv = c43 * MuY * (Vy - real(0.5) * Ux - real(0.5) * wz) + MuX * (Uy + Vx) +
                  MuZ * (Wy + vz);
        real sw = c43 * MuZ * (wz - real(0.5) * Ux - real(0.5) * Vy) + MuX * (Wx + uz) +
                  MuY * (Wy + vz);
        real qn = ph.cpPr * (MuX * Tx + MuY * Ty + MuZ * Tz[t]);
        real quad = c43 * (Ux * Ux + Vy * Vy + wz * wz) + Uy * Uy + uz * uz + Vx * Vx +
                    vz * vz + Wx * Wx + Wy * Wy - c43 * (Ux * Vy + Ux * wz + Vy * wz) +
                    real(2) * (Uy * Vx + uz * Wx + vz * Wy);
        real vt2 = c43 * (Uu * MuX * Ux + Vv * MuY * Vy + Ww * MuZ * wz) -
                   c23 * Uu * MuX * (Vy + wz) - c23 * Vv * MuY * (Ux + wz) -
                   c23 * Ww * MuZ * (Ux + Vy) + Uu * MuY * (Uy + Vx) + Uu * MuZ * (uz + Wx) +
                   Vv * MuX * (Uy + Vx) + Vv * MuZ * (vz + Wy) + Ww * MuX * (uz + Wx) +
                   Ww * MuY * (vz + Wy);
        real rhs[5];
        rhs[0] = ldo(F.R[0], i) + R0[t];
        rhs[1] = ldo(F.R[1], i) + Ru[t] + su;
        rhs[2] = ldo(F.R[2], i) + Rv[t] + sv;
        rhs[3] = ldo(F.R[3], i) + Rw[t] + sw;
        rhs[4] = ldo(F.R[4], i) + RE[t] - ph.h_ref * wz + qn + m_ * quad + vt2;
        real qs_[5] = {0, 0, 0, 0, 0};
        if (QS) {
            qs_[0] = ldo(q0, i);
            qs_[1] = ldo(q1, i);
            qs_[2] = ldo(q2, i);
            qs_[3] = ldo(q3, i);
            qs_[4] = ldo(q4, i);
        }
        if (QS && F.sigma) {
            const real sg = ldo(F.sigma, i);
#pragma unroll
            for (int v = 0; v < 5; v++) rhs[v] += sg * (ldo(F.avg[v], i) - qs_[v]);
            rhs[4] += sg * (ldo(F.avg[4], i) - qs_[4]);
        }
        real k[5];
#pragma unroll
        for (int v = 0; v < 5; v++) k[v] = dt * rhs[v];
        {
            const bool plane = line_plane || (sa.sponge_face[4] && kk == 0) ||
                               (sa.sponge_face[5] && kk == sa.Nz - 1);
            if (plane) {
#pragma unroll
                for (int v = 0; v < 5; v++) k[v] = 0;
            }
        }
        if (sa.diag) {
            // per-point inputs of the step diagnostics that come from this
            // stage's derivatives (CSolver_AWS.cpp:2190-2225)
            const real dil = Ux + Vy + wz;
            const real S00 = real(0.5) * (Ux + Ux) - c13 * dil;
            const real S11 = real(0.5) * (Vy + Vy) - c13 * dil;
            const real S22 = real(0.5) * (wz + wz) - c13 * dil;
            const real S01 = real(0.5) * (Uy + Vx), S02 = real(0.5) * (uz + Wx),
                       S12 = real(0.5) * (vz + Wy);
            real td = S00 * Ux;
            td += S01 * Uy;
            td += S02 * uz;
            td += S01 * Vx;
            td += S11 * Vy;
            td += S12 * vz;
            td += S02 * Wx;
            td += S12 * Wy;
            td += S22 * wz;
            const real wx_ = Wy - vz, wy_ = uz - Wx, wz_ = Vx - Uy;
            sto(F.diag[0], i, td);
            sto(F.diag[1], i, wx_ * wx_ + wy_ * wy_ + wz_ * wz_);
            sto(F.diag[2], i, dil);
            sto(F.diag[3], i, (Ux * Ux + Vy * Vy + wz * wz) * c13);
        }
        if (sa.mode == 1) {
#pragma unroll
            for (int v = 0; v < 5; v++) sto(F.k2[v], i, k[v]);
            continue;
        }
        if (sa.rk == 4) {
            // RK4.hpp:45-111: *2 accumulates the stages, *k is the next
            // stage state
#pragma unroll
            for (int v = 0; v < 5; v++) {
                const real qn1 = ldo(F.q[v], i);
                const real a0 = first ? qn1 : ldo(F.acc[v], i);
                sto(F.acc[v], i, a0 + ra * k[v]);
                if (has_next) sto(F.qk[v], i, qn1 + rb * k[v]);
            }
        } else {
            // TVDRK3.hpp:225-260: stages 1-2 write *k, stage 3 writes *2
#pragma unroll
            for (int v = 0; v < 5; v++) {
                const real qn1 = ldo(F.q[v], i);
                sto(tvd_to_k ? F.qk[v] : F.acc[v], i, ta * qn1 + tb * qs_[v] + tc * k[v]);
            }
        }
    }
}

template <typename real, bool UNI>
hipError_t launch_zpass(const OpDev<real> *d1, const OpDev<real> *d2, const LaunchGeom &lg,
                        const LineMap &lm, const Fields<real> &F, const Phys<real> &ph,
                        const StageArgs &sa, const StepScalars *sc, hipStream_t st)
{
    dim3 grid((lm.nlines + lg.LX - 1) / lg.LX), block(lg.threads);
    if (F.sigma || sa.rk == 3)
        hipLaunchKernelGGL((k_zpass<real, SEG_DERIV, UNI, true>), grid, block, lg.shm, st, d1, d2,
                           lm, F, ph, sa, sc);
    else
        hipLaunchKernelGGL((k_zpass<real, SEG_DERIV, UNI, false>), grid, block, lg.shm, st, d1,
                           d2, lm, F, ph, sa, sc);
    return hipGetLastError();
}

#define CS3D_INST_Z(real, UNI)                                                                  \
    template hipError_t launch_zpass<real, UNI>(const OpDev<real> *, const OpDev<real> *,          \
                                                const LaunchGeom &, const LineMap &,               \
                                                const Fields<real> &, const Phys<real> &,          \
                                                const StageArgs &, const StepScalars *,            \
                                                hipStream_t);
CS3D_INST_Z(float, true)
CS3D_INST_Z(float, false)
CS3D_INST_Z(double, true)
CS3D_INST_Z(double, false)

}  // namespace cs3d
