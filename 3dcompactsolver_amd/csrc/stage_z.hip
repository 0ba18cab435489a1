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

template <typename real, int M, bool UNI>
__global__ __launch_bounds__(sweep_threads<real>()) void k_zpass(
    const OpDev<real> *__restrict__ d1, const OpDev<real> *__restrict__ d2, LineMap lm,
    Fields<real> F, Phys<real> ph, StageArgs sa, const StepScalars *__restrict__ sc)
{
    CS3D_LANE_SETUP(MAP_COL, d1)
    constexpr int HW = 2, W = M + 2 * HW;
    const bool use_k = sa.rkStep > 1;
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
    for (int t = 0; t < W; t++) w[t] = (a4[t] + pp[t]) * CS3D_W(t);
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) RE[t] = -r[t];
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = a3[t] * CS3D_W(t) + pp[t];
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) Rw[t] = -r[t];
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, a3, r);
#pragma unroll
    for (int t = 0; t < M; t++) R0[t] = -r[t];
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = a1[t] * CS3D_W(t);
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) Ru[t] = -r[t];
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = a2[t] * CS3D_W(t);
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) Rv[t] = -r[t];

    real mu[M], amu[M], Uc[M], Vc[M], Wc[M];
    real Uz[M], Vz[M], Wz[M], Tz[M];
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = (pp[t] - a0[t] * ph.RT0s) * ir[t] * ph.iR;
#pragma unroll
    for (int t = 0; t < M; t++) {
        visc(ph, ph.T0s + w[HW + t], mu[t], amu[t]);
        Uc[t] = CS3D_U(HW + t);
        Vc[t] = CS3D_V(HW + t);
        Wc[t] = CS3D_W(HW + t);
    }
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, Tz);
    compact_op<OP_D2, real, M, HW, UNI, MAP_COL>(d2, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) RE[t] += ph.cpPr * (mu[t] * r[t]);
    const real c43 = real(4) / real(3), c13 = real(1) / real(3), c23 = real(2) / real(3);
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = CS3D_U(t);
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, Uz);
    compact_op<OP_D2, real, M, HW, UNI, MAP_COL>(d2, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) {
        Ru[t] += mu[t] * r[t];
        RE[t] += mu[t] * (Uc[t] * r[t]);
    }
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = CS3D_V(t);
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, Vz);
    compact_op<OP_D2, real, M, HW, UNI, MAP_COL>(d2, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) {
        Rv[t] += mu[t] * r[t];
        RE[t] += mu[t] * (Vc[t] * r[t]);
    }
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = CS3D_W(t);
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, Wz);
    compact_op<OP_D2, real, M, HW, UNI, MAP_COL>(d2, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) {
        Rw[t] += mu[t] * (c43 * r[t]);
        RE[t] += mu[t] * (Wc[t] * (c43 * r[t]));
    }
    // mixed derivatives from the stored x- and y-derivatives
    load_window<real, M, HW>([&](int p) { return F.dX[0][base + p * sp]; }, c.s, n, per, w);
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, r);  // Uxz
#pragma unroll
    for (int t = 0; t < M; t++) {
        Rw[t] += mu[t] * (c13 * r[t]);
        RE[t] += mu[t] * (c13 * (Wc[t] * r[t]));
    }
    load_window<real, M, HW>([&](int p) { return F.dX[2][base + p * sp]; }, c.s, n, per, w);
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, r);  // Wxz
#pragma unroll
    for (int t = 0; t < M; t++) {
        Ru[t] += mu[t] * (c13 * r[t]);
        RE[t] += mu[t] * (c13 * (Uc[t] * r[t]));
    }
    load_window<real, M, HW>([&](int p) { return F.dY[1][base + p * sp]; }, c.s, n, per, w);
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, r);  // Vyz
#pragma unroll
    for (int t = 0; t < M; t++) {
        Rw[t] += mu[t] * (c13 * r[t]);
        RE[t] += mu[t] * (c13 * (Wc[t] * r[t]));
    }
    load_window<real, M, HW>([&](int p) { return F.dY[2][base + p * sp]; }, c.s, n, per, w);
    compact_op<OP_D1, real, M, HW, UNI, MAP_COL>(d1, c, w, r);  // Wyz
#pragma unroll
    for (int t = 0; t < M; t++) {
        Rv[t] += mu[t] * (c13 * r[t]);
        RE[t] += mu[t] * (c13 * (Vc[t] * r[t]));
    }
    if (!active) return;

    const real dt = (real)sc->dt;
    const int Nx = sa.Nx, Ny = sa.Ny;
    const long Nxy = (long)Nx * Ny;
#pragma unroll
    for (int t = 0; t < M; t++) {
        if (t >= c.len) continue;
        const long i = base + (long)(c.s + t) * sp;
        const real Ux = F.dX[0][i], Vx = F.dX[1][i], Wx = F.dX[2][i], Tx = F.dX[3][i];
        const real Uy = F.dY[0][i], Vy = F.dY[1][i], Wy = F.dY[2][i], Ty = F.dY[3][i];
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
        rhs[0] = F.R[0][i] + R0[t];
        rhs[1] = F.R[1][i] + Ru[t] + su;
        rhs[2] = F.R[2][i] + Rv[t] + sv;
        rhs[3] = F.R[3][i] + Rw[t] + sw;
        rhs[4] = F.R[4][i] + RE[t] - ph.h_ref * wz + qn + m_ * quad + vt2;
        real qs_[5];
        if (F.sigma || (sa.rk == 3 && sa.rkStep > 1)) {
            qs_[0] = q0[i];
            qs_[1] = q1[i];
            qs_[2] = q2[i];
            qs_[3] = q3[i];
            qs_[4] = q4[i];
        }
        if (F.sigma) {
            const real sg = F.sigma[i];
#pragma unroll
            for (int v = 0; v < 5; v++) rhs[v] += sg * (F.avg[v][i] - qs_[v]);
            rhs[4] += sg * (F.avg[4][i] - qs_[4]);
        }
        real k[5];
#pragma unroll
        for (int v = 0; v < 5; v++) k[v] = dt * rhs[v];
        {
            const int ii = (int)(i % Nx), jj = (int)((i / Nx) % Ny), kk = (int)(i / Nxy);
            const bool plane = (sa.sponge_face[0] && ii == 0) || (sa.sponge_face[1] && ii == Nx - 1) ||
                               (sa.sponge_face[2] && jj == 0) || (sa.sponge_face[3] && jj == Ny - 1) ||
                               (sa.sponge_face[4] && kk == 0) || (sa.sponge_face[5] && kk == sa.Nz - 1);
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
            F.diag[0][i] = td;
            F.diag[1][i] = wx_ * wx_ + wy_ * wy_ + wz_ * wz_;
            F.diag[2][i] = dil;
            F.diag[3][i] = (Ux * Ux + Vy * Vy + wz * wz) / real(3);
        }
        if (sa.mode == 1) {
#pragma unroll
            for (int v = 0; v < 5; v++) F.k2[v][i] = k[v];
            continue;
        }
        if (sa.rk == 4) {
            const real aa = (sa.rkStep == 1 || sa.rkStep == 4) ? real(1) / real(6) : real(1) / real(3);
            const real bb = sa.rkStep == 3 ? real(1) : real(0.5);
#pragma unroll
            for (int v = 0; v < 5; v++) {
                const real qn1 = F.q[v][i];
                if (sa.rkStep == 1)
                    F.acc[v][i] = qn1 + aa * k[v];
                else
                    F.acc[v][i] = F.acc[v][i] + aa * k[v];
                if (sa.rkStep < 4) F.qk[v][i] = qn1 + bb * k[v];
            }
        } else {
#pragma unroll
            for (int v = 0; v < 5; v++) {
                const real qn1 = F.q[v][i];
                if (sa.rkStep == 1)
                    F.qk[v][i] = qn1 + k[v];
                else if (sa.rkStep == 2)
                    F.qk[v][i] = real(0.75) * qn1 + real(0.25) * qs_[v] + real(0.25) * k[v];
                else
                    F.acc[v][i] = (real(1) / real(3)) * qn1 + c23 * qs_[v] + c23 * k[v];
            }
        }
    }
}

template <typename real, bool UNI>
hipError_t launch_zpass(const OpDev<real> *d1, const OpDev<real> *d2, int threads, int LX,
                        const LineMap &lm, const Fields<real> &F, const Phys<real> &ph,
                        const StageArgs &sa, const StepScalars *sc, hipStream_t st)
{
    dim3 grid((lm.nlines + LX - 1) / LX), block(threads);
    size_t shm = 2 * (size_t)threads * sizeof(real);
    hipLaunchKernelGGL((k_zpass<real, SEG_DERIV, UNI>), grid, block, shm, st, d1, d2, lm, F, ph,
                       sa, sc);
    return hipGetLastError();
}

#define CS3D_INST_Z(real, UNI)                                                               \
    template hipError_t launch_zpass<real, UNI>(const OpDev<real> *, const OpDev<real> *, int,  \
                                                int, const LineMap &, const Fields<real> &,     \
                                                const Phys<real> &, const StageArgs &,          \
                                                const StepScalars *, hipStream_t);
CS3D_INST_Z(float, true)
CS3D_INST_Z(float, false)
CS3D_INST_Z(double, true)
CS3D_INST_Z(double, false)

}  // namespace cs3d
