// stage_x.hip -- X sweep of one RK stage (lines along x, contiguous in
// memory, MAP_ROW lanes).
//
// From the stage state: the x Euler fluxes (CSolver_AWS.cpp:861-867) and
// their first derivatives (:915-922), the first and second x-derivatives of
// U, V, W, T (:877-912); stores Ux Vx Wx Tx and starts the right-hand side
// with every term that is linear in an x-derivative.
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

#define XU(t) (b1[t] * ir[t])
#define XV(t) (b2[t] * ir[t])
#define XW(t) (b3[t] * ir[t])

template <typename real, int M, bool UNI, bool STALE>
__global__ __launch_bounds__(sweep_threads<real>()) void k_xpass(
    const OpDev<real> *__restrict__ d1, const OpDev<real> *__restrict__ d2, LineMap lm,
    Fields<real> F, Phys<real> ph, int use_k, StageArgs sa)
{
    CS3D_LANE_SETUP(MAP_ROW, d1)
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
    // primitive sources: the state itself, or the pre-reset copy on sponge
    // planes (STALE)
    real b0[W], b1[W], b2[W], b3[W], b4[W];
#pragma unroll
    for (int t = 0; t < W; t++) {
        b0[t] = a0[t];
        b1[t] = a1[t];
        b2[t] = a2[t];
        b3[t] = a3[t];
        b4[t] = a4[t];
    }
    if (STALE) {
        const int Lc = active ? L : lm.nlines - 1;
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
            if (plane) {
                const long i = base + p * sp;
                b0[t] = F.qstale[0][i];
                b1[t] = F.qstale[1][i];
                b2[t] = F.qstale[2][i];
                b3[t] = F.qstale[3][i];
                b4[t] = F.qstale[4][i];
            }
        }
    }
#pragma unroll
    for (int t = 0; t < W; t++) ir[t] = real(1) / (ph.rho_ref + b0[t]);
    real pp[W], w[W], r[M];
#pragma unroll
    for (int t = 0; t < W; t++) pp[t] = pprime(ph, b1, b2, b3, b4, ir, t);
    real R0[M], Ru[M], Rv[M], Rw[M], RE[M];
    // energy flux (rhoE + p)U = h_ref U + (rhoE' + p')U; the h_ref part is
    // added below as h_ref * Ux
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = (a4[t] + pp[t]) * XU(t);
    compact_op<OP_D1, real, M, HW, UNI, MAP_ROW>(d1, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) RE[t] = -r[t];
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = a1[t] * XU(t) + pp[t];
    compact_op<OP_D1, real, M, HW, UNI, MAP_ROW>(d1, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) Ru[t] = -r[t];
    compact_op<OP_D1, real, M, HW, UNI, MAP_ROW>(d1, c, a1, r);
#pragma unroll
    for (int t = 0; t < M; t++) R0[t] = -r[t];
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = a2[t] * XU(t);
    compact_op<OP_D1, real, M, HW, UNI, MAP_ROW>(d1, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) Rv[t] = -r[t];
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = a3[t] * XU(t);
    compact_op<OP_D1, real, M, HW, UNI, MAP_ROW>(d1, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) Rw[t] = -r[t];

    auto store = [&](real *dst, const real(&v)[M]) {
        if (active) {
#pragma unroll
            for (int t = 0; t < M; t++)
                if (t < c.len) dst[base + (long)(c.s + t) * sp] = v[t];
        }
    };
    // temperature perturbation T' and the viscosity at the segment's points
    real mu[M];
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = (pp[t] - b0[t] * ph.RT0s) * ir[t] * ph.iR;
#pragma unroll
    for (int t = 0; t < M; t++) {
        real amu;
        visc(ph, ph.T0s + w[HW + t], mu[t], amu);
    }
    compact_op<OP_D1, real, M, HW, UNI, MAP_ROW>(d1, c, w, r);
    store(F.dX[3], r);
    compact_op<OP_D2, real, M, HW, UNI, MAP_ROW>(d2, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) RE[t] += ph.cpPr * (mu[t] * r[t]);
    const real c43 = real(4) / real(3);
    // U: Ux, Uxx
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = XU(t);
    compact_op<OP_D1, real, M, HW, UNI, MAP_ROW>(d1, c, w, r);
    store(F.dX[0], r);
#pragma unroll
    for (int t = 0; t < M; t++) RE[t] -= ph.h_ref * r[t];
    compact_op<OP_D2, real, M, HW, UNI, MAP_ROW>(d2, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) {
        Ru[t] += mu[t] * (c43 * r[t]);
        RE[t] += mu[t] * (w[HW + t] * (c43 * r[t]));
    }
    // V: Vx, Vxx
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = XV(t);
    compact_op<OP_D1, real, M, HW, UNI, MAP_ROW>(d1, c, w, r);
    store(F.dX[1], r);
    compact_op<OP_D2, real, M, HW, UNI, MAP_ROW>(d2, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) {
        Rv[t] += mu[t] * r[t];
        RE[t] += mu[t] * (w[HW + t] * r[t]);
    }
    // W: Wx, Wxx
#pragma unroll
    for (int t = 0; t < W; t++) w[t] = XW(t);
    compact_op<OP_D1, real, M, HW, UNI, MAP_ROW>(d1, c, w, r);
    store(F.dX[2], r);
    compact_op<OP_D2, real, M, HW, UNI, MAP_ROW>(d2, c, w, r);
#pragma unroll
    for (int t = 0; t < M; t++) {
        Rw[t] += mu[t] * r[t];
        RE[t] += mu[t] * (w[HW + t] * r[t]);
    }
    store(F.R[0], R0);
    store(F.R[1], Ru);
    store(F.R[2], Rv);
    store(F.R[3], Rw);
    store(F.R[4], RE);
}

template <typename real, bool UNI>
hipError_t launch_xpass(const OpDev<real> *d1, const OpDev<real> *d2, int threads, int LX,
                        const LineMap &lm, const Fields<real> &F, const Phys<real> &ph,
                        int use_k, const StageArgs &sa, hipStream_t st)
{
    dim3 grid((lm.nlines + LX - 1) / LX), block(threads);
    size_t shm = 2 * (size_t)threads * sizeof(real);
    if (sa.stale)
        hipLaunchKernelGGL((k_xpass<real, SEG_DERIV, UNI, true>), grid, block, shm, st, d1, d2, lm,
                           F, ph, use_k, sa);
    else
        hipLaunchKernelGGL((k_xpass<real, SEG_DERIV, UNI, false>), grid, block, shm, st, d1, d2,
                           lm, F, ph, use_k, sa);
    return hipGetLastError();
}

#define CS3D_INST_X(real, UNI)                                                              \
    template hipError_t launch_xpass<real, UNI>(const OpDev<real> *, const OpDev<real> *, int, \
                                                int, const LineMap &, const Fields<real> &,    \
                                                const Phys<real> &, int, const StageArgs &,    \
                                                hipStream_t);
CS3D_INST_X(float, true)
CS3D_INST_X(float, false)
CS3D_INST_X(double, true)
CS3D_INST_X(double, false)

}  // namespace cs3d
