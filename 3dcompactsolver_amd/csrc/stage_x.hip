// stage_x.hip -- last sweep of one RK stage: lines along x (contiguous in
// memory, MAP_ROW lanes, whole 16-byte chunks), with the pointwise
// assembly and the RK update fused in.
//
// From the stage state: the x Euler fluxes (CSolver_AWS.cpp:861-867) and
// their first derivatives (:915-922), the first and second x-derivatives of
// U, V, W, T (:877-912), and the mixed Uxy, Vxy, Uxz, Wxz as x-derivatives
// of the stored Uy, Vy, Uz, Wz (the reference takes them along y and z of
// x-derivatives, :991-1003 and :1161-1192; operators along different axes
// commute).  Then, per point, the terms of solveContinuity .. solveEnergy
// (:1246-1504) that are products of first derivatives, the sponge source
// (twice in the energy, :1491/:1495), the wall and sponge plane rules of
// postStepBCHandling (:1506-1670), and the stage update of
// RK4::updateConservedData (RK4.hpp:45-111) or TVDRK3::updateConservedData
// (TVDRK3.hpp:225-260).  Solves: a batch of 9 (5 flux + 4 mixed), then
// 4 first + 4 second derivatives.
//
// The reference's x block forms its fluxes and derivatives from the
// primitive arrays as they stand (CSolver_AWS.cpp:861-912).  At the first
// stage those were computed (updateNonConservedData, :1913-1925) before
// updateSponge reset the sponge planes of the state (:1955-2021), so on
// those planes the primitives belong to the state before the reset while the
// conserved fluxes use the state after it.  The STALE variant reproduces
// this from the copy of the pre-reset plane values that the sponge update
// keeps, and differentiates the stale-based mixed inputs of the first two
// sweeps (F.mix).
#include "stage_common.hpp"

namespace cs3d {

// QS: the stage state at the point stays in registers for the update (the
// sponge source, the TVDRK3 combinations of stages 2 and 3, and *1 itself
// at the first stage, where the stage state is *1)
// FILT: last stage of a step that filters, on uniform periodic x lines: the
// new state is filtered along x in place of the x filter sweep and written
// to *1 (fx: the x filter with this sweep's segmentation)
// CMP: compensated FP32 mode: the state, the RK4 accumulator and TVDRK3's
// stage state carry low parts (F.qlo, F.acclo, F.qklo); the update forms
// hi + lo + c dt R in FP64 and splits the result back into hi and lo
template <typename real, int M, bool UNI, bool STALE, bool QS, bool FILT, bool CMP = false>
__global__ __launch_bounds__(X_SWEEP_THREADS) void k_xpass(
    const OpDev<real> *__restrict__ d1, const OpDev<real> *__restrict__ d2,
    const OpDev<real> *__restrict__ fx, LineMap lm, Fields<real> F, Phys<real> ph, StageArgs sa,
    const StepScalars *__restrict__ sc)
{
    constexpr int HW = 2, W = M + 2 * HW;
    CS3D_LANE_SETUP(MAP_ROW, d1, FILT ? xfilt_kr(d1, d2, fx) : sweep_kr(d1, d2), SWEEP_PLANES)
    const bool use_k = sa.rkStep > 1;
    const QSet<real> qs = use_k ? F.qk : F.q;
    real a0[W], a1[W], a2[W], a3[W], a4[W];
    real ir[W], U[W], V[W], Wv[W], pp[W], Tp[W];
    // (halos taken from neighbouring lanes with ds_bpermute instead of the
    // chunk loads measured 15% slower: the permutes wait for each load)
    win.load_quad(qs.m, a0, a1, a2, a3);
    win.load(qs.e, a4);
    // a conserved set's values on this lane's segment
    auto load_set_seg = [&](const QSet<real> &q5, real(&r)[5][M]) {
        win.load_seg_quad(q5.m, r[0], r[1], r[2], r[3]);
        win.load_seg(q5.e, r[4]);
    };
    auto store_set_seg = [&](const QSet<real> &q5, const real(&r)[5][M]) {
        win.template store_quad<false>(q5.m, r[0], r[1], r[2], r[3], c.len);
        win.template store<true>(q5.e, r[4], c.len);
    };
#pragma unroll
    for (int t = 0; t < W; t++)
        prim(ph, a0[t], a1[t], a2[t], a3[t], a4[t], ir[t], U[t], V[t], Wv[t], pp[t], Tp[t]);
    // lines along x: L = j + Ny k, the point index is i
    const int lj = Lc % sa.Ny, lk = Lc / sa.Ny;
    const bool line_plane = (sa.sponge_face[2] && lj == 0) ||
                            (sa.sponge_face[3] && lj == sa.Ny - 1) ||
                            (sa.sponge_face[4] && lk == 0) ||
                            (sa.sponge_face[5] && lk == sa.Nz - 1);
    // whole x lines on a y or z wall: the normal direction of the last such
    // face (z faces come after y faces), -1 when none
    const int line_wall = ((sa.wall_face[4] && lk == 0) || (sa.wall_face[5] && lk == sa.Nz - 1))
                              ? 2
                          : ((sa.wall_face[2] && lj == 0) || (sa.wall_face[3] && lj == sa.Ny - 1))
                              ? 1
                              : -1;
    if (STALE) {
        // primitives on sponge planes from the pre-reset copy
#pragma unroll
        for (int t = 0; t < W; t++) {
            int p = c.s - HW + t;
            p += p < 0 ? n : 0;
            p -= p >= n ? n : 0;
            const bool plane = line_plane || (sa.sponge_face[0] && p == 0) ||
                               (sa.sponge_face[1] && p == n - 1);
            if (plane) {
                real qo[5];
                qset_load_at(F.qstale, win.at(t), qo);
                prim(ph, qo[0], qo[1], qo[2], qo[3], qo[4], ir[t], U[t], V[t], Wv[t], pp[t], Tp[t]);
            }
        }
    }
    const real c43 = real(4) / real(3), c13 = real(1) / real(3), c23 = real(2) / real(3);
    real mu[M], amu[M], Uc[M], Vc[M], Wc[M], qs_[5][M];
#pragma unroll
    for (int t = 0; t < M; t++) {
        visc(ph, ph.T0s + Tp[HW + t], mu[t], amu[t]);
        Uc[t] = U[HW + t];
        Vc[t] = V[HW + t];
        Wc[t] = Wv[HW + t];
        qs_[0][t] = a0[HW + t];
        qs_[1][t] = a1[HW + t];
        qs_[2][t] = a2[HW + t];
        qs_[3][t] = a3[HW + t];
        qs_[4][t] = a4[HW + t];
    }
    // the stored y and z derivatives that are differentiated along x (their
    // stale-based counterparts in the STALE variant) and their own values
    real Uy[M], Vy[M], uz[M], wz[M], Wy[M], Ty[M];
    real R0[M], Ru[M], Rv[M], Rw[M], RE[M];
    {
        // x fluxes (the h_ref U part of the energy flux is added as h_ref Ux)
        // and the mixed derivatives
        real r[9][M], f[W];
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
        if constexpr (STALE) {
            win.load(F.mix[0], f);
            rhs<OP_D1, real, M, HW, UNI>(d1, c, f, r[5]);
            win.load(F.mix[1], f);
            rhs<OP_D1, real, M, HW, UNI>(d1, c, f, r[6]);
            win.load(F.mix[2], f);
            rhs<OP_D1, real, M, HW, UNI>(d1, c, f, r[7]);
            win.load(F.mix[3], f);
            rhs<OP_D1, real, M, HW, UNI>(d1, c, f, r[8]);
        } else {
            // Uy with Ty, Vy with Wy, Uz with Wz
            real fb[W];
            win.load_pair(F.dYut, f, fb);
#pragma unroll
            for (int t = 0; t < M; t++) {
                Uy[t] = f[HW + t];
                Ty[t] = fb[HW + t];
            }
            rhs<OP_D1, real, M, HW, UNI>(d1, c, f, r[5]);
            win.load_pair(F.dYvw, f, fb);
#pragma unroll
            for (int t = 0; t < M; t++) {
                Vy[t] = f[HW + t];
                Wy[t] = fb[HW + t];
            }
            rhs<OP_D1, real, M, HW, UNI>(d1, c, f, r[6]);
            win.load_pair(F.dZuw, f, fb);
#pragma unroll
            for (int t = 0; t < M; t++) {
                uz[t] = f[HW + t];
                wz[t] = fb[HW + t];
            }
            rhs<OP_D1, real, M, HW, UNI>(d1, c, f, r[7]);
            rhs<OP_D1, real, M, HW, UNI>(d1, c, fb, r[8]);
        }
        solve_batch<9, kc_deriv<real>(), real, M, UNI>(d1, c, r);
#pragma unroll
        for (int t = 0; t < M; t++) {
            RE[t] = -r[0][t];
            Ru[t] = -r[1][t];
            R0[t] = -r[2][t];
            Rv[t] = -r[3][t];
            Rw[t] = -r[4][t];
            const real Uxy = r[5][t], Vxy = r[6][t], Uxz = r[7][t], Wxz = r[8][t];
            Rv[t] += mu[t] * (c13 * Uxy);
            RE[t] += mu[t] * (c13 * (Vc[t] * Uxy));
            Ru[t] += mu[t] * (c13 * Vxy);
            RE[t] += mu[t] * (c13 * (Uc[t] * Vxy));
            Rw[t] += mu[t] * (c13 * Uxz);
            RE[t] += mu[t] * (c13 * (Wc[t] * Uxz));
            Ru[t] += mu[t] * (c13 * Wxz);
            RE[t] += mu[t] * (c13 * (Uc[t] * Wxz));
        }
    }
    if (STALE) {
        // the assembly takes the current first derivatives
        win.load_seg_pair(F.dYut, Uy, Ty);
        win.load_seg_pair(F.dYvw, Vy, Wy);
        win.load_seg_pair(F.dZuw, uz, wz);
    }
    // the rest of the point values, whole segments at a time, issued before
    // the second solve batch so that they arrive during it (inactive lanes
    // address the last line and read in bounds)
    real vz[M], Tz[M], Rp[5][M];
    win.load_seg_pair(F.dZvt, vz, Tz);
    {
        real rz[5][M];
        win.load_seg_quad(F.Ry4, Rp[0], Rp[1], Rp[2], Rp[3]);
        win.load_seg(F.RyE, Rp[4]);
        win.load_seg_quad(F.Rz4, rz[0], rz[1], rz[2], rz[3]);
        win.load_seg(F.RzE, rz[4]);
#pragma unroll
        for (int v = 0; v < 5; v++)
#pragma unroll
            for (int t = 0; t < M; t++) Rp[v][t] += rz[v][t];
    }
    // first and second x-derivatives of T', U, V, W
    real r[8][M];
    rhs<OP_D1, real, M, HW, UNI>(d1, c, Tp, r[0]);
    rhs<OP_D1, real, M, HW, UNI>(d1, c, U, r[1]);
    rhs<OP_D1, real, M, HW, UNI>(d1, c, V, r[2]);
    rhs<OP_D1, real, M, HW, UNI>(d1, c, Wv, r[3]);
    rhs<OP_D2, real, M, HW, UNI>(d2, c, Tp, r[4]);
    rhs<OP_D2, real, M, HW, UNI>(d2, c, U, r[5]);
    rhs<OP_D2, real, M, HW, UNI>(d2, c, V, r[6]);
    rhs<OP_D2, real, M, HW, UNI>(d2, c, Wv, r[7]);
    solve_d12<4, 4, real, M, UNI>(d1, d2, c, r);
#pragma unroll
    for (int t = 0; t < M; t++) {
        RE[t] += ph.cpPr * (mu[t] * r[4][t]);
        RE[t] -= ph.h_ref * r[1][t];
        Ru[t] += mu[t] * (c43 * r[5][t]);
        RE[t] += mu[t] * (Uc[t] * (c43 * r[5][t]));
        Rv[t] += mu[t] * r[6][t];
        RE[t] += mu[t] * (Vc[t] * r[6][t]);
        Rw[t] += mu[t] * r[7][t];
        RE[t] += mu[t] * (Wc[t] * r[7][t]);
    }
    // inactive lanes (a partial last workgroup) stop here, except where the
    // fused filter still needs every lane at its barriers; they then skip
    // every store
    if constexpr (!FILT) {
        if (!active) return;
    }
    // the update's inputs (the step's base state, the RK4 accumulator)
    // issued before the pointwise assembly, so that they arrive during it
    // (FP32 3.5% faster steps).  The fused-filter sweep, which runs at
    // RK4's last stage, takes only the accumulator.  The STALE sweep runs
    // only at a step's first stage (the launcher checks it), where the base
    // state is the stage state already in registers and RK4 has no
    // accumulator yet: it has nothing to issue early.  FP64, at one
    // wavefront per SIMD with its registers full, issues only the
    // accumulator early, and only where a stage needs it (not in the QS
    // variants: RK4's first stage and TVDRK3, which keeps none); issuing
    // the base state early as well overflows into scratch.
    constexpr bool HOIST = !STALE && (sizeof(real) == 4 || !QS);
    constexpr bool HOIST_BASE = HOIST && !FILT && sizeof(real) == 4;
    real qb[HOIST_BASE ? 5 : 1][M], ab[HOIST ? 5 : 1][M];
    if constexpr (HOIST) {
        if (sa.mode != 1) {
            const bool first_stage = sa.rkStep == 1;
            if constexpr (HOIST_BASE) {
                if (QS && first_stage) {
#pragma unroll
                    for (int v = 0; v < 5; v++)
#pragma unroll
                        for (int t = 0; t < M; t++) qb[v][t] = qs_[v][t];
                } else {
                    load_set_seg(F.q, qb);
                }
            }
            if (sa.rk == 4 && !first_stage) load_set_seg(F.acc, ab);
        }
    }

    const real dt = (real)sc->dt;
    real k[5][M];
#pragma unroll
    for (int t = 0; t < M; t++) {
        const real Tx = r[0][t], Ux = r[1][t], Vx = r[2][t], Wx = r[3][t];
        const real MuX = amu[t] * Tx, MuY = amu[t] * Ty[t], MuZ = amu[t] * Tz[t];
        const real Uu = Uc[t], Vv = Vc[t], Ww = Wc[t], m_ = mu[t];
        const real uy = Uy[t], vy = Vy[t], wy = Wy[t], uz_ = uz[t], vz_ = vz[t], wz_ = wz[t];
        real su = c43 * MuX * (Ux - real(0.5) * vy - real(0.5) * wz_) + MuY * (uy + Vx) +
                  MuZ * (Wx + uz_);
        real sv = c43 * MuY * (vy - real(0.5) * Ux - real(0.5) * wz_) + MuX * (uy + Vx) +
                  MuZ * (wy + vz_);
        real sw = c43 * MuZ * (wz_ - real(0.5) * Ux - real(0.5) * vy) + MuX * (Wx + uz_) +
                  MuY * (wy + vz_);
        real qn = ph.cpPr * (MuX * Tx + MuY * Ty[t] + MuZ * Tz[t]);
        real quad = c43 * (Ux * Ux + vy * vy + wz_ * wz_) + uy * uy + uz_ * uz_ + Vx * Vx +
                    vz_ * vz_ + Wx * Wx + wy * wy - c43 * (Ux * vy + Ux * wz_ + vy * wz_) +
                    real(2) * (uy * Vx + uz_ * Wx + vz_ * wy);
        real vt2 = c43 * (Uu * MuX * Ux + Vv * MuY * vy + Ww * MuZ * wz_) -
                   c23 * Uu * MuX * (vy + wz_) - c23 * Vv * MuY * (Ux + wz_) -
                   c23 * Ww * MuZ * (Ux + vy) + Uu * MuY * (uy + Vx) + Uu * MuZ * (uz_ + Wx) +
                   Vv * MuX * (uy + Vx) + Vv * MuZ * (vz_ + wy) + Ww * MuX * (uz_ + Wx) +
                   Ww * MuY * (vz_ + wy);
        real rhs_[5];
        rhs_[0] = Rp[0][t] + R0[t];
        rhs_[1] = Rp[1][t] + Ru[t] + su;
        rhs_[2] = Rp[2][t] + Rv[t] + sv;
        rhs_[3] = Rp[3][t] + Rw[t] + sw;
        rhs_[4] = Rp[4][t] + RE[t] + qn + m_ * quad + vt2;
        // table plans: points past the segment's length are never stored
        // and their offsets may leave the field, so nothing is read there
        const bool valid = UNI || t < c.len;
        if (QS && F.sigma && valid) {
            const unsigned i = win.mid(t);
            const real sg = ldo(F.sigma, i);
            real av[5];
            qset_load_at(F.avg, i, av);
#pragma unroll
            for (int v = 0; v < 5; v++) rhs_[v] += sg * (av[v] - qs_[v][t]);
            rhs_[4] += sg * (av[4] - qs_[4][t]);
        }
        const int ii = c.s + t;
        const bool plane = line_plane || (sa.sponge_face[0] && ii == 0) ||
                           (sa.sponge_face[1] && ii == sa.Nx - 1);
#pragma unroll
        for (int v = 0; v < 5; v++) k[v][t] = plane ? real(0) : dt * rhs_[v];
        // wall planes keep only the normal continuity flux term of the last
        // wall face through the point (faces X0..Z1 in turn); a sponge plane
        // stays zero
        int wn = line_wall;
        if (wn < 0 && ((sa.wall_face[0] && ii == 0) || (sa.wall_face[1] && ii == sa.Nx - 1)))
            wn = 0;
        if (wn >= 0 && !plane) {
            real cn = R0[t];
            if (wn > 0) cn = valid ? ldo_part<4>(wn == 1 ? F.Ry4 : F.Rz4, win.mid(t), 0) : real(0);
            k[0][t] = dt * cn;
#pragma unroll
            for (int v = 1; v < 5; v++) k[v][t] = real(0);
        }
        if (sa.diag && valid && active) {
            // per-point inputs of the step diagnostics that come from this
            // stage's derivatives (CSolver_AWS.cpp:2190-2225)
            const unsigned i = win.mid(t);
            const real dil = Ux + vy + wz_;
            const real S00 = real(0.5) * (Ux + Ux) - c13 * dil;
            const real S11 = real(0.5) * (vy + vy) - c13 * dil;
            const real S22 = real(0.5) * (wz_ + wz_) - c13 * dil;
            const real S01 = real(0.5) * (uy + Vx), S02 = real(0.5) * (uz_ + Wx),
                       S12 = real(0.5) * (vz_ + wy);
            real td = S00 * Ux;
            td += S01 * uy;
            td += S02 * uz_;
            td += S01 * Vx;
            td += S11 * vy;
            td += S12 * vz_;
            td += S02 * Wx;
            td += S12 * wy;
            td += S22 * wz_;
            const real wx_ = wy - vz_, wy_ = uz_ - Wx, wzv = Vx - uy;
            sto(F.diag[0], i, td);
            sto(F.diag[1], i, wx_ * wx_ + wy_ * wy_ + wzv * wzv);
            sto(F.diag[2], i, dil);
            sto(F.diag[3], i, (Ux * Ux + vy * vy + wz_ * wz_) * c13);
        }
    }
    if (sa.mode == 1) {
#pragma unroll
        for (int v = 0; v < 5; v++) win.template store<true>(F.k2[v], k[v], c.len);
        return;
    }
    // stage-update coefficients, uniform over the launch
    const bool first = STALE || sa.rkStep == 1, has_next = sa.rkStep < 4;
    const real ra = (sa.rkStep == 1 || sa.rkStep == 4) ? real(1) / real(6) : real(1) / real(3);
    const real rb = sa.rkStep == 3 ? real(1) : real(0.5);
    const bool tvd_to_k = sa.rkStep < 3;
    const real ta = sa.rkStep == 1 ? real(1) : sa.rkStep == 2 ? real(0.75) : real(1) / real(3);
    const real tb = sa.rkStep == 1 ? real(0) : sa.rkStep == 2 ? real(0.25) : real(2) / real(3);
    const real tc = sa.rkStep == 1 ? real(1) : sa.rkStep == 2 ? real(0.25) : real(2) / real(3);
    // RK4's last stage takes only the accumulator, not the step's base
    // state: the fused-filter sweep, which runs only there, skips that load
    // (TVDRK3's last stage still takes it, here)
    const bool need_base = !FILT || sa.rk != 4 || has_next || first;
    real fin[FILT ? 5 : 1][M];  // the step's new state (FILT)
    // the step's base state (*1) and RK4's accumulator where they are not in
    // registers already (the hoisted loads above, or the stage state at the
    // first stage)
    real ql[HOIST_BASE ? 1 : 5][M], al[HOIST ? 1 : 5][M];
    if constexpr (!HOIST_BASE) {
        if (!(QS && first) && need_base) load_set_seg(F.q, ql);
    }
    if constexpr (!HOIST) {
        if (sa.rk == 4 && !first) load_set_seg(F.acc, al);
    }
    auto base = [&](int v, int t) -> real {
        if constexpr (HOIST_BASE) {
            return qb[v][t];
        } else {
            if (QS && first) return qs_[v][t];
            return need_base ? ql[v][t] : real(0);
        }
    };
    auto accum = [&](int v, int t) -> real {
        if constexpr (HOIST)
            return ab[v][t];
        else
            return al[v][t];
    };
    // compensated mode: the low parts of the base state, the accumulator and
    // TVDRK3's stage state, and the update coefficients in FP64
    static_assert(!CMP || sizeof(real) == 4, "the compensated mode is an FP32 mode");
    real blo[CMP ? 5 : 1][M], alo[CMP ? 5 : 1][M], slo[CMP ? 5 : 1][M];
    real finlo[CMP && FILT ? 5 : 1][M], eAL[CMP ? M : 1];
    if constexpr (CMP) {
#pragma unroll
        for (int v = 0; v < 5; v++)
#pragma unroll
            for (int t = 0; t < M; t++) blo[v][t] = alo[v][t] = slo[v][t] = real(0);
        if (need_base) load_set_seg(F.qlo, blo);
        if (sa.rk == 4 && !first) load_set_seg(F.acclo, alo);
        if (sa.rk == 3 && sa.rkStep > 1) load_set_seg(F.qklo, slo);
    }
    const double rad = (sa.rkStep == 1 || sa.rkStep == 4) ? 1.0 / 6.0 : 1.0 / 3.0;
    const double rbd = sa.rkStep == 3 ? 1.0 : 0.5;
    const double tad = sa.rkStep == 1 ? 1.0 : sa.rkStep == 2 ? 0.75 : 1.0 / 3.0;
    const double tbd = sa.rkStep == 1 ? 0.0 : sa.rkStep == 2 ? 0.25 : 2.0 / 3.0;
    const double tcd = sa.rkStep == 1 ? 1.0 : sa.rkStep == 2 ? 0.25 : 2.0 / 3.0;
    const QSet<real> out_lo = (sa.rk == 4 || !tvd_to_k) ? F.acclo : F.qklo;
    // point by point: each point's quad goes out as soon as it is formed;
    // the rhoE' values of the segment go out together at the end.  The quad
    // stores are ordinary ones: a wavefront's store of one point per lane
    // covers every fourth 16 bytes, which the cache merges with the other
    // three points' stores; as non-temporal stores the FP32 x sweep took
    // 37% longer (FP64 44%)
    real eA[M], eQ[M];
    const QSet<real> out_a = (sa.rk == 4 || !tvd_to_k) ? F.acc : F.qkw;
    const bool store_q = sa.rk == 4 && has_next;
#pragma unroll
    for (int t = 0; t < M; t++) {
        const bool st_ok = UNI || t < c.len;
        real b[5], a[5];
#pragma unroll
        for (int v = 0; v < 5; v++) b[v] = base(v, t);
        if constexpr (CMP) {
            real al_[5];
            if (sa.rk == 4) {
#pragma unroll
                for (int v = 0; v < 5; v++) {
                    const double prev = first ? two_part(b[v], blo[v][t])
                                              : two_part(accum(v, t), alo[v][t]);
                    split_part(prev + rad * (double)k[v][t], a[v], al_[v]);
                }
                if (store_q) {
#pragma unroll
                    for (int v = 0; v < 5; v++)
                        b[v] = (real)(two_part(b[v], blo[v][t]) + rbd * (double)k[v][t]);
                    if (st_ok) win.template store_quad_at<false>(F.qkw.m, t, b[0], b[1], b[2], b[3]);
                    eQ[t] = b[4];
                }
            } else {
#pragma unroll
                for (int v = 0; v < 5; v++)
                    split_part(tad * two_part(b[v], blo[v][t]) +
                                   tbd * two_part(qs_[v][t], slo[v][t]) + tcd * (double)k[v][t],
                               a[v], al_[v]);
            }
            if constexpr (FILT) {
#pragma unroll
                for (int v = 0; v < 5; v++) finlo[v][t] = al_[v];
            } else {
                if (st_ok) win.template store_quad_at<false>(out_lo.m, t, al_[0], al_[1], al_[2], al_[3]);
                eAL[t] = al_[4];
            }
        } else if (sa.rk == 4) {
            // RK4.hpp:45-111: *2 accumulates the stages, *k is the next
            // stage state
#pragma unroll
            for (int v = 0; v < 5; v++) a[v] = (first ? b[v] : accum(v, t)) + ra * k[v][t];
            if (store_q) {
#pragma unroll
                for (int v = 0; v < 5; v++) b[v] += rb * k[v][t];
                if (st_ok) win.template store_quad_at<false>(F.qkw.m, t, b[0], b[1], b[2], b[3]);
                eQ[t] = b[4];
            }
        } else {
            // TVDRK3.hpp:225-260: stages 1-2 write *k, stage 3 writes *2
#pragma unroll
            for (int v = 0; v < 5; v++) a[v] = ta * b[v] + tb * qs_[v][t] + tc * k[v][t];
        }
        if constexpr (FILT) {
#pragma unroll
            for (int v = 0; v < 5; v++) fin[v][t] = a[v];
        } else {
            if (st_ok) win.template store_quad_at<false>(out_a.m, t, a[0], a[1], a[2], a[3]);
            eA[t] = a[4];
        }
    }
    if constexpr (!FILT) win.template store<true>(out_a.e, eA, c.len);
    if constexpr (CMP && !FILT) win.template store<true>(out_lo.e, eAL, c.len);
    if (store_q) win.template store<true>(F.qkw.e, eQ, c.len);
    if constexpr (FILT) {
        // filterConservedData's x sweep (CSolver_AWS.cpp:1682-1893) on the
        // new state in registers: out = f - A^-1 (A - B) f.  The halo (four
        // points on each side: the whole neighbouring segments) comes
        // through the LDS ring, free past the solves; lines are periodic
        // and uniform, so the neighbours are segments g -+ 1 mod P.
        static_assert(M == 4, "the fused filter takes whole neighbouring segments as halo");
        constexpr int FH = 4, FW = M + 2 * FH;
        real *stg = c.xb;
        const int T = blockDim.x, me = threadIdx.x;
        __syncthreads();
#pragma unroll
        for (int v = 0; v < 5; v++)
#pragma unroll
            for (int t = 0; t < M; t++) stg[(v * M + t) * T + me] = fin[v][t];
        __syncthreads();
        const int ml = c.line * P + (c.g == 0 ? P - 1 : c.g - 1);
        const int mr = c.line * P + (c.g == P - 1 ? 0 : c.g + 1);
        real r[5][M];
#pragma unroll
        for (int v = 0; v < 5; v++) {
            real w[FW];
#pragma unroll
            for (int t = 0; t < M; t++) {
                w[t] = stg[(v * M + t) * T + ml];
                w[FH + t] = fin[v][t];
                w[FH + M + t] = stg[(v * M + t) * T + mr];
            }
            rhs<OP_FILTER, real, M, FH, true>(fx, c, w, r[v]);
        }
        if (c.P == 64) {
            solve_batch_wave<5, 0, real, M, MAP_ROW>(fx, fx, c, r);
        } else {
            // the exchange ring starts at the staging area: every lane of
            // the workgroup must have read its halos before any lane writes
            // the ring (c.P is uniform, so the barrier is too)
            __syncthreads();
            solve_batch<5, kc_filter_x<real>(), real, M, true>(fx, c, r);
        }
        if (active) {
            if constexpr (CMP) {
                // the correction from the high parts, subtracted from hi + lo
#pragma unroll
                for (int v = 0; v < 5; v++)
#pragma unroll
                    for (int t = 0; t < M; t++)
                        split_part(two_part(fin[v][t], finlo[v][t]) - (double)r[v][t], fin[v][t],
                                   finlo[v][t]);
                store_set_seg(F.qlo, finlo);
            } else {
#pragma unroll
                for (int v = 0; v < 5; v++)
#pragma unroll
                    for (int t = 0; t < M; t++) fin[v][t] -= r[v][t];
            }
            store_set_seg(F.q, fin);
        }
    }
}

template <typename real, bool UNI, bool STALE, bool FILT, bool CMP>
static hipError_t launch_x5(const OpDev<real> *d1, const OpDev<real> *d2, const OpDev<real> *fx,
                            const LaunchGeom &lg, const LineMap &lm, const Fields<real> &F,
                            const Phys<real> &ph, const StageArgs &sa, const StepScalars *sc,
                            hipStream_t st)
{
    dim3 grid((lm.nlines + lg.LX - 1) / lg.LX), block(lg.threads);
    if (STALE || F.sigma || sa.rk == 3 || sa.rkStep == 1)
        hipLaunchKernelGGL((k_xpass<real, SEG_DERIV, UNI, STALE, true, FILT, CMP>), grid, block,
                           lg.shm, st, d1, d2, fx, lm, F, ph, sa, sc);
    else
        hipLaunchKernelGGL((k_xpass<real, SEG_DERIV, UNI, STALE, false, FILT, CMP>), grid, block,
                           lg.shm, st, d1, d2, fx, lm, F, ph, sa, sc);
    return hipGetLastError();
}

template <typename real, bool UNI, bool STALE, bool FILT>
static hipError_t launch_x4(const OpDev<real> *d1, const OpDev<real> *d2, const OpDev<real> *fx,
                            const LaunchGeom &lg, const LineMap &lm, const Fields<real> &F,
                            const Phys<real> &ph, const StageArgs &sa, const StepScalars *sc,
                            hipStream_t st)
{
    // the stale primitives exist only at the first stage after a sponge
    // update, and the STALE kernel takes the first stage as given
    if (STALE && sa.rkStep != 1) return hipErrorInvalidValue;
    // the compensated variants: FP32, whole updates only
    if (sa.compensated && sa.mode == 0) {
        if constexpr (sizeof(real) == 4)
            return launch_x5<real, UNI, STALE, FILT, true>(d1, d2, fx, lg, lm, F, ph, sa, sc, st);
        return hipErrorInvalidValue;
    }
    return launch_x5<real, UNI, STALE, FILT, false>(d1, d2, fx, lg, lm, F, ph, sa, sc, st);
}

template <typename real, bool UNI>
hipError_t launch_xpass(const OpDev<real> *d1, const OpDev<real> *d2, const OpDev<real> *fx,
                        const LaunchGeom &lg, const LineMap &lm, const Fields<real> &F,
                        const Phys<real> &ph, const StageArgs &sa, const StepScalars *sc,
                        hipStream_t st)
{
    if (lg.threads > X_SWEEP_THREADS) return hipErrorInvalidConfiguration;
    if (fx) {
        // the fused filter: uniform periodic lines only, never a stale
        // stage (the last stage of a step), never the right-hand side alone
        if (!UNI || sa.stale || sa.mode != 0) return hipErrorInvalidValue;
        if constexpr (UNI)
            return launch_x4<real, true, false, true>(d1, d2, fx, lg, lm, F, ph, sa, sc, st);
    }
    return sa.stale ? launch_x4<real, UNI, true, false>(d1, d2, nullptr, lg, lm, F, ph, sa, sc, st)
                    : launch_x4<real, UNI, false, false>(d1, d2, nullptr, lg, lm, F, ph, sa, sc, st);
}

#define CS3D_INST_X(real, UNI)                                                                  \
    template hipError_t launch_xpass<real, UNI>(const OpDev<real> *, const OpDev<real> *,          \
                                                const OpDev<real> *, const LaunchGeom &,           \
                                                const LineMap &,                                   \
                                                const Fields<real> &, const Phys<real> &,          \
                                                const StageArgs &, const StepScalars *,            \
                                                hipStream_t);
CS3D_INST_X(float, true)
CS3D_INST_X(float, false)
CS3D_INST_X(double, true)
CS3D_INST_X(double, false)

}  // namespace cs3d
