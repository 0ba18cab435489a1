// adjoint.hip -- kernels of the discrete adjoint of the stage right-hand
// side (no reference counterpart: the reference has no adjoint): the fused
// five line passes of periodic grids without sponges or walls, and the
// pointwise kernels of the general path (non-periodic lines, sponges,
// walls), which applies its derivatives with the standalone and transposed
// operators.
//
// The forward right-hand side (CSolver_AWS.cpp:1246-1504, away from sponge
// and wall planes) is, with u = m/rho, p = (gamma-1)(E - |m|^2/(2 rho)),
// T = p/(rho R), mu = mu_ref T^0.76, a = dmu/dT = 0.76 mu/T:
//   R_rho = -sum_j D_j(m_j)
//   R_i   = -sum_j D_j(m_i u_j + p delta_ij) + Rv_i
//   R_E   = -sum_j D_j((E + p) u_j) + sum_i u_i Rv_i + mu QUAD
//           + cp/Pr (a |grad T|^2 + mu lap T)
// where Rv_i = mu (4/3 d_ii u_i + sum_{j!=i} d_jj u_i + 1/3 mixed terms)
//            + a (terms linear in grad T times velocity gradients)
// and QUAD the quadratic form of the velocity gradients.  Every
// derivative is a compact operator applied to a field; the adjoint applies
// the transposed operators (on periodic lines D1 is skew-adjoint and D2
// self-adjoint, which the fused passes use; the general path applies the
// transposed operators themselves) to the cotangent seeds these kernels
// form, then chains back through the pointwise maps.  The cotangent v of the stage increment dt*R enters
// already multiplied by dt.
#include "adjoint.hpp"

namespace cs3d {

// One point of the cotangent seeds of the derivative fields, and of the
// pointwise cotangents of U, V, W, T from the viscous and heat-flux terms.
// First and second derivatives: index 3 * f + j with f = U, V, W, T and
// j = x, y, z; mixed: Uxy, Uxz, Vxy, Vyz, Wxz, Wyz.
template <typename real>
struct SeedPt {
    real U, V, W, Tp, mu, a, v1, v2, v3, vE;
    real d1[12], d2[12], mx[6];
};
enum { SD_S = 0, SD_GQ = 3, SD_G1 = 4, SD_UB = SD_G1 + G1_COUNT, SD_VB, SD_WB, SD_TB, SD_COUNT };

template <typename real>
__device__ __forceinline__ void seed_point(const Phys<real> &ph, const SeedPt<real> &p,
                                           real (&o)[SD_COUNT])
{
    const real c43 = real(4) / real(3), c23 = real(2) / real(3), c13 = real(1) / real(3);
    const real U = p.U, V = p.V, W = p.W;
    const real T = ph.T0s + p.Tp, mu = p.mu, a = p.a;
    const real v1 = p.v1, v2 = p.v2, v3 = p.v3, vE = p.vE;
    const real w1 = v1 + U * vE, w2 = v2 + V * vE, w3 = v3 + W * vE;
    const real Ux = p.d1[0], Uy = p.d1[1], Uz = p.d1[2];
    const real Vx = p.d1[3], Vy = p.d1[4], Vz = p.d1[5];
    const real Wx = p.d1[6], Wy = p.d1[7], Wz = p.d1[8];
    const real Tx = p.d1[9], Ty = p.d1[10], Tz = p.d1[11];
    const real Uxx = p.d2[0], Uyy = p.d2[1], Uzz = p.d2[2];
    const real Vxx = p.d2[3], Vyy = p.d2[4], Vzz = p.d2[5];
    const real Wxx = p.d2[6], Wyy = p.d2[7], Wzz = p.d2[8];
    const real Txx = p.d2[9], Tyy = p.d2[10], Tzz = p.d2[11];
    const real Uxy = p.mx[0], Uxz = p.mx[1], Vxy = p.mx[2];
    const real Vyz = p.mx[3], Wxz = p.mx[4], Wyz = p.mx[5];

    // second and mixed derivatives: multiples of s_i = mu w_i and of
    // gq = cp/Pr mu vE
    o[SD_S + 0] = mu * w1;
    o[SD_S + 1] = mu * w2;
    o[SD_S + 2] = mu * w3;
    o[SD_GQ] = ph.cpPr * mu * vE;

    // first derivatives: the a-terms of the momenta (weights w), the
    // quadratic form (weight mu vE) and the heat-flux square (cp/Pr a vE)
    const real P1 = c43 * (Ux - real(0.5) * Vy - real(0.5) * Wz);
    const real P2 = c43 * (Vy - real(0.5) * Ux - real(0.5) * Wz);
    const real P3 = c43 * (Wz - real(0.5) * Ux - real(0.5) * Vy);
    const real Sxy = Uy + Vx, Sxz = Wx + Uz, Syz = Wy + Vz;
    const real mq = mu * vE;
    o[SD_G1 + G1_UX] = a * (c43 * Tx * w1 - c23 * Ty * w2 - c23 * Tz * w3) +
                     mq * (real(8) / real(3) * Ux - c43 * (Vy + Wz));
    o[SD_G1 + G1_VY] = a * (-c23 * Tx * w1 + c43 * Ty * w2 - c23 * Tz * w3) +
                     mq * (real(8) / real(3) * Vy - c43 * (Ux + Wz));
    o[SD_G1 + G1_WZ] = a * (-c23 * Tx * w1 - c23 * Ty * w2 + c43 * Tz * w3) +
                     mq * (real(8) / real(3) * Wz - c43 * (Ux + Vy));
    o[SD_G1 + G1_XY] = a * (Ty * w1 + Tx * w2) + real(2) * mq * Sxy;  // Uy and Vx
    o[SD_G1 + G1_XZ] = a * (Tz * w1 + Tx * w3) + real(2) * mq * Sxz;  // Uz and Wx
    o[SD_G1 + G1_YZ] = a * (Tz * w2 + Ty * w3) + real(2) * mq * Syz;  // Vz and Wy
    const real ht = real(2) * ph.cpPr * a * vE;
    o[SD_G1 + G1_TX] = a * (P1 * w1 + Sxy * w2 + Sxz * w3) + ht * Tx;
    o[SD_G1 + G1_TY] = a * (Sxy * w1 + P2 * w2 + Syz * w3) + ht * Ty;
    o[SD_G1 + G1_TZ] = a * (Sxz * w1 + Syz * w2 + P3 * w3) + ht * Tz;

    // cotangents of mu and a, then of T through mu(T) and a(T)
    const real quad = c43 * (Ux * Ux + Vy * Vy + Wz * Wz) + Uy * Uy + Uz * Uz + Vx * Vx +
                      Vz * Vz + Wx * Wx + Wy * Wy - c43 * (Ux * Vy + Ux * Wz + Vy * Wz) +
                      real(2) * (Uy * Vx + Uz * Wx + Vz * Wy);
    const real L1 = c43 * Uxx + Uyy + Uzz + c13 * Vxy + c13 * Wxz;
    const real L2 = Vxx + c43 * Vyy + Vzz + c13 * Uxy + c13 * Wyz;
    const real L3 = Wxx + Wyy + c43 * Wzz + c13 * Uxz + c13 * Vyz;
    const real A1 = Tx * P1 + Ty * Sxy + Tz * Sxz;
    const real A2 = Ty * P2 + Tx * Sxy + Tz * Syz;
    const real A3 = Tz * P3 + Tx * Sxz + Ty * Syz;
    const real mub = w1 * L1 + w2 * L2 + w3 * L3 + vE * quad +
                     ph.cpPr * (Txx + Tyy + Tzz) * vE;
    const real ab = A1 * w1 + A2 * w2 + A3 * w3 + ph.cpPr * (Tx * Tx + Ty * Ty + Tz * Tz) * vE;
    o[SD_TB] = mub * a + ab * (real(-0.24) * a / T);
    // the energy equation's sum_i u_i Rv_i
    o[SD_UB] = vE * (mu * L1 + a * A1);
    o[SD_VB] = vE * (mu * L2 + a * A2);
    o[SD_WB] = vE * (mu * L3 + a * A3);
}

// Chain the primitive cotangents (U, V, W, T) and the flux cotangents
// Fb[3 * c + j] = D_j v_c of one point back to the conserved state (q in
// the stored perturbation form).
template <typename real>
__device__ __forceinline__ void chain_point(const Phys<real> &ph, const real (&q)[5],
                                            const real (&Fb)[15], const real (&ubi)[3], real Tbi,
                                            real (&res)[5])
{
    const real rho = ph.rho_ref + q[0];
    const real ir = real(1) / rho;
    const real m[3] = {q[1], q[2], q[3]};
    const real u[3] = {m[0] * ir, m[1] * ir, m[2] * ir};
    const real E = ph.rhoE_ref + q[4];
    const real pp = ph.gm1 * (q[4] - real(0.5) * (m[0] * u[0] + m[1] * u[1] + m[2] * u[2]));
    const real p = ph.p_ref + pp;
    const real T = ph.T0s + (pp - q[0] * ph.RT0s) * ir * ph.iR;
    real mb[3], ub[3] = {ubi[0], ubi[1], ubi[2]};
    real pb = 0, Eb = 0, rb = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) mb[k] = Fb[k];  // continuity fluxes m_j
#pragma unroll
    for (int ii = 0; ii < 3; ii++) {  // momentum fluxes m_i u_j + p d_ij
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const real fb = Fb[3 * (ii + 1) + j];
            mb[ii] += fb * u[j];
            ub[j] += fb * m[ii];
            if (ii == j) pb += fb;
        }
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {  // energy fluxes (E + p) u_j
        const real fb = Fb[12 + j];
        Eb += fb * u[j];
        pb += fb * u[j];
        ub[j] += fb * (E + p);
    }
    // T = p / (rho R)
    const real Tb = Tbi;
    pb += Tb * ir * ph.iR;
    rb -= Tb * T * ir;
    // p = (gamma - 1)(E - |m|^2 / (2 rho))
    Eb += ph.gm1 * pb;
    const real u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    rb += ph.gm1 * real(0.5) * u2 * pb;
    for (int k = 0; k < 3; k++) mb[k] -= ph.gm1 * u[k] * pb;
    // u = m / rho
    for (int k = 0; k < 3; k++) {
        mb[k] += ub[k] * ir;
        rb -= ub[k] * u[k] * ir;
    }
    res[0] = rb;
    res[1] = mb[0];
    res[2] = mb[1];
    res[3] = mb[2];
    res[4] = Eb;
}

// ------------------------------------------------------------ windows

// The window of the stage state (perturbation form, a quad and the energy)
// and the primitives U, V, W, T' at every window point, by the forward's
// own primitive evaluation.
template <typename real, int W, typename WinT>
__device__ __forceinline__ void prim_windows(const WinT &win, const QSet<real> &q,
                                             const Phys<real> &ph, real (&pw)[4][W],
                                             real (&qw)[5][W])
{
    win.load_quad(q.m, qw[0], qw[1], qw[2], qw[3]);
    win.load(q.e, qw[4]);
#pragma unroll
    for (int t = 0; t < W; t++) {
        real ir, pp;
        prim(ph, qw[0][t], qw[1][t], qw[2][t], qw[3][t], qw[4][t], ir, pw[0][t], pw[1][t],
             pw[2][t], pp, pw[3][t]);
    }
}

// Windows of values computed in registers on row (x) lines, through the LDS
// staging area: each lane puts its own points at their positions along the
// line (field f of line l at (f * LX + l) * n), and after a barrier reads
// the window of a field, wrapping round the periodic line.  Any
// segmentation (ragged lines included); the area needs NF * LX * n <=
// NF * M * lanes elements.  The caller places the barriers: before the
// first put (the area is the exchange ring's), after the last put, and
// after the last read.
template <typename real>
__device__ __forceinline__ void stage_put(const Lane<real, MAP_ROW> &c, int n, int f, int t,
                                          real v)
{
    if (t < c.len) c.xb[(f * c.LX + c.line) * n + c.s + t] = v;
}
template <typename real, int M, int HW>
__device__ __forceinline__ void stage_get(const Lane<real, MAP_ROW> &c, int n, int f,
                                          real (&w)[M + 2 * HW])
{
    const real *st = c.xb + (f * c.LX + c.line) * n;
#pragma unroll
    for (int t = 0; t < M + 2 * HW; t++) {
        int p = c.s - HW + t;
        p += p < 0 ? n : 0;
        p -= p >= n ? n : 0;
        w[t] = st[p];
    }
}

// one point's values of a quad, pair or single field
template <typename real, typename WinT>
__device__ __forceinline__ void pt_quad(const WinT &win, const real *f, int t, real *o)
{
    typedef real vec4 __attribute__((ext_vector_type(4)));
    const vec4 v = *(const vec4 *)WinT::quad_at(f, win.mid(t));
    o[0] = v[0];
    o[1] = v[1];
    o[2] = v[2];
    o[3] = v[3];
}
template <typename real, typename WinT>
__device__ __forceinline__ void pt_pair(const WinT &win, const real *f, int t, real *o)
{
    typedef real vec2 __attribute__((ext_vector_type(2)));
    const vec2 v = *(const vec2 *)WinT::pair_at(f, win.mid(t));
    o[0] = v[0];
    o[1] = v[1];
}

// ------------------------------------------------------------ recompute passes

// y: D1 and D2 of U, V, W, T' along y -> (Uy, Ty), (Vy, Wy), (Uyy, Vyy, Wyy, Tyy)
template <typename real, int M, bool UNI>
__global__ __launch_bounds__(adj_rec_threads<real>(UNI)) void k_adj_yrec(
    const OpDev<real> *__restrict__ d1, const OpDev<real> *__restrict__ d2, LineMap lm,
    AdjVjp<real> a, Phys<real> ph)
{
    constexpr int HW = 2, W = M + 2 * HW;
    CS3D_LANE_SETUP(MAP_COL, d1, sweep_kr(d1, d2), ADJ_R_PLANES)
    (void)n;
    (void)per;
    real r[8][M];
    {
        real pw[4][W], qw[5][W];
        prim_windows(win, a.q, ph, pw, qw);
#pragma unroll
        for (int f = 0; f < 4; f++) {
            rhs<OP_D1, real, M, HW, UNI>(d1, c, pw[f], r[f]);
            rhs<OP_D2, real, M, HW, UNI>(d2, c, pw[f], r[4 + f]);
        }
    }
    solve_d12<4, 4, real, M, UNI>(d1, d2, c, r);
    if (!active) return;
    win.template store_pair<true>(a.rec.dYut, r[0], r[3], c.len);
    win.template store_pair<true>(a.rec.dYvw, r[1], r[2], c.len);
    win.template store_quad<true>(a.rec.dY2, r[4], r[5], r[6], r[7], c.len);
}

// z: D1 and D2 of U, V, W, T' and D1 of the stored Vy, Wy along z ->
// (Uz, Vz, Wz, Tz), (Uzz, Vzz, Wzz, Tzz), (Vyz, Wyz)
template <typename real, int M, bool UNI>
__global__ __launch_bounds__(adj_rec_threads<real>(UNI)) void k_adj_zrec(
    const OpDev<real> *__restrict__ d1, const OpDev<real> *__restrict__ d2, LineMap lm,
    AdjVjp<real> a, Phys<real> ph)
{
    constexpr int HW = 2, W = M + 2 * HW;
    CS3D_LANE_SETUP(MAP_COL, d1, sweep_kr(d1, d2), ADJ_R_PLANES)
    (void)n;
    (void)per;
    real r[10][M];  // D1 of U V W T Vy Wy | D2 of U V W T
    {
        real pw[4][W], qw[5][W];
        prim_windows(win, a.q, ph, pw, qw);
#pragma unroll
        for (int f = 0; f < 4; f++) {
            rhs<OP_D1, real, M, HW, UNI>(d1, c, pw[f], r[f]);
            rhs<OP_D2, real, M, HW, UNI>(d2, c, pw[f], r[6 + f]);
        }
    }
    {
        real mw[2][W];
        win.load_pair(a.rec.dYvw, mw[0], mw[1]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, mw[0], r[4]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, mw[1], r[5]);
    }
    solve_d12<6, 4, real, M, UNI>(d1, d2, c, r);
    if (!active) return;
    win.template store_quad<true>(a.rec.dZ1, r[0], r[1], r[2], r[3], c.len);
    win.template store_quad<true>(a.rec.dZ2, r[6], r[7], r[8], r[9], c.len);
    win.template store_pair<true>(a.rec.mYZ, r[4], r[5], c.len);
}

// ------------------------------------------------------------ x pass

// The x derivatives of the stage (D1 of U, V, W, T and of the stored Uy, Uz,
// Vy, Wz for the mixed derivatives Uxy, Uxz, Vxy, Wxz; D2 of U, V, W, T)
// in one batch; the cotangent seeds at every point from them and the
// stored y and z derivatives; then every x-direction transposed term in a
// second batch over windows of the seeds staged through LDS: D_x of the
// first-derivative seeds g1UX, g1XY, g1XZ, g1TX and of the flux cotangents
// v_c, D2_x of s0, s1, s2, gq, and Xs_i = D_x s_i for the mixed transposes
// the column passes finish.  The pointwise cotangents (ub, vb, wb, Tb) and
// these x terms are chained to the conserved state (resx).  The mixed
// derivatives are taken as D_x D_y U etc.: periodic operators along
// different axes commute.  Values needed after a solve batch are read again
// per point rather than held across it (they come from the caches).
template <typename real, int M, bool UNI>
__global__ __launch_bounds__(ADJ_X_THREADS) void k_adj_x(const OpDev<real> *__restrict__ d1,
                                                         const OpDev<real> *__restrict__ d2,
                                                         LineMap lm, AdjVjp<real> a,
                                                         Phys<real> ph)
{
    constexpr int HW = 2, W = M + 2 * HW;
    CS3D_LANE_SETUP(MAP_ROW, d1, sweep_kr(d1, d2), ADJ_X_PLANES)
    (void)per;
    // r: Ux Vx Wx Tx | Uxy Uxz Vxy Wxz | Uxx Vxx Wxx Txx; own values of the
    // state and of Uy Ty Vy Wy Uz Vz Wz Tz (held across the first solve
    // batch on uniform plans; the table variants, whose solves need more
    // registers, read them again per point)
    constexpr bool HOLD = UNI;
    real r[12][M], qo[5][M], yz[HOLD ? 8 : 1][M];
    {
        real pw[4][W], qw[5][W];
        prim_windows(win, a.q, ph, pw, qw);
#pragma unroll
        for (int f = 0; f < 5; f++)
#pragma unroll
            for (int t = 0; t < M; t++) qo[f][t] = qw[f][HW + t];
#pragma unroll
        for (int f = 0; f < 4; f++) {
            rhs<OP_D1, real, M, HW, UNI>(d1, c, pw[f], r[f]);
            rhs<OP_D2, real, M, HW, UNI>(d2, c, pw[f], r[8 + f]);
        }
    }
    {
        real w0[W], w1[W], w2[W], w3[W];
        win.load_pair(a.rec.dYut, w0, w1);
        if constexpr (HOLD) {
#pragma unroll
            for (int t = 0; t < M; t++) {
                yz[0][t] = w0[HW + t];
                yz[1][t] = w1[HW + t];
            }
        }
        rhs<OP_D1, real, M, HW, UNI>(d1, c, w0, r[4]);
        win.load_pair(a.rec.dYvw, w0, w1);
        if constexpr (HOLD) {
#pragma unroll
            for (int t = 0; t < M; t++) {
                yz[2][t] = w0[HW + t];
                yz[3][t] = w1[HW + t];
            }
        }
        rhs<OP_D1, real, M, HW, UNI>(d1, c, w0, r[6]);
        win.load_quad(a.rec.dZ1, w0, w1, w2, w3);
        if constexpr (HOLD) {
#pragma unroll
            for (int t = 0; t < M; t++) {
                yz[4][t] = w0[HW + t];
                yz[5][t] = w1[HW + t];
                yz[6][t] = w2[HW + t];
                yz[7][t] = w3[HW + t];
            }
        }
        rhs<OP_D1, real, M, HW, UNI>(d1, c, w0, r[5]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, w2, r[7]);
    }
    solve_d12<8, 4, real, M, UNI>(d1, d2, c, r);
    // x: D1 of g1UX g1XY g1XZ g1TX s0 s1 s2, D1 of v0..v4 at XV, D2 of s0 s1
    // s2 gq at X2.  FP32 solves them in one batch; FP64 in two (7 + 4, then
    // 5), whose carry reads (11 terms each) would not fit the registers at
    // once.  The v windows come first: their centres are the seeds' v.
    constexpr bool ONE = sizeof(real) == 4;
    constexpr int XV = ONE ? 7 : 11, X2 = ONE ? 12 : 7;
    real x[16][M], vo[4][M];
    {
        real vw[5][W];
        win.load_quad(a.v.m, vw[0], vw[1], vw[2], vw[3]);
        win.load(a.v.e, vw[4]);
#pragma unroll
        for (int k = 0; k < 5; k++)
#pragma unroll
            for (int t = 0; t < W; t++) vw[k][t] *= a.cv;
#pragma unroll
        for (int k = 0; k < 5; k++) rhs<OP_D1, real, M, HW, UNI>(d1, c, vw[k], x[XV + k]);
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int t = 0; t < M; t++) vo[k][t] = vw[k + 1][HW + t];
    }
    // seeds, point by point: the eight the x solves need go to the window
    // staging, the ones needed after the solve wait in LDS past the ring and
    // the staging (the host sizes the area: adj_x_lds) at this lane's slots
    constexpr int NS = ADJ_X_STAGE;
    const int stg[NS] = {SD_G1 + G1_UX, SD_G1 + G1_XY, SD_G1 + G1_XZ, SD_G1 + G1_TX,
                         SD_S,          SD_S + 1,      SD_S + 2,      SD_GQ};
    enum { K_UB, K_VB, K_WB, K_TB, K_XZ, K_WZ, K_TZ, K_XY, K_VY, K_TY, K_YZ,
           K_S0, K_S1, K_S2, K_GQ, K_COUNT };
    static_assert(K_COUNT == ADJ_X_KEEP, "kept values and their LDS area differ");
    const int kept[K_COUNT] = {SD_UB,         SD_VB,         SD_WB,         SD_TB,
                               SD_G1 + G1_XZ, SD_G1 + G1_WZ, SD_G1 + G1_TZ, SD_G1 + G1_XY,
                               SD_G1 + G1_VY, SD_G1 + G1_TY, SD_G1 + G1_YZ, SD_S,
                               SD_S + 1,      SD_S + 2,      SD_GQ};
    real *kp = xb + (2 * c.cap > NS * M * c.LX * c.P ? 2 * c.cap : NS * M * c.LX * c.P);
    const int T = c.LX * c.P, me = c.line * c.P + c.g;
    auto kslot = [&](int k, int t) -> real & { return kp[(k * M + t) * T + me]; };
    __syncthreads();
#pragma unroll
    for (int t = 0; t < M; t++) {
        real y2[4], z2[4], myz[2], yt[8];
        pt_quad(win, a.rec.dY2, t, y2);
        pt_quad(win, a.rec.dZ2, t, z2);
        pt_pair(win, a.rec.mYZ, t, myz);
        if constexpr (HOLD) {
#pragma unroll
            for (int k = 0; k < 8; k++) yt[k] = yz[k][t];
        } else {
            pt_pair(win, a.rec.dYut, t, yt);
            pt_pair(win, a.rec.dYvw, t, yt + 2);
            pt_quad(win, a.rec.dZ1, t, yt + 4);
        }
        SeedPt<real> p;
        real ir, pp;
        prim(ph, qo[0][t], qo[1][t], qo[2][t], qo[3][t], qo[4][t], ir, p.U, p.V, p.W, pp, p.Tp);
        visc(ph, ph.T0s + p.Tp, p.mu, p.a);
        p.v1 = vo[0][t];
        p.v2 = vo[1][t];
        p.v3 = vo[2][t];
        p.vE = vo[3][t];
#pragma unroll
        for (int f = 0; f < 4; f++) {
            p.d1[3 * f] = r[f][t];
            p.d2[3 * f] = r[8 + f][t];
            p.d2[3 * f + 1] = y2[f];
            p.d2[3 * f + 2] = z2[f];
        }
        p.d1[1] = yt[0];   // Uy
        p.d1[2] = yt[4];   // Uz
        p.d1[4] = yt[2];   // Vy
        p.d1[5] = yt[5];   // Vz
        p.d1[7] = yt[3];   // Wy
        p.d1[8] = yt[6];   // Wz
        p.d1[10] = yt[1];  // Ty
        p.d1[11] = yt[7];  // Tz
        p.mx[0] = r[4][t];    // Uxy
        p.mx[1] = r[5][t];    // Uxz
        p.mx[2] = r[6][t];    // Vxy
        p.mx[3] = myz[0];     // Vyz
        p.mx[4] = r[7][t];    // Wxz
        p.mx[5] = myz[1];     // Wyz
        real o[SD_COUNT];
        seed_point(ph, p, o);
#pragma unroll
        for (int f = 0; f < NS; f++) stage_put(c, n, f, t, o[stg[f]]);
#pragma unroll
        for (int k = 0; k < K_COUNT; k++) kslot(k, t) = o[kept[k]];
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < NS; f++) {
        real w[W];
        stage_get<real, M, HW>(c, n, f, w);
        if (f < 7) rhs<OP_D1, real, M, HW, UNI>(d1, c, w, x[f]);
        if (f >= 4) rhs<OP_D2, real, M, HW, UNI>(d2, c, w, x[X2 + f - 4]);
    }
    __syncthreads();
    if constexpr (ONE) {
        solve_d12<12, 4, real, M, UNI>(d1, d2, c, x);
    } else {
        solve_d12<7, 4, real, M, UNI>(d1, d2, c, *reinterpret_cast<real(*)[11][M]>(&x[0]));
        solve_d12<5, 0, real, M, UNI>(d1, d2, c, *reinterpret_cast<real(*)[5][M]>(&x[11]));
    }
    const real c43 = real(4) / real(3);
    real res[5][M];
#pragma unroll
    for (int t = 0; t < M; t++) {
        real q[5], Fb[15], ub[3];
#pragma unroll
        for (int k = 0; k < 5; k++) {
            q[k] = qo[k][t];
            Fb[3 * k] = x[XV + k][t];
            Fb[3 * k + 1] = 0;
            Fb[3 * k + 2] = 0;
        }
        ub[0] = kslot(K_UB, t) - x[0][t] + c43 * x[X2][t];
        ub[1] = kslot(K_VB, t) - x[1][t] + x[X2 + 1][t];
        ub[2] = kslot(K_WB, t) - x[2][t] + x[X2 + 2][t];
        const real Tb = kslot(K_TB, t) - x[3][t] + x[X2 + 3][t];
        real rp[5];
        chain_point(ph, q, Fb, ub, Tb, rp);
#pragma unroll
        for (int k = 0; k < 5; k++) res[k][t] = rp[k];
    }
    if (!active) return;
    win.store_quad(a.resx.m, res[0], res[1], res[2], res[3], c.len);
    win.store(a.resx.e, res[4], c.len);
    real g[K_COUNT][M];
#pragma unroll
    for (int k = K_XZ; k < K_COUNT; k++)
#pragma unroll
        for (int t = 0; t < M; t++) g[k][t] = kslot(k, t);
    win.store_quad(a.sd.sq, g[K_S0], g[K_S1], g[K_S2], g[K_GQ], c.len);
    win.store_quad(a.sd.gz, g[K_XZ], g[K_WZ], g[K_TZ], x[6], c.len);
    win.store_quad(a.sd.gy, g[K_XY], g[K_VY], g[K_TY], x[5], c.len);
    win.store_pair(a.sd.gs, g[K_YZ], x[4], c.len);
}

// ------------------------------------------------------------ column passes

// z: the z-direction transposed terms, D1 of v_c (flux), D2 of s0..gq, D1
// of g1XZ, g1WZ, g1TZ, Xs2, g1YZ, Xs0 and of s1, s2 (mzz, for the y pass);
// chained and added to resx -> resxz.
template <typename real, int M, bool UNI>
__global__ __launch_bounds__(adj_col_threads<real>()) void k_adj_z(
    const OpDev<real> *__restrict__ d1, const OpDev<real> *__restrict__ d2, LineMap lm,
    AdjVjp<real> a, Phys<real> ph)
{
    constexpr int HW = 2, W = M + 2 * HW;
    CS3D_LANE_SETUP(MAP_COL, d1, sweep_kr(d1, d2), ADJ_T_PLANES)
    (void)n;
    (void)per;
    real f1[9][M];  // D1 of v0..v4 | D2 of s0 s1 s2 gq
    real f2[8][M];  // D1 of g1XZ g1WZ g1TZ Xs2 g1YZ Xs0 s1 s2
    {
        real w[4][W];
        win.load_quad(a.sd.sq, w[0], w[1], w[2], w[3]);
#pragma unroll
        for (int k = 0; k < 4; k++) rhs<OP_D2, real, M, HW, UNI>(d2, c, w[k], f1[5 + k]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, w[1], f2[6]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, w[2], f2[7]);
        win.load_quad(a.sd.gz, w[0], w[1], w[2], w[3]);
#pragma unroll
        for (int k = 0; k < 4; k++) rhs<OP_D1, real, M, HW, UNI>(d1, c, w[k], f2[k]);
        win.load_pair(a.sd.gs, w[0], w[1]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, w[0], f2[4]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, w[1], f2[5]);
        win.load_quad(a.v.m, w[0], w[1], w[2], w[3]);
#pragma unroll
        for (int k = 0; k < 4; k++) rhs<OP_D1, real, M, HW, UNI>(d1, c, w[k], f1[k]);
        win.load(a.v.e, w[0]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, w[0], f1[4]);
#pragma unroll
        for (int k = 0; k < 5; k++)
#pragma unroll
            for (int t = 0; t < M; t++) f1[k][t] *= a.cv;
    }
    solve_d12<5, 4, real, M, UNI>(d1, d2, c, f1);
    solve_d12<8, 0, real, M, UNI>(d1, d2, c, f2);
    real qv[5][M], rx[5][M];
    win.load_seg_quad(a.q.m, qv[0], qv[1], qv[2], qv[3]);
    win.load_seg(a.q.e, qv[4]);
    win.load_seg_quad(a.resx.m, rx[0], rx[1], rx[2], rx[3]);
    win.load_seg(a.resx.e, rx[4]);
    const real c13 = real(1) / real(3), c43 = real(4) / real(3);
#pragma unroll
    for (int t = 0; t < M; t++) {
        real q[5], Fb[15], ub[3];
#pragma unroll
        for (int k = 0; k < 5; k++) {
            q[k] = qv[k][t];
            Fb[3 * k] = 0;
            Fb[3 * k + 1] = 0;
            Fb[3 * k + 2] = f1[k][t];
        }
        ub[0] = f1[5][t] - f2[0][t] + c13 * f2[3][t];
        ub[1] = f1[6][t] - f2[4][t];
        ub[2] = c43 * f1[7][t] - f2[1][t] + c13 * f2[5][t];
        const real Tb = f1[8][t] - f2[2][t];
        real rp[5];
        chain_point(ph, q, Fb, ub, Tb, rp);
#pragma unroll
        for (int k = 0; k < 5; k++) rx[k][t] += rp[k];
    }
    if (!active) return;
    win.template store_quad<true>(a.resxz.m, rx[0], rx[1], rx[2], rx[3], c.len);
    win.template store<true>(a.resxz.e, rx[4], c.len);
    win.template store_pair<true>(a.sd.mzz, f2[6], f2[7], c.len);
}

// y: the y-direction transposed terms, D1 of v_c, D2 of s0..gq, D1 of
// g1XY, g1VY, g1TY, Xs1, g1YZ, Xs0 and of D_z s1, D_z s2; chained and added
// to resxz -> res, then the epilogue (AdjVjp).
template <typename real, int M, bool UNI>
__global__ __launch_bounds__(adj_col_threads<real>()) void k_adj_y(
    const OpDev<real> *__restrict__ d1, const OpDev<real> *__restrict__ d2, LineMap lm,
    AdjVjp<real> a, Phys<real> ph)
{
    constexpr int HW = 2, W = M + 2 * HW;
    CS3D_LANE_SETUP(MAP_COL, d1, sweep_kr(d1, d2), ADJ_T_PLANES)
    (void)n;
    (void)per;
    real f1[9][M];  // D1 of v0..v4 | D2 of s0 s1 s2 gq
    real f2[8][M];  // D1 of g1XY g1VY g1TY Xs1 g1YZ Xs0 Dz s1 Dz s2
    {
        real w[4][W];
        win.load_quad(a.sd.sq, w[0], w[1], w[2], w[3]);
#pragma unroll
        for (int k = 0; k < 4; k++) rhs<OP_D2, real, M, HW, UNI>(d2, c, w[k], f1[5 + k]);
        win.load_quad(a.sd.gy, w[0], w[1], w[2], w[3]);
#pragma unroll
        for (int k = 0; k < 4; k++) rhs<OP_D1, real, M, HW, UNI>(d1, c, w[k], f2[k]);
        win.load_pair(a.sd.gs, w[0], w[1]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, w[0], f2[4]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, w[1], f2[5]);
        win.load_pair(a.sd.mzz, w[0], w[1]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, w[0], f2[6]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, w[1], f2[7]);
        win.load_quad(a.v.m, w[0], w[1], w[2], w[3]);
#pragma unroll
        for (int k = 0; k < 4; k++) rhs<OP_D1, real, M, HW, UNI>(d1, c, w[k], f1[k]);
        win.load(a.v.e, w[0]);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, w[0], f1[4]);
#pragma unroll
        for (int k = 0; k < 5; k++)
#pragma unroll
            for (int t = 0; t < M; t++) f1[k][t] *= a.cv;
    }
    solve_d12<5, 4, real, M, UNI>(d1, d2, c, f1);
    solve_d12<8, 0, real, M, UNI>(d1, d2, c, f2);
    real qv[5][M], rx[5][M];
    win.load_seg_quad(a.q.m, qv[0], qv[1], qv[2], qv[3]);
    win.load_seg(a.q.e, qv[4]);
    win.load_seg_quad(a.resxz.m, rx[0], rx[1], rx[2], rx[3]);
    win.load_seg(a.resxz.e, rx[4]);
    const real c13 = real(1) / real(3), c43 = real(4) / real(3);
#pragma unroll
    for (int t = 0; t < M; t++) {
        real q[5], Fb[15], ub[3];
#pragma unroll
        for (int k = 0; k < 5; k++) {
            q[k] = qv[k][t];
            Fb[3 * k] = 0;
            Fb[3 * k + 1] = f1[k][t];
            Fb[3 * k + 2] = 0;
        }
        ub[0] = f1[5][t] - f2[0][t] + c13 * f2[3][t];
        ub[1] = c43 * f1[6][t] - f2[1][t] + c13 * (f2[5][t] + f2[7][t]);
        ub[2] = f1[7][t] - f2[4][t] + c13 * f2[6][t];
        const real Tb = f1[8][t] - f2[2][t];
        real rp[5];
        chain_point(ph, q, Fb, ub, Tb, rp);
#pragma unroll
        for (int k = 0; k < 5; k++) rx[k][t] += rp[k];
    }
    if (!active) return;
    if (a.out.m) {
        win.store_quad(a.out.m, rx[0], rx[1], rx[2], rx[3], c.len);
        win.store(a.out.e, rx[4], c.len);
    }
    if (a.acc.m) {
        real b[5][M];
        if (a.acc_base.m) {
            win.load_seg_quad(a.acc_base.m, b[0], b[1], b[2], b[3]);
            win.load_seg(a.acc_base.e, b[4]);
        } else {
#pragma unroll
            for (int k = 0; k < 5; k++)
#pragma unroll
                for (int t = 0; t < M; t++) b[k][t] = 0;
        }
#pragma unroll
        for (int k = 0; k < 5; k++)
#pragma unroll
            for (int t = 0; t < M; t++) b[k][t] += rx[k][t];
        win.store_quad(a.acc.m, b[0], b[1], b[2], b[3], c.len);
        win.store(a.acc.e, b[4], c.len);
    }
    if (a.vn.m) {
        real l[5][M];
        win.load_seg_quad(a.la.m, l[0], l[1], l[2], l[3]);
        win.load_seg(a.la.e, l[4]);
#pragma unroll
        for (int k = 0; k < 5; k++)
#pragma unroll
            for (int t = 0; t < M; t++) l[k][t] = a.ca * l[k][t] + a.cb * rx[k][t];
        win.store_quad(a.vn.m, l[0], l[1], l[2], l[3], c.len);
        win.store(a.vn.e, l[4], c.len);
    }
}

template <typename real, bool UNI>
static hipError_t launch_pass(int pass, const OpDev<real> *d1, const OpDev<real> *d2,
                              const LaunchGeom &g, const LineMap &lm, const AdjVjp<real> &a,
                              const Phys<real> &ph, hipStream_t st)
{
    constexpr int M = SEG_DERIV;
    dim3 grid((lm.nlines + g.LX - 1) / g.LX), block(g.threads);
    switch (pass) {
    case ADJ_YREC:
        hipLaunchKernelGGL((k_adj_yrec<real, M, UNI>), grid, block, g.shm, st, d1, d2, lm, a, ph);
        break;
    case ADJ_ZREC:
        hipLaunchKernelGGL((k_adj_zrec<real, M, UNI>), grid, block, g.shm, st, d1, d2, lm, a, ph);
        break;
    case ADJ_X:
        hipLaunchKernelGGL((k_adj_x<real, M, UNI>), grid, block, g.shm, st, d1, d2, lm, a, ph);
        break;
    case ADJ_Z:
        hipLaunchKernelGGL((k_adj_z<real, M, UNI>), grid, block, g.shm, st, d1, d2, lm, a, ph);
        break;
    default:
        hipLaunchKernelGGL((k_adj_y<real, M, UNI>), grid, block, g.shm, st, d1, d2, lm, a, ph);
    }
    return hipGetLastError();
}

// the lanes each pass's launch bound allows
template <typename real>
static int pass_threads(int pass, bool uni)
{
    if (pass == ADJ_X) return ADJ_X_THREADS;
    if (pass == ADJ_YREC || pass == ADJ_ZREC) return adj_rec_threads<real>(uni);
    return adj_col_threads<real>();
}

template <typename real>
hipError_t launch_adj_pass(int pass, const OpDev<real> *d1, const OpDev<real> *d2, int uni, int M,
                           const LaunchGeom &g, const LineMap &lm, const AdjVjp<real> &a,
                           const Phys<real> &ph, hipStream_t st)
{
    if (M != SEG_DERIV || pass < ADJ_YREC || pass > ADJ_Y) return hipErrorInvalidValue;
    if (g.LX < 1 || g.threads < 1 || g.threads > pass_threads<real>(pass, uni != 0))
        return hipErrorInvalidConfiguration;
    return uni ? launch_pass<real, true>(pass, d1, d2, g, lm, a, ph, st)
               : launch_pass<real, false>(pass, d1, d2, g, lm, a, ph, st);
}

// ------------------------------------------------------------ pointwise

// Gradient of J = (1/N) sum rho |u|^2 / 2 with respect to the conserved
// state, and the map of a conserved cotangent to the initial primitive
// velocities at fixed rho and p: m_i = rho u_i, E = p/(gamma-1) + rho|u|^2/2.
template <typename real>
__global__ void k_adj_ke_seed(QSet<real> q, Phys<real> ph, QSet<real> out, long N)
{
    const real iN = real(1) / real(N);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        real q5[5];
        qset_load(q, i, q5);
        const real rho = ph.rho_ref + q5[0];
        const real u = q5[1] / rho, v = q5[2] / rho, w = q5[3] / rho;
        const real o[5] = {-real(0.5) * (u * u + v * v + w * w) * iN, u * iN, v * iN, w * iN, 0};
        qset_store(out, i, o);
    }
}

template <typename real>
__global__ void k_adj_to_velocity(QSet<real> q, QSet<real> qb, Phys<real> ph, real *gU, real *gV,
                                  real *gW, long N)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        real q5[5], b[5];
        qset_load(q, i, q5);
        qset_load(qb, i, b);
        const real rho = ph.rho_ref + q5[0];
        gU[i] = rho * (b[1] + (q5[1] / rho) * b[4]);
        gV[i] = rho * (b[2] + (q5[2] / rho) * b[4]);
        gW[i] = rho * (b[3] + (q5[3] / rho) * b[4]);
    }
}

static dim3 grid_for(long N)
{
    long nb = (N + 255) / 256;
    if (nb > 8192) nb = 8192;
    return dim3((unsigned)nb);
}

template <typename real>
hipError_t launch_adj_ke_seed(const QSet<real> &q, const Phys<real> &ph, const QSet<real> &out,
                              long N, hipStream_t st)
{
    hipLaunchKernelGGL((k_adj_ke_seed<real>), grid_for(N), dim3(256), 0, st, q, ph, out, N);
    return hipGetLastError();
}
template <typename real>
hipError_t launch_adj_to_velocity(const QSet<real> &q, const QSet<real> &qb, const Phys<real> &ph,
                                  real *gU, real *gV, real *gW, long N, hipStream_t st)
{
    hipLaunchKernelGGL((k_adj_to_velocity<real>), grid_for(N), dim3(256), 0, st, q, qb, ph, gU, gV,
                       gW, N);
    return hipGetLastError();
}

// ------------------------------------------------------------ general path
//
// Stage products on any boundary set the forward covers without walls
// (non-periodic lines with their closures, sponge faces): the derivative
// fields of the stage state are recomputed with the standalone operators,
// the seeds formed point by point, every transposed term applied with the
// transposed operators (adjoint.hpp: GenAdj), and the chain to the
// conserved state taken point by point with the sponge terms and the RK
// epilogue.

// whether flat point i lies on a sponge face plane
__device__ __forceinline__ bool on_sponge_plane(const StageArgs &sa, long i)
{
    const int ii = (int)(i % sa.Nx), jj = (int)((i / sa.Nx) % sa.Ny),
              kk = (int)(i / ((long)sa.Nx * sa.Ny));
    return (sa.sponge_face[0] && ii == 0) || (sa.sponge_face[1] && ii == sa.Nx - 1) ||
           (sa.sponge_face[2] && jj == 0) || (sa.sponge_face[3] && jj == sa.Ny - 1) ||
           (sa.sponge_face[4] && kk == 0) || (sa.sponge_face[5] && kk == sa.Nz - 1);
}

// the normal direction of the last wall face through flat point i (faces
// X0..Z1 in turn, as postStepBCHandling visits them), -1 off the walls
__device__ __forceinline__ int wall_row_dir(const StageArgs &sa, long i)
{
    const int ii = (int)(i % sa.Nx), jj = (int)((i / sa.Nx) % sa.Ny),
              kk = (int)(i / ((long)sa.Nx * sa.Ny));
    if ((sa.wall_face[4] && kk == 0) || (sa.wall_face[5] && kk == sa.Nz - 1)) return 2;
    if ((sa.wall_face[2] && jj == 0) || (sa.wall_face[3] && jj == sa.Ny - 1)) return 1;
    if ((sa.wall_face[0] && ii == 0) || (sa.wall_face[1] && ii == sa.Nx - 1)) return 0;
    return -1;
}

// the state the x block's primitives come from at point i: the stage state,
// or at the first stage after a sponge update the pre-reset values on the
// sponge planes
template <typename real>
__device__ __forceinline__ void xset_state(const GenAdj<real> &g, long i, const real (&q)[5],
                                           real (&qx)[5], bool &split)
{
    split = g.stale && on_sponge_plane(g.sa, i);
#pragma unroll
    for (int k = 0; k < 5; k++) qx[k] = q[k];
    if (split) qset_load(g.qstale, i, qx);
}

template <typename real>
__global__ __launch_bounds__(256) void k_gadj_prims(GenAdj<real> g, Phys<real> ph, long N)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        real q[5], qx[5], ir, pp;
        bool split;
        qset_load(g.q, i, q);
        xset_state(g, i, q, qx, split);
        prim(ph, q[0], q[1], q[2], q[3], q[4], ir, g.pY[0][i], g.pY[1][i], g.pY[2][i], pp,
             g.pY[3][i]);
        if (g.pX[0] != g.pY[0])
            prim(ph, qx[0], qx[1], qx[2], qx[3], qx[4], ir, g.pX[0][i], g.pX[1][i], g.pX[2][i],
                 pp, g.pX[3][i]);
    }
}

// Seeds of every derivative field (seed_point), the scaled and masked stage
// cotangent, and the starting values of the primitive cotangents (the
// pointwise terms; the transposed operators add the rest).
template <typename real>
__global__ __launch_bounds__(256) void k_gadj_seeds(GenAdj<real> g, Phys<real> ph, long N)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        real q[5], v[5];
        qset_load(g.q, i, q);
        qset_load(g.v, i, v);
        const real mask = on_sponge_plane(g.sa, i) ? real(0) : g.cv;
        // postStepBCHandling: a wall plane keeps only -dt D_n(rho u_n) of
        // its last wall face (X0..Z1); every other term and seed is zero
        const int wd = wall_row_dir(g.sa, i);
        if (g.vc[0] != g.vs[0]) {
#pragma unroll
            for (int d = 0; d < 3; d++) g.vc[d][i] = (wd < 0 || wd == d) ? v[0] * mask : real(0);
        }
        if (wd >= 0)
#pragma unroll
            for (int k = 0; k < 5; k++) v[k] = 0;
        SeedPt<real> p;
        real ir, pp, Tp;
        prim(ph, q[0], q[1], q[2], q[3], q[4], ir, p.U, p.V, p.W, pp, Tp);
        p.Tp = g.pX[3][i];
        visc(ph, ph.T0s + p.Tp, p.mu, p.a);
#pragma unroll
        for (int k = 0; k < 5; k++) {
            v[k] *= mask;
            g.vs[k][i] = v[k];
        }
        p.v1 = v[1];
        p.v2 = v[2];
        p.v3 = v[3];
        p.vE = v[4];
#pragma unroll
        for (int k = 0; k < 12; k++) {
            p.d1[k] = g.d1[k][i];
            p.d2[k] = g.d2[k][i];
        }
#pragma unroll
        for (int k = 0; k < 6; k++) p.mx[k] = g.mx[k][i];
        real o[SD_COUNT];
        seed_point(ph, p, o);
#pragma unroll
        for (int k = 0; k < 3; k++) g.s[k][i] = o[SD_S + k];
        g.s[3][i] = o[SD_GQ];
#pragma unroll
        for (int k = 0; k < G1_COUNT; k++) g.g1[k][i] = o[SD_G1 + k];
        g.bY[0][i] = o[SD_UB];
        g.bY[1][i] = o[SD_VB];
        g.bY[2][i] = o[SD_WB];
        g.bY[3][i] = 0;
#pragma unroll
        for (int k = 0; k < 3; k++) g.bX[k][i] = 0;
        g.bX[3][i] = o[SD_TB];
    }
}

// Primitive cotangents (U, V, W, T) of one set into its conserved
// cotangents: T = p / (rho R), p = (gamma - 1)(E - |m|^2 / (2 rho)) as a
// function of the conserved values, u = m / rho.  ub and pb carry the flux
// parts already.
template <typename real>
__device__ __forceinline__ void chain_prims(const Phys<real> &ph, const real (&q)[5],
                                            real (&ub)[3], real pb, real Tb, real (&res)[5])
{
    const real rho = ph.rho_ref + q[0];
    const real ir = real(1) / rho;
    const real u[3] = {q[1] * ir, q[2] * ir, q[3] * ir};
    const real pp = ph.gm1 * (q[4] - real(0.5) * (q[1] * u[0] + q[2] * u[1] + q[3] * u[2]));
    const real T = ph.T0s + (pp - q[0] * ph.RT0s) * ir * ph.iR;
    pb += Tb * ir * ph.iR;
    res[0] -= Tb * T * ir;
    res[4] += ph.gm1 * pb;
    res[0] += ph.gm1 * real(0.5) * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * pb;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        res[1 + k] += ub[k] * ir - ph.gm1 * u[k] * pb;
        res[0] -= ub[k] * u[k] * ir;
    }
}

// The chain to the conserved state: the flux cotangents F_d[c] of direction
// d (velocity and pressure of the x block's set along x, of the stage state
// along y and z; the conserved factors always of the stage state), the
// primitive cotangents of both sets, the sponge terms; then the epilogue.
template <typename real>
__global__ __launch_bounds__(256) void k_gadj_chain(GenAdj<real> g, AdjVjp<real> a, Phys<real> ph, long N)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        real q[5], qx[5];
        bool split;
        qset_load(g.q, i, q);
        xset_state(g, i, q, qx, split);
        real res[5] = {0, 0, 0, 0, 0}, rsx[5] = {0, 0, 0, 0, 0};
        real ub[3] = {g.bY[0][i], g.bY[1][i], g.bY[2][i]}, pb = 0;
        real ubx[3] = {g.bX[0][i], g.bX[1][i], g.bX[2][i]}, pbx = 0;
        const real rho = ph.rho_ref + q[0], ir = real(1) / rho;
        const real E = ph.rhoE_ref + q[4];
        const real m[3] = {q[1], q[2], q[3]};
        const real rhox = ph.rho_ref + qx[0], irx = real(1) / rhox;
        const real pY = ph.p_ref + ph.gm1 * (q[4] - real(0.5) * (m[0] * m[0] + m[1] * m[1] +
                                                                m[2] * m[2]) * ir);
        const real pX = ph.p_ref + ph.gm1 * (qx[4] - real(0.5) * (qx[1] * qx[1] + qx[2] * qx[2] +
                                                                 qx[3] * qx[3]) * irx);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            real Fb[5];
#pragma unroll
            for (int c = 0; c < 5; c++) Fb[c] = g.F[d][c][i];
            const real ud = d == 0 ? qx[1] * irx : m[d] * ir;
            const real p = d == 0 ? pX : pY;
            real &ubd = d == 0 ? ubx[0] : ub[d];
            real &pbd = d == 0 ? pbx : pb;
            res[1 + d] += Fb[0];  // continuity flux m_d
#pragma unroll
            for (int k = 0; k < 3; k++) {  // momentum fluxes m_k u_d + p d_kd
                res[1 + k] += Fb[1 + k] * ud;
                ubd += Fb[1 + k] * m[k];
            }
            pbd += Fb[1 + d];
            res[4] += Fb[4] * ud;  // energy flux (E + p) u_d
            pbd += Fb[4] * ud;
            ubd += Fb[4] * (E + p);
        }
        chain_prims(ph, q, ub, pb, g.bY[3][i], res);
        chain_prims(ph, qx, ubx, pbx, g.bX[3][i], rsx);
        if (!split) {
#pragma unroll
            for (int k = 0; k < 5; k++) {
                res[k] += rsx[k];
                rsx[k] = 0;
            }
        }
        if (g.stale) qset_store(g.res_stale, i, rsx);
        if (g.sigma) {
            // the source is part of the right-hand side that the sponge
            // planes and the wall rows zero
            real w[5], ra[5];
            qset_load(g.v, i, w);
            const real mask =
                on_sponge_plane(g.sa, i) || wall_row_dir(g.sa, i) >= 0 ? real(0) : g.cv;
            qset_load(g.resA, i, ra);
#pragma unroll
            for (int k = 0; k < 5; k++) {
                const real sp = (k == 4 ? real(2) : real(1)) * g.sigma[i] * mask * w[k];
                res[k] -= sp;
                ra[k] += sp;
            }
            qset_store(g.resA, i, ra);
        }
        if (a.out.m) qset_store(a.out, i, res);
        if (a.acc.m) {
            real b[5] = {0, 0, 0, 0, 0};
            if (a.acc_base.m) qset_load(a.acc_base, i, b);
#pragma unroll
            for (int k = 0; k < 5; k++) b[k] += res[k];
            qset_store(a.acc, i, b);
        }
        if (a.vn.m) {
            real l[5];
            qset_load(a.la, i, l);
#pragma unroll
            for (int k = 0; k < 5; k++) l[k] = a.ca * l[k] + a.cb * res[k];
            qset_store(a.vn, i, l);
        }
    }
}

// Reverse of updateSponge (CSolver_AWS.cpp:1937-2031) after a step whose
// filtered state was qf and whose averages were A: lq (cotangent of the
// new state, in place) becomes that of qf, lA (in place) that of A.  The
// planes' values came from the new averages; a = A + eps (qf - A) before
// the energy blend a4' = epsP a4 + (1 - epsP)(pS/(gamma - 1) + |a_m|^2 /
// (2 a_rho)).  ls (null: none): the stale cotangent of qf from the next
// step's first stage.
template <typename real>
__global__ __launch_bounds__(256) void k_gadj_sponge_post(QSet<real> qf, QSet<real> A, QSet<real> lq, QSet<real> lA,
                                   QSet<real> ls, StageArgs sa, Phys<real> ph, real eps,
                                   real epsP, long N)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        real f[5], av[5], l[5], la[5];
        qset_load(qf, i, f);
        qset_load(A, i, av);
        qset_load(lq, i, l);
        qset_load(lA, i, la);
        const bool plane = on_sponge_plane(sa, i);
        real a[5];
#pragma unroll
        for (int k = 0; k < 5; k++) {
            a[k] = av[k] + eps * (f[k] - av[k]);
            if (plane) {
                la[k] += l[k];
                l[k] = 0;
            }
        }
        const real rho = ph.rho_ref + a[0];
        const real c = (real(1) - epsP) * la[4];
        la[4] = epsP * la[4];
#pragma unroll
        for (int k = 1; k < 4; k++) la[k] += c * a[k] / rho;
        la[0] -= c * real(0.5) * (a[1] * a[1] + a[2] * a[2] + a[3] * a[3]) / (rho * rho);
        if (ls.m) {
            real s[5];
            qset_load(ls, i, s);
#pragma unroll
            for (int k = 0; k < 5; k++) l[k] += s[k];
        }
#pragma unroll
        for (int k = 0; k < 5; k++) {
            l[k] += eps * la[k];
            la[k] *= real(1) - eps;
        }
        qset_store(lq, i, l);
        qset_store(lA, i, la);
    }
}

// b += a over a conserved set
template <typename real>
__global__ __launch_bounds__(256) void k_gadj_add(QSet<real> b, QSet<real> a, long N)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        real x[5], y[5];
        qset_load(b, i, x);
        qset_load(a, i, y);
#pragma unroll
        for (int k = 0; k < 5; k++) x[k] += y[k];
        qset_store(b, i, x);
    }
}

template <typename real>
__global__ __launch_bounds__(256) void k_gadj_lincomb(QSet<real> out, real ca, QSet<real> a,
                                                      real cb, QSet<real> b, long N)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        real x[5], y[5] = {0, 0, 0, 0, 0};
        qset_load(a, i, x);
        if (b.m) qset_load(b, i, y);
#pragma unroll
        for (int k = 0; k < 5; k++) x[k] = ca * x[k] + (b.m ? cb * y[k] : real(0));
        qset_store(out, i, x);
    }
}

template <typename real>
hipError_t launch_gadj_lincomb(const QSet<real> &out, real ca, const QSet<real> &a, real cb,
                               const QSet<real> &b, long N, hipStream_t st)
{
    hipLaunchKernelGGL((k_gadj_lincomb<real>), grid_for(N), dim3(256), 0, st, out, ca, a, cb, b, N);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_gadj(int which, const GenAdj<real> &g, const AdjVjp<real> &a,
                       const Phys<real> &ph, long N, hipStream_t st)
{
    const dim3 grid = grid_for(N), block(256);
    switch (which) {
    case GADJ_PRIMS: hipLaunchKernelGGL((k_gadj_prims<real>), grid, block, 0, st, g, ph, N); break;
    case GADJ_SEEDS: hipLaunchKernelGGL((k_gadj_seeds<real>), grid, block, 0, st, g, ph, N); break;
    case GADJ_CHAIN:
        hipLaunchKernelGGL((k_gadj_chain<real>), grid, block, 0, st, g, a, ph, N);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename real>
hipError_t launch_gadj_sponge_post(const QSet<real> &qf, const QSet<real> &A,
                                   const QSet<real> &lq, const QSet<real> &lA,
                                   const QSet<real> &ls, const StageArgs &sa,
                                   const Phys<real> &ph, double eps, double epsP, long N,
                                   hipStream_t st)
{
    hipLaunchKernelGGL((k_gadj_sponge_post<real>), grid_for(N), dim3(256), 0, st, qf, A, lq, lA,
                       ls, sa, ph, (real)eps, (real)epsP, N);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_gadj_add(const QSet<real> &b, const QSet<real> &a, long N, hipStream_t st)
{
    hipLaunchKernelGGL((k_gadj_add<real>), grid_for(N), dim3(256), 0, st, b, a, N);
    return hipGetLastError();
}

#define CS3D_INST_ADJ(real)                                                                       \
    template hipError_t launch_adj_pass<real>(int, const OpDev<real> *, const OpDev<real> *, int,  \
                                              int, const LaunchGeom &, const LineMap &,          \
                                              const AdjVjp<real> &, const Phys<real> &,          \
                                              hipStream_t);                                      \
    template hipError_t launch_adj_ke_seed<real>(const QSet<real> &, const Phys<real> &,           \
                                                 const QSet<real> &, long, hipStream_t);         \
    template hipError_t launch_adj_to_velocity<real>(const QSet<real> &, const QSet<real> &,       \
                                                     const Phys<real> &, real *, real *, real *, \
                                                     long, hipStream_t);                         \
    template hipError_t launch_gadj<real>(int, const GenAdj<real> &, const AdjVjp<real> &,         \
                                          const Phys<real> &, long, hipStream_t);                \
    template hipError_t launch_gadj_sponge_post<real>(                                            \
        const QSet<real> &, const QSet<real> &, const QSet<real> &, const QSet<real> &,          \
        const QSet<real> &, const StageArgs &, const Phys<real> &, double, double, long,         \
        hipStream_t);                                                                            \
    template hipError_t launch_gadj_add<real>(const QSet<real> &, const QSet<real> &, long,        \
                                              hipStream_t);                                      \
    template hipError_t launch_gadj_lincomb<real>(const QSet<real> &, real, const QSet<real> &,    \
                                                  real, const QSet<real> &, long, hipStream_t);
CS3D_INST_ADJ(float)
CS3D_INST_ADJ(double)

}  // namespace cs3d
