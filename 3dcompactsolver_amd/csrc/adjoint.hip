// adjoint.hip -- pointwise kernels of the discrete adjoint of the stage
// right-hand side (no reference counterpart: the reference has no adjoint).
//
// The forward right-hand side (CSolver_AWS.cpp:1246-1504, periodic, no
// sponge) is, with u = m/rho, p = (gamma-1)(E - |m|^2/(2 rho)),
// T = p/(rho R), mu = mu_ref T^0.76, a = dmu/dT = 0.76 mu/T:
//   R_rho = -sum_j D_j(m_j)
//   R_i   = -sum_j D_j(m_i u_j + p delta_ij) + Rv_i
//   R_E   = -sum_j D_j((E + p) u_j) + sum_i u_i Rv_i + mu QUAD
//           + cp/Pr (a |grad T|^2 + mu lap T)
// where Rv_i = mu (4/3 d_ii u_i + sum_{j!=i} d_jj u_i + 1/3 mixed terms)
//            + a (terms linear in grad T times velocity gradients)
// and QUAD the quadratic form of the velocity gradients.  Every
// derivative is a compact operator applied to a field; the adjoint applies
// the transposed operators (periodic D1 is skew-adjoint, D2 self-adjoint)
// to the cotangent seeds these kernels form, then chains back through the
// pointwise maps.  The cotangent v of the stage increment dt*R enters
// already multiplied by dt.
#include "adjoint.hpp"

namespace cs3d {

template <typename real>
__global__ void k_adj_prim(QSet<real> q, Phys<real> ph, AdjPrim<real> o, long N)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        // the forward's own primitive and viscosity evaluations
        real ir, U, V, W, pp, Tp, mu, amu, q5[5];
        qset_load(q, i, q5);
        prim(ph, q5[0], q5[1], q5[2], q5[3], q5[4], ir, U, V, W, pp, Tp);
        visc(ph, ph.T0s + Tp, mu, amu);
        o.U[i] = U;
        o.V[i] = V;
        o.W[i] = W;
        o.Tp[i] = Tp;
        o.mu[i] = mu;
        o.amu[i] = amu;
    }
}

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

// out = sum_k c[k] * in[k] (n terms)
template <typename real>
__global__ void k_adj_lincomb(AdjComb<real> a, long N)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        real s = a.base ? a.base[i] : real(0);
        for (int k = 0; k < a.n; k++) s += a.c[k] * a.in[k][i];
        a.out[i] = s;
    }
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

// N1 first-derivative and N2 second-derivative line solves of one direction
// through one exchange (see LineTerms).
template <int N1, int N2, bool ACC, typename real, int M, bool UNI, int MAP>
__global__ __launch_bounds__(line_terms_threads<real>()) void k_line_terms(
    const OpDev<real> *__restrict__ d1, const OpDev<real> *__restrict__ d2, LineMap lm,
    LineTerms<real> a)
{
    constexpr int HW = 2, W = M + 2 * HW, N = N1 + N2;
    CS3D_LANE_SETUP(MAP, d1, sweep_kr(d1, d2), N)
    (void)n;
    (void)per;
    real r[N][M];
#pragma unroll
    for (int t = 0; t < N; t++) {
        real w[W];
        win.load(a.in[t], w);
        if (t < N1)
            rhs<OP_D1, real, M, HW, UNI>(d1, c, w, r[t]);
        else
            rhs<OP_D2, real, M, HW, UNI>(d2, c, w, r[t]);
    }
    solve_d12<N1, N2, real, M, UNI>(d1, d2, c, r);
    if (ACC) {
        real o[M];
#pragma unroll
        for (int t = 0; t < M; t++) o[t] = 0;
#pragma unroll
        for (int b = 0; b < 2; b++)
            if (a.base[b]) {
                real bv[M];
                win.load_seg(a.base[b], bv);
#pragma unroll
                for (int t = 0; t < M; t++) o[t] += a.bc[b] * bv[t];
            }
#pragma unroll
        for (int k = 0; k < N; k++)
#pragma unroll
            for (int t = 0; t < M; t++) o[t] += a.c[k] * r[k][t];
        if (active) win.store(a.out[0], o, c.len);
    } else {
#pragma unroll
        for (int k = 0; k < N; k++) {
#pragma unroll
            for (int t = 0; t < M; t++) r[k][t] *= a.c[k];
            if (active) win.store(a.out[k], r[k], c.len);
        }
    }
}

template <int N1, int N2, bool ACC, typename real, bool UNI>
static hipError_t launch_lt(const OpDev<real> *d1, const OpDev<real> *d2, const LaunchGeom &g,
                            int map, const LineMap &lm, const LineTerms<real> &a, hipStream_t st)
{
    constexpr int M = SEG_DERIV;
    dim3 grid((lm.nlines + g.LX - 1) / g.LX), block(g.threads);
    if (map == MAP_ROW)
        hipLaunchKernelGGL((k_line_terms<N1, N2, ACC, real, M, UNI, MAP_ROW>), grid, block, g.shm,
                           st, d1, d2, lm, a);
    else
        hipLaunchKernelGGL((k_line_terms<N1, N2, ACC, real, M, UNI, MAP_COL>), grid, block, g.shm,
                           st, d1, d2, lm, a);
    return hipGetLastError();
}

template <int N1, int N2, bool ACC, typename real>
static hipError_t launch_lt2(const OpDev<real> *d1, const OpDev<real> *d2, int uni,
                             const LaunchGeom &g, int map, const LineMap &lm,
                             const LineTerms<real> &a, hipStream_t st)
{
    return uni ? launch_lt<N1, N2, ACC, real, true>(d1, d2, g, map, lm, a, st)
               : launch_lt<N1, N2, ACC, real, false>(d1, d2, g, map, lm, a, st);
}

template <typename real>
hipError_t launch_line_terms(int n1, int n2, bool acc, const OpDev<real> *d1,
                             const OpDev<real> *d2, int uni, int M, const LaunchGeom &g, int map,
                             const LineMap &lm, const LineTerms<real> &a, hipStream_t st)
{
    if (M != SEG_DERIV) return hipErrorInvalidValue;
    if (g.LX < 1 || g.threads < 1 || g.threads > line_terms_threads<real>())
        return hipErrorInvalidConfiguration;
#define CS3D_LT(A, B, C)                                                                   \
    if (n1 == A && n2 == B && acc == C) return launch_lt2<A, B, C, real>(d1, d2, uni, g, map, lm, a, st);
    CS3D_LT(4, 4, false)
    CS3D_LT(2, 0, false)
    CS3D_LT(4, 0, false)
    CS3D_LT(3, 0, false)
    CS3D_LT(5, 0, false)
    CS3D_LT(1, 1, true)
    CS3D_LT(2, 1, true)
    CS3D_LT(1, 0, true)
#undef CS3D_LT
    return hipErrorInvalidValue;
}

// Gradient of J = (1/N) sum rho |u|^2 / 2 with respect to the conserved
// state, and the map of a conserved cotangent to the initial primitive
// velocities at fixed rho and p: m_i = rho u_i, E = p/(gamma-1) + rho|u|^2/2.
// The x-direction derivatives of the stage (D1 of U, V, W, T and of Uy, Uz,
// Vy, Wz for the mixed derivatives, D2 of U, V, W, T) through one batched
// solve, then the cotangent seeds at every point from them and the stored
// y and z derivatives.  The mixed derivatives are taken as D_x D_y U etc.:
// periodic operators along different axes commute.
template <typename real, int M, bool UNI>
__global__ __launch_bounds__(ADJ_X_THREADS) void k_adj_xseed(const OpDev<real> *__restrict__ d1,
                                                             const OpDev<real> *__restrict__ d2,
                                                             LineMap lm, AdjXSeed<real> in,
                                                             AdjSeedOut<real> o, Phys<real> ph)
{
    constexpr int HW = 2, W = M + 2 * HW;
    CS3D_LANE_SETUP(MAP_ROW, d1, sweep_kr(d1, d2), 12)
    (void)n;
    (void)per;
    // r: Ux Vx Wx Tx | Uxy Uxz Vxy Wxz | Uxx Vxx Wxx Txx; own: the
    // segment's own values of the eight window fields
    real r[12][M], own[8][M];
    {
        const real *src[8] = {in.U, in.V, in.W, in.Tp, in.dy1[0], in.dz1[0], in.dy1[1], in.dz1[2]};
#pragma unroll
        for (int f = 0; f < 8; f++) {
            real w[W];
            win.load(src[f], w);
#pragma unroll
            for (int t = 0; t < M; t++) own[f][t] = w[HW + t];
            rhs<OP_D1, real, M, HW, UNI>(d1, c, w, r[f]);
            if (f < 4) rhs<OP_D2, real, M, HW, UNI>(d2, c, w, r[8 + f]);
        }
    }
    solve_d12<8, 4, real, M, UNI>(d1, d2, c, r);
    // the remaining per-point inputs, whole segments at a time
    enum { P_MU, P_A, P_V1, P_V2, P_V3, P_VE, P_WY, P_TY, P_VZ, P_TZ, P_VYZ, P_WYZ, P_D2 };
    real pt[P_D2 + 8][M];
    {
        const real *src[P_D2 + 8] = {in.mu, in.amu, in.v[1], in.v[2], in.v[3], in.v[4],
                                     in.dy1[2], in.dy1[3], in.dz1[1], in.dz1[3], in.Vyz, in.Wyz,
                                     in.dy2[0], in.dz2[0], in.dy2[1], in.dz2[1],
                                     in.dy2[2], in.dz2[2], in.dy2[3], in.dz2[3]};
#pragma unroll
        for (int k = 0; k < P_D2 + 8; k++) win.load_seg(src[k], pt[k]);
    }
    real out[SD_COUNT][M];
#pragma unroll
    for (int t = 0; t < M; t++) {
        SeedPt<real> p;
        p.U = own[0][t];
        p.V = own[1][t];
        p.W = own[2][t];
        p.Tp = own[3][t];
        p.mu = pt[P_MU][t];
        p.a = pt[P_A][t];
        p.v1 = pt[P_V1][t];
        p.v2 = pt[P_V2][t];
        p.v3 = pt[P_V3][t];
        p.vE = pt[P_VE][t];
#pragma unroll
        for (int f = 0; f < 4; f++) {
            p.d1[3 * f] = r[f][t];
            p.d2[3 * f] = r[8 + f][t];
            p.d2[3 * f + 1] = pt[P_D2 + 2 * f][t];
            p.d2[3 * f + 2] = pt[P_D2 + 2 * f + 1][t];
        }
        p.d1[1] = own[4][t];  // Uy
        p.d1[2] = own[5][t];  // Uz
        p.d1[4] = own[6][t];  // Vy
        p.d1[5] = pt[P_VZ][t];
        p.d1[7] = pt[P_WY][t];
        p.d1[8] = own[7][t];  // Wz
        p.d1[10] = pt[P_TY][t];
        p.d1[11] = pt[P_TZ][t];
        p.mx[0] = r[4][t];  // Uxy
        p.mx[1] = r[5][t];  // Uxz
        p.mx[2] = r[6][t];  // Vxy
        p.mx[3] = pt[P_VYZ][t];
        p.mx[4] = r[7][t];  // Wxz
        p.mx[5] = pt[P_WYZ][t];
        real sd[SD_COUNT];
        seed_point(ph, p, sd);
#pragma unroll
        for (int k = 0; k < SD_COUNT; k++) out[k][t] = sd[k];
    }
    if (!active) return;
    real *dst[SD_COUNT];
#pragma unroll
    for (int k = 0; k < 3; k++) dst[SD_S + k] = o.s[k];
    dst[SD_GQ] = o.gq;
#pragma unroll
    for (int k = 0; k < G1_COUNT; k++) dst[SD_G1 + k] = o.g1[k];
    dst[SD_UB] = o.ub;
    dst[SD_VB] = o.vb;
    dst[SD_WB] = o.wb;
    dst[SD_TB] = o.Tb;
#pragma unroll
    for (int k = 0; k < SD_COUNT; k++) win.store(dst[k], out[k], c.len);
}

// The x-direction transposed terms of the primitive cotangents and the x
// flux cotangents through one batched solve, then the chain back to the
// conserved state at every point, with its epilogue (AdjFinal).
template <typename real, int M, bool UNI>
__global__ __launch_bounds__(ADJ_X_THREADS) void k_adj_xfinal(const OpDev<real> *__restrict__ d1,
                                                              const OpDev<real> *__restrict__ d2,
                                                              LineMap lm, AdjFinal<real> f,
                                                              Phys<real> ph)
{
    constexpr int HW = 2, W = M + 2 * HW;
    CS3D_LANE_SETUP(MAP_ROW, d1, sweep_kr(d1, d2), 14)
    (void)n;
    (void)per;
    // r: D_x of cU cV m1 gxz gTx v0..v4 | D2_x of s1 s2 s3 gq
    real r[14][M];
    {
        const real *in1[10] = {f.cU, f.cV, f.m1, f.gxz, f.gTx, f.v[0], f.v[1], f.v[2], f.v[3], f.v[4]};
#pragma unroll
        for (int k = 0; k < 10; k++) {
            real w[W];
            win.load(in1[k], w);
            rhs<OP_D1, real, M, HW, UNI>(d1, c, w, r[k]);
        }
        const real *in2[4] = {f.s[0], f.s[1], f.s[2], f.gq};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            real w[W];
            win.load(in2[k], w);
            rhs<OP_D2, real, M, HW, UNI>(d2, c, w, r[10 + k]);
        }
    }
    solve_d12<10, 4, real, M, UNI>(d1, d2, c, r);
    const real c13 = real(1) / real(3), c43 = real(4) / real(3);
    // the per-point inputs, whole segments at a time: tot + totY, D_y v,
    // D_z v, q
    real pt[19][M];
    {
        const real *src[14] = {f.tot[0], f.tot[1], f.tot[2], f.tot[3], f.Fy[0], f.Fy[1], f.Fy[2],
                               f.Fy[3], f.Fy[4], f.Fz[0], f.Fz[1], f.Fz[2], f.Fz[3], f.Fz[4]};
#pragma unroll
        for (int k = 0; k < 14; k++) win.load_seg(src[k], pt[k]);
        win.load_seg_quad(f.q.m, pt[14], pt[15], pt[16], pt[17]);
        win.load_seg(f.q.e, pt[18]);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            real ty[M];
            win.load_seg(f.totY[k], ty);
#pragma unroll
            for (int t = 0; t < M; t++) pt[k][t] += ty[t];
        }
    }
    real res[5][M];
#pragma unroll
    for (int t = 0; t < M; t++) {
        real q[5], Fb[15], ub[3];
#pragma unroll
        for (int k = 0; k < 5; k++) {
            q[k] = pt[14 + k][t];
            Fb[3 * k] = r[5 + k][t];
            Fb[3 * k + 1] = pt[4 + k][t];
            Fb[3 * k + 2] = pt[9 + k][t];
        }
        ub[0] = pt[0][t] + r[0][t] + c43 * r[10][t];
        ub[1] = pt[1][t] + r[1][t] + r[11][t];
        ub[2] = pt[2][t] + c13 * r[2][t] - r[3][t] + r[12][t];
        const real Tb = pt[3][t] - r[4][t] + r[13][t];
        real rp[5];
        chain_point(ph, q, Fb, ub, Tb, rp);
#pragma unroll
        for (int k = 0; k < 5; k++) res[k][t] = rp[k];
    }
    if (!active) return;
#pragma unroll
    for (int k = 0; k < 5; k++) {
        if (f.out[k]) win.store(f.out[k], res[k], c.len);
        if (f.acc[k]) {
            real b[M];
            if (f.acc_base[k])
                win.load_seg(f.acc_base[k], b);
            else
#pragma unroll
                for (int t = 0; t < M; t++) b[t] = 0;
#pragma unroll
            for (int t = 0; t < M; t++) b[t] += res[k][t];
            win.store(f.acc[k], b, c.len);
        }
        if (f.vn[k]) {
            real l[M];
            win.load_seg(f.la[k], l);
#pragma unroll
            for (int t = 0; t < M; t++) l[t] = f.ca * l[t] + f.cb * res[k][t];
            win.store(f.vn[k], l, c.len);
        }
    }
}

template <typename real, bool UNI>
static hipError_t launch_x(int which, const OpDev<real> *d1, const OpDev<real> *d2,
                           const LaunchGeom &g, const LineMap &lm, const AdjXSeed<real> *in,
                           const AdjSeedOut<real> *so, const AdjFinal<real> *fin,
                           const Phys<real> &ph, hipStream_t st)
{
    constexpr int M = SEG_DERIV;
    dim3 grid((lm.nlines + g.LX - 1) / g.LX), block(g.threads);
    if (which == 0)
        hipLaunchKernelGGL((k_adj_xseed<real, M, UNI>), grid, block, g.shm, st, d1, d2, lm, *in,
                           *so, ph);
    else
        hipLaunchKernelGGL((k_adj_xfinal<real, M, UNI>), grid, block, g.shm, st, d1, d2, lm, *fin,
                           ph);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_adj_xseed(const OpDev<real> *d1, const OpDev<real> *d2, int uni, int M,
                            const LaunchGeom &g, const LineMap &lm, const AdjXSeed<real> &in,
                            const AdjSeedOut<real> &o, const Phys<real> &ph, hipStream_t st)
{
    if (M != SEG_DERIV) return hipErrorInvalidValue;
    if (g.LX < 1 || g.threads < 1 || g.threads > ADJ_X_THREADS) return hipErrorInvalidConfiguration;
    return uni ? launch_x<real, true>(0, d1, d2, g, lm, &in, &o, nullptr, ph, st)
               : launch_x<real, false>(0, d1, d2, g, lm, &in, &o, nullptr, ph, st);
}

template <typename real>
hipError_t launch_adj_xfinal(const OpDev<real> *d1, const OpDev<real> *d2, int uni, int M,
                             const LaunchGeom &g, const LineMap &lm, const AdjFinal<real> &f,
                             const Phys<real> &ph, hipStream_t st)
{
    if (M != SEG_DERIV) return hipErrorInvalidValue;
    if (g.LX < 1 || g.threads < 1 || g.threads > ADJ_X_THREADS) return hipErrorInvalidConfiguration;
    return uni ? launch_x<real, true>(1, d1, d2, g, lm, nullptr, nullptr, &f, ph, st)
               : launch_x<real, false>(1, d1, d2, g, lm, nullptr, nullptr, &f, ph, st);
}

template <typename real>
__global__ void k_adj_ke_seed(QSet<real> q, Phys<real> ph, Ptr5<real> out, long N)
{
    const real iN = real(1) / real(N);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        real q5[5];
        qset_load(q, i, q5);
        const real rho = ph.rho_ref + q5[0];
        const real u = q5[1] / rho, v = q5[2] / rho, w = q5[3] / rho;
        out.out[0][i] = -real(0.5) * (u * u + v * v + w * w) * iN;
        out.out[1][i] = u * iN;
        out.out[2][i] = v * iN;
        out.out[3][i] = w * iN;
        out.out[4][i] = 0;
    }
}

template <typename real>
__global__ void k_adj_to_velocity(QSet<real> q, Ptr5<real> qb, Phys<real> ph, real *gU, real *gV,
                                  real *gW, long N)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        real q5[5];
        qset_load(q, i, q5);
        const real rho = ph.rho_ref + q5[0];
        const real Eb = qb.in[4][i];
        gU[i] = rho * (qb.in[1][i] + (q5[1] / rho) * Eb);
        gV[i] = rho * (qb.in[2][i] + (q5[2] / rho) * Eb);
        gW[i] = rho * (qb.in[3][i] + (q5[3] / rho) * Eb);
    }
}

static dim3 grid_for(long N)
{
    long nb = (N + 255) / 256;
    if (nb > 8192) nb = 8192;
    return dim3((unsigned)nb);
}

template <typename real>
hipError_t launch_adj_prim(const QSet<real> &q, const Phys<real> &ph, const AdjPrim<real> &o,
                           long N, hipStream_t st)
{
    hipLaunchKernelGGL((k_adj_prim<real>), grid_for(N), dim3(256), 0, st, q, ph, o, N);
    return hipGetLastError();
}
template <typename real>
hipError_t launch_adj_lincomb(const AdjComb<real> &a, long N, hipStream_t st)
{
    hipLaunchKernelGGL((k_adj_lincomb<real>), grid_for(N), dim3(256), 0, st, a, N);
    return hipGetLastError();
}
template <typename real>
hipError_t launch_adj_ke_seed(const QSet<real> &q, const Phys<real> &ph, const Ptr5<real> &out,
                              long N, hipStream_t st)
{
    hipLaunchKernelGGL((k_adj_ke_seed<real>), grid_for(N), dim3(256), 0, st, q, ph, out, N);
    return hipGetLastError();
}
template <typename real>
hipError_t launch_adj_to_velocity(const QSet<real> &q, const Ptr5<real> &qb, const Phys<real> &ph,
                                  real *gU, real *gV, real *gW, long N, hipStream_t st)
{
    hipLaunchKernelGGL((k_adj_to_velocity<real>), grid_for(N), dim3(256), 0, st, q, qb, ph, gU, gV,
                       gW, N);
    return hipGetLastError();
}

#define CS3D_INST_ADJ(real)                                                                       \
    template hipError_t launch_adj_prim<real>(const QSet<real> &, const Phys<real> &,              \
                                              const AdjPrim<real> &, long, hipStream_t);         \
    template hipError_t launch_adj_xseed<real>(const OpDev<real> *, const OpDev<real> *, int, int, \
                                               const LaunchGeom &, const LineMap &,              \
                                               const AdjXSeed<real> &, const AdjSeedOut<real> &, \
                                               const Phys<real> &, hipStream_t);                 \
    template hipError_t launch_adj_lincomb<real>(const AdjComb<real> &, long, hipStream_t);        \
    template hipError_t launch_line_terms<real>(int, int, bool, const OpDev<real> *,               \
                                                const OpDev<real> *, int, int, const LaunchGeom &, \
                                                int, const LineMap &, const LineTerms<real> &,     \
                                                hipStream_t);                                      \
    template hipError_t launch_adj_xfinal<real>(const OpDev<real> *, const OpDev<real> *, int,     \
                                                int, const LaunchGeom &, const LineMap &,        \
                                                const AdjFinal<real> &, const Phys<real> &,      \
                                                hipStream_t);                                    \
    template hipError_t launch_adj_ke_seed<real>(const QSet<real> &, const Phys<real> &,           \
                                                 const Ptr5<real> &, long, hipStream_t);         \
    template hipError_t launch_adj_to_velocity<real>(const QSet<real> &, const Ptr5<real> &,       \
                                                     const Phys<real> &, real *, real *, real *, \
                                                     long, hipStream_t);
CS3D_INST_ADJ(float)
CS3D_INST_ADJ(double)

}  // namespace cs3d
