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
__global__ void k_adj_prim(Ptr5<real> q, Phys<real> ph, AdjPrim<real> o, long N)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        // the forward's own primitive and viscosity evaluations
        real ir, U, V, W, pp, Tp, mu, amu;
        prim(ph, q.in[0][i], q.in[1][i], q.in[2][i], q.in[3][i], q.in[4][i], ir, U, V, W, pp, Tp);
        visc(ph, ph.T0s + Tp, mu, amu);
        o.U[i] = U;
        o.V[i] = V;
        o.W[i] = W;
        o.Tp[i] = Tp;
        o.mu[i] = mu;
        o.amu[i] = amu;
    }
}

// Cotangent seeds of the derivative fields and the pointwise cotangents of
// U, V, W, T from the viscous and heat-flux terms.  Index of the input first
// and second derivative fields: 3 * f + j with f = U, V, W, T and
// j = x, y, z; mixed: Uxy, Uxz, Vxy, Vyz, Wxz, Wyz.
template <typename real>
__global__ void k_adj_seed(AdjSeedIn<real> in, AdjSeedOut<real> o, Phys<real> ph, long N)
{
    const real c43 = real(4) / real(3), c23 = real(2) / real(3), c13 = real(1) / real(3);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        const real U = in.U[i], V = in.V[i], W = in.W[i];
        const real T = ph.T0s + in.Tp[i], mu = in.mu[i], a = in.amu[i];
        const real v1 = in.v[1][i], v2 = in.v[2][i], v3 = in.v[3][i], vE = in.v[4][i];
        const real w1 = v1 + U * vE, w2 = v2 + V * vE, w3 = v3 + W * vE;
        const real Ux = in.d1[0][i], Uy = in.d1[1][i], Uz = in.d1[2][i];
        const real Vx = in.d1[3][i], Vy = in.d1[4][i], Vz = in.d1[5][i];
        const real Wx = in.d1[6][i], Wy = in.d1[7][i], Wz = in.d1[8][i];
        const real Tx = in.d1[9][i], Ty = in.d1[10][i], Tz = in.d1[11][i];
        const real Uxx = in.d2[0][i], Uyy = in.d2[1][i], Uzz = in.d2[2][i];
        const real Vxx = in.d2[3][i], Vyy = in.d2[4][i], Vzz = in.d2[5][i];
        const real Wxx = in.d2[6][i], Wyy = in.d2[7][i], Wzz = in.d2[8][i];
        const real Txx = in.d2[9][i], Tyy = in.d2[10][i], Tzz = in.d2[11][i];
        const real Uxy = in.mx[0][i], Uxz = in.mx[1][i], Vxy = in.mx[2][i];
        const real Vyz = in.mx[3][i], Wxz = in.mx[4][i], Wyz = in.mx[5][i];

        // second and mixed derivatives: multiples of s_i = mu w_i and of
        // gq = cp/Pr mu vE
        o.s[0][i] = mu * w1;
        o.s[1][i] = mu * w2;
        o.s[2][i] = mu * w3;
        o.gq[i] = ph.cpPr * mu * vE;

        // first derivatives: the a-terms of the momenta (weights w), the
        // quadratic form (weight mu vE) and the heat-flux square (cp/Pr a vE)
        const real P1 = c43 * (Ux - real(0.5) * Vy - real(0.5) * Wz);
        const real P2 = c43 * (Vy - real(0.5) * Ux - real(0.5) * Wz);
        const real P3 = c43 * (Wz - real(0.5) * Ux - real(0.5) * Vy);
        const real Sxy = Uy + Vx, Sxz = Wx + Uz, Syz = Wy + Vz;
        const real mq = mu * vE;
        o.g1[G1_UX][i] = a * (c43 * Tx * w1 - c23 * Ty * w2 - c23 * Tz * w3) +
                         mq * (real(8) / real(3) * Ux - c43 * (Vy + Wz));
        o.g1[G1_VY][i] = a * (-c23 * Tx * w1 + c43 * Ty * w2 - c23 * Tz * w3) +
                         mq * (real(8) / real(3) * Vy - c43 * (Ux + Wz));
        o.g1[G1_WZ][i] = a * (-c23 * Tx * w1 - c23 * Ty * w2 + c43 * Tz * w3) +
                         mq * (real(8) / real(3) * Wz - c43 * (Ux + Vy));
        o.g1[G1_XY][i] = a * (Ty * w1 + Tx * w2) + real(2) * mq * Sxy;  // Uy and Vx
        o.g1[G1_XZ][i] = a * (Tz * w1 + Tx * w3) + real(2) * mq * Sxz;  // Uz and Wx
        o.g1[G1_YZ][i] = a * (Tz * w2 + Ty * w3) + real(2) * mq * Syz;  // Vz and Wy
        const real ht = real(2) * ph.cpPr * a * vE;
        o.g1[G1_TX][i] = a * (P1 * w1 + Sxy * w2 + Sxz * w3) + ht * Tx;
        o.g1[G1_TY][i] = a * (Sxy * w1 + P2 * w2 + Syz * w3) + ht * Ty;
        o.g1[G1_TZ][i] = a * (Sxz * w1 + Syz * w2 + P3 * w3) + ht * Tz;

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
        o.Tb[i] = mub * a + ab * (real(-0.24) * a / T);
        // the energy equation's sum_i u_i Rv_i
        o.ub[i] = vE * (mu * L1 + a * A1);
        o.vb[i] = vE * (mu * L2 + a * A2);
        o.wb[i] = vE * (mu * L3 + a * A3);
    }
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
// Fb[3 * c + j] = D_j v_c back to the conserved state.
template <typename real>
__global__ void k_adj_final(Ptr5<real> q, AdjFinal<real> f, Phys<real> ph, long N)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        const real rho = ph.rho_ref + q.in[0][i];
        const real ir = real(1) / rho;
        const real m[3] = {q.in[1][i], q.in[2][i], q.in[3][i]};
        const real u[3] = {m[0] * ir, m[1] * ir, m[2] * ir};
        const real E = ph.rhoE_ref + q.in[4][i];
        const real pp = ph.gm1 * (q.in[4][i] - real(0.5) * (m[0] * u[0] + m[1] * u[1] + m[2] * u[2]));
        const real p = ph.p_ref + pp;
        const real T = ph.T0s + (pp - q.in[0][i] * ph.RT0s) * ir * ph.iR;
        real mb[3], ub[3] = {f.ub[0][i], f.ub[1][i], f.ub[2][i]};
        real pb = 0, Eb = 0, rb = 0;
        for (int k = 0; k < 3; k++) mb[k] = f.Fb[k][i];  // continuity fluxes m_j
        for (int ii = 0; ii < 3; ii++) {                  // momentum fluxes m_i u_j + p d_ij
            for (int j = 0; j < 3; j++) {
                const real Fb = f.Fb[3 * (ii + 1) + j][i];
                mb[ii] += Fb * u[j];
                ub[j] += Fb * m[ii];
                if (ii == j) pb += Fb;
            }
        }
        for (int j = 0; j < 3; j++) {  // energy fluxes (E + p) u_j
            const real Fb = f.Fb[12 + j][i];
            Eb += Fb * u[j];
            pb += Fb * u[j];
            ub[j] += Fb * (E + p);
        }
        // T = p / (rho R)
        const real Tb = f.Tb[i];
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
        const real res[5] = {rb, mb[0], mb[1], mb[2], Eb};
#pragma unroll
        for (int c = 0; c < 5; c++) {
            if (f.out[c]) f.out[c][i] = res[c];
            if (f.acc[c]) f.acc[c][i] = (f.acc_base[c] ? f.acc_base[c][i] : real(0)) + res[c];
            if (f.vn[c]) f.vn[c][i] = f.ca * f.la[c][i] + f.cb * res[c];
        }
    }
}

// N1 first-derivative and N2 second-derivative line solves of one direction
// through one exchange (see LineTerms).
template <int N1, int N2, bool ACC, typename real, int M, bool UNI, int MAP>
__global__ __launch_bounds__(line_terms_threads<real>()) void k_line_terms(
    const OpDev<real> *__restrict__ d1, const OpDev<real> *__restrict__ d2, LineMap lm,
    LineTerms<real> a)
{
    constexpr int HW = 2, W = M + 2 * HW, N = N1 + N2;
    constexpr int KC = kc_deriv<real>();
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
    solve_batch2<N1, N2, KC, real, M, UNI>(d1, d2, c, r);
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
template <typename real>
__global__ void k_adj_ke_seed(Ptr5<real> q, Phys<real> ph, Ptr5<real> out, long N)
{
    const real iN = real(1) / real(N);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        const real rho = ph.rho_ref + q.in[0][i];
        const real u = q.in[1][i] / rho, v = q.in[2][i] / rho, w = q.in[3][i] / rho;
        out.out[0][i] = -real(0.5) * (u * u + v * v + w * w) * iN;
        out.out[1][i] = u * iN;
        out.out[2][i] = v * iN;
        out.out[3][i] = w * iN;
        out.out[4][i] = 0;
    }
}

template <typename real>
__global__ void k_adj_to_velocity(Ptr5<real> q, Ptr5<real> qb, Phys<real> ph, real *gU, real *gV,
                                  real *gW, long N)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        const real rho = ph.rho_ref + q.in[0][i];
        const real Eb = qb.in[4][i];
        gU[i] = rho * (qb.in[1][i] + (q.in[1][i] / rho) * Eb);
        gV[i] = rho * (qb.in[2][i] + (q.in[2][i] / rho) * Eb);
        gW[i] = rho * (qb.in[3][i] + (q.in[3][i] / rho) * Eb);
    }
}

static dim3 grid_for(long N)
{
    long nb = (N + 255) / 256;
    if (nb > 8192) nb = 8192;
    return dim3((unsigned)nb);
}

template <typename real>
hipError_t launch_adj_prim(const Ptr5<real> &q, const Phys<real> &ph, const AdjPrim<real> &o,
                           long N, hipStream_t st)
{
    hipLaunchKernelGGL((k_adj_prim<real>), grid_for(N), dim3(256), 0, st, q, ph, o, N);
    return hipGetLastError();
}
template <typename real>
hipError_t launch_adj_seed(const AdjSeedIn<real> &in, const AdjSeedOut<real> &o,
                           const Phys<real> &ph, long N, hipStream_t st)
{
    hipLaunchKernelGGL((k_adj_seed<real>), grid_for(N), dim3(256), 0, st, in, o, ph, N);
    return hipGetLastError();
}
template <typename real>
hipError_t launch_adj_lincomb(const AdjComb<real> &a, long N, hipStream_t st)
{
    hipLaunchKernelGGL((k_adj_lincomb<real>), grid_for(N), dim3(256), 0, st, a, N);
    return hipGetLastError();
}
template <typename real>
hipError_t launch_adj_final(const Ptr5<real> &q, const AdjFinal<real> &f, const Phys<real> &ph,
                            long N, hipStream_t st)
{
    hipLaunchKernelGGL((k_adj_final<real>), grid_for(N), dim3(256), 0, st, q, f, ph, N);
    return hipGetLastError();
}
template <typename real>
hipError_t launch_adj_ke_seed(const Ptr5<real> &q, const Phys<real> &ph, const Ptr5<real> &out,
                              long N, hipStream_t st)
{
    hipLaunchKernelGGL((k_adj_ke_seed<real>), grid_for(N), dim3(256), 0, st, q, ph, out, N);
    return hipGetLastError();
}
template <typename real>
hipError_t launch_adj_to_velocity(const Ptr5<real> &q, const Ptr5<real> &qb, const Phys<real> &ph,
                                  real *gU, real *gV, real *gW, long N, hipStream_t st)
{
    hipLaunchKernelGGL((k_adj_to_velocity<real>), grid_for(N), dim3(256), 0, st, q, qb, ph, gU, gV,
                       gW, N);
    return hipGetLastError();
}

#define CS3D_INST_ADJ(real)                                                                       \
    template hipError_t launch_adj_prim<real>(const Ptr5<real> &, const Phys<real> &,              \
                                              const AdjPrim<real> &, long, hipStream_t);         \
    template hipError_t launch_adj_seed<real>(const AdjSeedIn<real> &, const AdjSeedOut<real> &,   \
                                              const Phys<real> &, long, hipStream_t);            \
    template hipError_t launch_adj_lincomb<real>(const AdjComb<real> &, long, hipStream_t);        \
    template hipError_t launch_line_terms<real>(int, int, bool, const OpDev<real> *,               \
                                                const OpDev<real> *, int, int, const LaunchGeom &, \
                                                int, const LineMap &, const LineTerms<real> &,     \
                                                hipStream_t);                                      \
    template hipError_t launch_adj_final<real>(const Ptr5<real> &, const AdjFinal<real> &,         \
                                               const Phys<real> &, long, hipStream_t);           \
    template hipError_t launch_adj_ke_seed<real>(const Ptr5<real> &, const Phys<real> &,           \
                                                 const Ptr5<real> &, long, hipStream_t);         \
    template hipError_t launch_adj_to_velocity<real>(const Ptr5<real> &, const Ptr5<real> &,       \
                                                     const Phys<real> &, real *, real *, real *, \
                                                     long, hipStream_t);
CS3D_INST_ADJ(float)
CS3D_INST_ADJ(double)

}  // namespace cs3d
