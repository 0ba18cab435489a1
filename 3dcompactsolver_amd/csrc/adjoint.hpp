// adjoint.hpp -- argument blocks and launchers of the adjoint kernels
// (adjoint.hip).  Field indices: first and second derivatives 3 * f + j with
// f = U, V, W, T and j = x, y, z; mixed derivatives Uxy, Uxz, Vxy, Vyz, Wxz,
// Wyz; flux cotangents 3 * c + j with c = rho, rhoU, rhoV, rhoW, rhoE.
#pragma once

#include "stage_common.hpp"

namespace cs3d {

template <typename real>
struct AdjPrim {
    real *U, *V, *W, *Tp, *mu, *amu;
};

// Inputs of the fused x seed kernel: the primitives, the stored y and z
// derivatives (U, V, W, T order), the two mixed derivatives that are not
// taken along x, and the stage cotangent v = dt * lam.
template <typename real>
struct AdjXSeed {
    const real *U, *V, *W, *Tp, *mu, *amu;
    const real *dy1[4], *dz1[4], *dy2[4], *dz2[4];
    const real *Vyz, *Wyz;
    const real *v[5];
};

// The cotangent seeds, deduplicated: every second-derivative and mixed
// seed is a multiple of s_i = mu w_i (w_i = v_i + u_i v_E) or of
// gq = cp/Pr mu v_E, and the first-derivative seeds of the symmetric pairs
// (Uy, Vx), (Uz, Wx), (Vz, Wy) coincide.
enum { G1_UX, G1_VY, G1_WZ, G1_XY, G1_XZ, G1_YZ, G1_TX, G1_TY, G1_TZ, G1_COUNT };

template <typename real>
struct AdjSeedOut {
    real *s[3];
    real *gq;
    real *g1[G1_COUNT];
    real *ub, *vb, *wb, *Tb;
};

constexpr int ADJ_COMB_MAX = 8;

template <typename real>
struct AdjComb {
    real *out;
    const real *base;  // may be null
    int n;
    real c[ADJ_COMB_MAX];
    const real *in[ADJ_COMB_MAX];
};

// Inputs of the fused x chain kernel and what is made of its result res,
// the conserved cotangent of the vector-Jacobian product: the x solves
// (D_x of cU, cV, m1/3, -gxz, -gTx and of v_c; D2_x of 4/3 s1, s2, s3, gq),
// the z and y parts of the primitive cotangents (tot, totY), the y and z flux
// cotangents D_y v_c, D_z v_c, and the state q.  Epilogue: out = res,
// acc = acc_base + res, vn = ca la + cb res (each skipped when its pointer
// is null; acc may alias acc_base; vn must not alias v).
template <typename real>
struct AdjFinal {
    const real *cU, *cV, *m1, *gxz, *gTx, *s[3], *gq, *v[5];
    const real *tot[4], *totY[4], *Fy[5], *Fz[5];
    QSet<real> q;
    real *out[5];
    real *acc[5];
    const real *acc_base[5];
    real *vn[5];
    const real *la[5];
    real ca, cb;
};

constexpr int ADJ_X_THREADS = 256;

// Batched compact line operators along one direction, N1 first-derivative
// terms then N2 second-derivative terms, term t acting on in[t] with weight
// c[t].  Separate form: out[t] = c[t] OP_t(in[t]).  Accumulated form:
// out[0] = sum_b bc[b] base[b] + sum_t c[t] OP_t(in[t]) (null bases skipped).
constexpr int LINE_TERMS_MAX = 8;
// most lanes per block of the batched line kernels (FP64 needs the larger
// register budget of half-size blocks)
template <typename real>
constexpr int line_terms_threads()
{
    return sizeof(real) == 4 ? 1024 : 512;
}

template <typename real>
struct LineTerms {
    const real *in[LINE_TERMS_MAX];
    real *out[LINE_TERMS_MAX];
    real c[LINE_TERMS_MAX];
    const real *base[2];
    real bc[2];
};

template <typename real>
hipError_t launch_adj_prim(const QSet<real> &q, const Phys<real> &ph, const AdjPrim<real> &o,
                           long N, hipStream_t st);
template <typename real>
hipError_t launch_adj_xseed(const OpDev<real> *d1, const OpDev<real> *d2, int uni, int M,
                            const LaunchGeom &g, const LineMap &lm, const AdjXSeed<real> &in,
                            const AdjSeedOut<real> &o, const Phys<real> &ph, hipStream_t st);
template <typename real>
hipError_t launch_adj_lincomb(const AdjComb<real> &a, long N, hipStream_t st);
template <typename real>
hipError_t launch_adj_xfinal(const OpDev<real> *d1, const OpDev<real> *d2, int uni, int M,
                             const LaunchGeom &g, const LineMap &lm, const AdjFinal<real> &f,
                             const Phys<real> &ph, hipStream_t st);
template <typename real>
hipError_t launch_adj_ke_seed(const QSet<real> &q, const Phys<real> &ph, const Ptr5<real> &out,
                              long N, hipStream_t st);
template <typename real>
hipError_t launch_line_terms(int n1, int n2, bool acc, const OpDev<real> *d1,
                             const OpDev<real> *d2, int uni, int M, const LaunchGeom &g, int map,
                             const LineMap &lm, const LineTerms<real> &a, hipStream_t st);
template <typename real>
hipError_t launch_adj_to_velocity(const QSet<real> &q, const Ptr5<real> &qb, const Phys<real> &ph,
                                  real *gU, real *gV, real *gW, long N, hipStream_t st);

}  // namespace cs3d
