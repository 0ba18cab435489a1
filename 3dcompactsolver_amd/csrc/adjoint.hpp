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

template <typename real>
struct AdjSeedIn {
    const real *U, *V, *W, *Tp, *mu, *amu;
    const real *d1[12];
    const real *d2[12];
    const real *mx[6];
    const real *v[5];  // cotangent of the stage increment dt*R, times dt
};

template <typename real>
struct AdjSeedOut {
    real *g1[12];
    real *g2[12];
    real *gm[6];
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

template <typename real>
struct AdjFinal {
    const real *Fb[15];
    const real *ub[3];
    const real *Tb;
    real *out[5];
};

template <typename real>
hipError_t launch_adj_prim(const Ptr5<real> &q, const Phys<real> &ph, const AdjPrim<real> &o,
                           long N, hipStream_t st);
template <typename real>
hipError_t launch_adj_seed(const AdjSeedIn<real> &in, const AdjSeedOut<real> &o,
                           const Phys<real> &ph, long N, hipStream_t st);
template <typename real>
hipError_t launch_adj_lincomb(const AdjComb<real> &a, long N, hipStream_t st);
template <typename real>
hipError_t launch_adj_final(const Ptr5<real> &q, const AdjFinal<real> &f, const Phys<real> &ph,
                            long N, hipStream_t st);
template <typename real>
hipError_t launch_adj_ke_seed(const Ptr5<real> &q, const Phys<real> &ph, const Ptr5<real> &out,
                              long N, hipStream_t st);
template <typename real>
hipError_t launch_adj_to_velocity(const Ptr5<real> &q, const Ptr5<real> &qb, const Phys<real> &ph,
                                  real *gU, real *gV, real *gW, long N, hipStream_t st);

}  // namespace cs3d
