// adjoint.hpp -- argument blocks and launchers of the adjoint kernels
// (adjoint.hip).  One vector-Jacobian product of the stage right-hand side
// is five line passes: the y and z derivative fields of the stage state
// recomputed (y, z), the x derivatives, the cotangent seeds and every
// x-direction transposed term (x), then the z and the y transposed terms,
// each pass chaining its share of the primitive and flux cotangents back to
// the conserved state (the chain is linear in them) and adding it to the
// previous pass's share; the y pass also forms the RK epilogue.  Cotangents
// of the conserved state are stored like the state: a quad (rho, rhoU, rhoV,
// rhoW) per point and the energy on its own (QSet).
#pragma once

#include "stage_common.hpp"

namespace cs3d {

// The derivative fields of the stage state the x pass reads: from the y
// pass (Uy, Ty), (Vy, Wy) pairs and the quad (Uyy, Vyy, Wyy, Tyy); from the
// z pass the quads (Uz, Vz, Wz, Tz), (Uzz, Vzz, Wzz, Tzz) and the pair
// (Vyz, Wyz) = D_z of (Vy, Wy).
template <typename real>
struct AdjRec {
    real *dYut, *dYvw, *dY2;
    real *dZ1, *dZ2, *mYZ;
};

// The cotangent seeds, deduplicated: every second-derivative and mixed
// seed is a multiple of s_i = mu w_i (w_i = v_i + u_i v_E) or of
// gq = cp/Pr mu v_E, and the first-derivative seeds of the symmetric pairs
// (Uy, Vx), (Uz, Wx), (Vz, Wy) coincide.
enum { G1_UX, G1_VY, G1_WZ, G1_XY, G1_XZ, G1_YZ, G1_TX, G1_TY, G1_TZ, G1_COUNT };

// What the x pass leaves for the column passes (Xs_i = D_x s_i, the x half
// of the mixed-derivative transposes): the quads sq = (s0, s1, s2, gq),
// gz = (g1XZ, g1WZ, g1TZ, Xs2), gy = (g1XY, g1VY, g1TY, Xs1), the pair
// gs = (g1YZ, Xs0); and the pair mzz = (D_z s1, D_z s2) the z pass leaves
// for the y pass.
template <typename real>
struct AdjSeeds {
    real *sq, *gz, *gy, *gs, *mzz;
};

// One product res = (d(dt R)/dQ)^T lam at the stage state q, with the stage
// cotangent dt * lam = cv * v given (the passes scale v as they read it);
// the passes add their shares through resx
// (x) and resxz (x + z).  Epilogue of the y pass, each part skipped when
// its set is null (m == nullptr): out = res, acc = acc_base + res (acc may
// alias acc_base; a null acc_base counts as zero), vn = ca la + cb res (vn
// must not alias v).
template <typename real>
struct AdjVjp {
    QSet<real> q, v;
    AdjRec<real> rec;
    AdjSeeds<real> sd;
    QSet<real> resx, resxz;
    QSet<real> out, acc, acc_base, vn, la;
    real cv, ca, cb;
};

// most lanes per workgroup: the x pass takes whole lines as the x sweep
// does; the transposed column passes hold 17 solves per lane, which 512
// lanes (FP32) and 256 (FP64) give the registers for; the recompute passes
// hold 8 and 10 (1024 lanes on uniform FP32 lines, else 512)
constexpr int ADJ_X_THREADS = 256;
template <typename real>
constexpr int adj_col_threads()
{
    return sizeof(real) == 4 ? 512 : 256;
}
template <typename real>
constexpr int adj_rec_threads(bool uni)
{
    return sizeof(real) == 4 && uni ? 1024 : 512;
}
// ring planes (values per exchange position) and LDS staging of each pass
constexpr int ADJ_X_PLANES = 16, ADJ_X_STAGE = 8;
constexpr int ADJ_T_PLANES = 9;   // transposed column passes: batches of 9 and 8
constexpr int ADJ_R_PLANES = 10;  // recompute passes: 8 (y) and 10 (z) solves

enum { ADJ_YREC, ADJ_ZREC, ADJ_X, ADJ_Z, ADJ_Y };

// values per point the x pass keeps in LDS across its second solve batch
constexpr int ADJ_X_KEEP = 15;
// LDS bytes of an x-pass workgroup of LX lines of P segments of M points:
// the exchange ring or the window staging, whichever is larger, then the
// kept values
inline size_t adj_x_lds(int P, int kr, int LX, int M, int esize)
{
    const size_t ring = ring_bytes(P, kr, LX, ADJ_X_PLANES, esize, false);
    const size_t stage = (size_t)ADJ_X_STAGE * M * LX * P * esize;
    return (ring > stage ? ring : stage) + (size_t)ADJ_X_KEEP * M * LX * P * esize;
}

// pass: ADJ_*; g from line_geom with the planes and staging above
template <typename real>
hipError_t launch_adj_pass(int pass, const OpDev<real> *d1, const OpDev<real> *d2, int uni, int M,
                           const LaunchGeom &g, const LineMap &lm, const AdjVjp<real> &a,
                           const Phys<real> &ph, hipStream_t st);
// ---- general path (any boundary set without walls: non-periodic lines,
// sponge faces).  One product: the primitives U, V, W, T' of the stage
// state (pY) and of the x block's set (pX: the stage state, or at the first
// stage after a sponge update the pre-reset values on the sponge planes;
// pX may alias pY), their derivative fields with the standalone operators
// (d1[3 f + j], d2[3 f + j] for f = U V W T and j = x y z, x derivatives of
// the pX set; mx = Uxy Uxz Vxy Vyz Wxz Wyz), the seeds (s: s0 s1 s2 gq; g1
// as G1_*), the masked and scaled stage cotangent vs, the primitive
// cotangents of both sets (bY, bX: U V W T) and the flux cotangents F[d][c]
// after the transposed operators, then the chain (res_stale: the x block's
// share on the sponge planes at a stale stage; resA += the cotangent of the
// sponge averages).
enum { GADJ_PRIMS, GADJ_SEEDS, GADJ_CHAIN };
template <typename real>
struct GenAdj {
    QSet<real> q, qstale, v;
    StageArgs sa;
    int stale;
    real cv;
    real *pY[4], *pX[4];
    real *d1[12], *d2[12], *mx[6];
    real *s[4], *g1[G1_COUNT], *vs[5];
    // the continuity cotangent of each direction's flux term: vs[0], except
    // on wall planes, which keep only the normal term of their last wall
    // face (vc[d] alias vs[0] without walls)
    real *vc[3];
    real *bY[4], *bX[4];
    real *F[3][5];
    real *sigma;            // null without sponges
    QSet<real> res_stale, resA;
};
template <typename real>
hipError_t launch_gadj(int which, const GenAdj<real> &g, const AdjVjp<real> &a,
                       const Phys<real> &ph, long N, hipStream_t st);
// reverse of the sponge update of one step (adjoint.hip)
template <typename real>
hipError_t launch_gadj_sponge_post(const QSet<real> &qf, const QSet<real> &A,
                                   const QSet<real> &lq, const QSet<real> &lA,
                                   const QSet<real> &ls, const StageArgs &sa,
                                   const Phys<real> &ph, double eps, double epsP, long N,
                                   hipStream_t st);
template <typename real>
hipError_t launch_gadj_add(const QSet<real> &b, const QSet<real> &a, long N, hipStream_t st);
// out = ca a + cb b (b may be null: cb ignored; out may alias a or b)
template <typename real>
hipError_t launch_gadj_lincomb(const QSet<real> &out, real ca, const QSet<real> &a, real cb,
                               const QSet<real> &b, long N, hipStream_t st);
// Reverse of one wall face's correction (walls.hip: k_wall_face): lam, the
// cotangent of the state after the face, becomes that of the state before
// it, st (in place); the face's momenta and energy were overwritten, its
// density kept, and its energy depends on the six inward temperatures.  At
// the first stage after a sponge update the pair: stS the pre-reset copy
// before the face and lamS its cotangent (both empty otherwise).
template <typename real>
hipError_t launch_wall_face_vjp(const QSet<real> &st, const QSet<real> &lam, const QSet<real> &stS,
                                const QSet<real> &lamS, const Phys<real> &ph, const StageArgs &sa,
                                const WallFace &w, hipStream_t s);

template <typename real>
hipError_t launch_adj_ke_seed(const QSet<real> &q, const Phys<real> &ph, const QSet<real> &out,
                              long N, hipStream_t st);
template <typename real>
hipError_t launch_adj_to_velocity(const QSet<real> &q, const QSet<real> &qb, const Phys<real> &ph,
                                  real *gU, real *gV, real *gW, long N, hipStream_t st);

}  // namespace cs3d
