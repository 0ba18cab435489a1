// plan.hpp -- constants of one compact operator (6th-order first or second
// derivative, or 8th-order filter) along one direction, factorised once on
// the host in FP64 and stored on the device in the working precision.
//
// The operator is out = A^-1 (B f) on every line of a field, A tridiagonal
// (Derivatives.hpp:123-161, Filter.hpp:83-104) and B the explicit stencil
// (Derivatives.cpp:119-247, Filter.cpp:5-96).  A line of n points is cut into
// P segments of at most M points; one GPU lane owns one segment of one line.
//
// Factorisation used for the solve (host side, see plan.cpp):
//  * periodic lines: A = alpha S^-1 + I + alpha S is circulant, and
//    A = (I + rho S^-1)(I + rho S) / (1 + rho^2) with
//    rho = (1 - sqrt(1 - 4 alpha^2)) / (2 alpha), |rho| < 1.  Each factor is a
//    cyclic first-order recurrence, stable in its own sweep direction.  This
//    replaces the Sherman-Morrison correction of Utils.cpp:22-46 (same
//    matrix, same solution; no second solve).
//  * non-periodic lines: the LU (Thomas) factorisation of Utils.cpp:4-20,
//    computed once instead of on every line.
// Either way the solve is two first-order linear recurrences
//     y_i = r_i - m_i y_{i-1}           (forward)
//     x_i = iw_i y_i + mu_i x_{i+1}     (backward)
// run segment-locally with zero carry, then corrected with the carry from
// the other segments: y_i += pf_i C_g, x_i += pb_i D_g, where C_g, D_g are
// short weighted sums over the other segments' end values (phi, psi).
#pragma once

#include <cstdint>
#include <vector>

namespace cs3d {

constexpr int MAX_SEG = 16;   // largest segment length compiled
constexpr int KMAX = 128;     // most cross-segment terms kept
constexpr int MAX_P = 256;    // most segments per line

enum OpKind { OP_FILTER = 0, OP_D1 = 1, OP_D2 = 2 };

template <typename real>
struct OpDev {
    int n;        // line length
    int M;        // segment capacity (compile-time M of the kernel)
    int P;        // number of segments
    int uniform;  // 1: every segment has M points and the per-point
                  //    constants depend on the local index only
    int cyclic;   // cross-segment recurrences wrap around (periodic)
    int kind;     // OpKind
    int H;        // interior stencil half width (2 or 4)
    int NB;       // boundary rows with their own stencil at each end
    int K;        // cross-segment terms kept (forward and backward)
    int Ks;       // table plans: row stride of phi and psi (>= K, zero past K)
    real scale;   // output scale (1 + rho^2 for periodic, 1 otherwise)
    real tap[5];  // interior stencil: D1 t1,t2 (odd), D2 t0,t1,t2, already
                  // divided by dh or dh^2; filter: tap[0] = (1 - 2 alpha_f)/256,
                  // the weight of the 8th difference in (A - B) f
    real bcl[4][7];  // row r < NB:        sum_u bcl[r][u] f[u]
    real bcr[4][7];  // row n - NB + r:    sum_u bcr[r][u] f[n - 7 + u]
                     // (for the filter these are rows of A - B: the filter
                     // output is f - A^-1 (A - B) f)
    // uniform path
    real um, umu, uiw;
    real upf[MAX_SEG];   // forward fix-up, local index t
    real upb[MAX_SEG];   // backward fix-up, local index t
    real uphi[KMAX];     // C_g = sum_k uphi[k-1] e_{g-k}
    real upsi[KMAX];     // D_g = sum_k upsi[k-1] b_{g+k}
    // the same carries summed over a whole line of 64 segments held by one
    // wavefront (uniform periodic plans): with lam the segment transfer
    // factor of the forward (0) or backward (1) recurrence,
    // uphi[k-1] = lam^(k-1) / (1 - lam^P); wlam[d][i] = lam^(2^i) and
    // wrow[d][k] = lam^k / (1 - lam^64)
    real wlam[2][4];
    real wrow[2][65];
    // table path (device pointers into the plan's own allocation)
    const real *m, *iw, *mu, *pf, *pb;  // per lane slot g * M + t (P * M each);
                                        // zero past a short segment's end
    const real *phi, *psi;              // [P][Ks]
    const int *sstart, *slen;           // [P]
};

// The transposed operator (the adjoint's): D^T g = B^T (A^-T g) and, for the
// filter, F^T g = g - (A - B)^T (A^-T g).  The solve with A^T runs on a plan
// built with its LHS transposed; B^T (or (A - B)^T) then acts on the solved
// line.  In the interior B is a Toeplitz stencil, so B^T is the same stencil
// (D1 antisymmetric: negated; D2 and the filter's difference symmetric).
// Near the ends of a non-periodic line the boundary rows of B make its
// columns irregular: columns j < JB (and j >= n - JB) are dense sums over
// the first (last) UB points, JB = max(7, NB + H) (the 7-point closure
// rows, and the interior rows reaching past them), UB = JB + H.
constexpr int TJB = 8;
constexpr int TUB = 12;

struct OpTHost {
    int n = 0, periodic = 0, kind = 0, JB = 0, UB = 0;
    double BTl[TJB][TUB] = {};  // column j < JB:     sum_u BTl[j][u] z[u]
    double BTr[TJB][TUB] = {};  // column n - JB + j: sum_u BTr[j][u] z[n - UB + u]
};

template <typename real>
struct OpTDev {
    int n, periodic, kind, JB, UB;
    real BTl[TJB][TUB];
    real BTr[TJB][TUB];
};

// Host-side FP64 description of a plan, before upload.
struct OpHost {
    int n = 0, M = 0, P = 0, uniform = 0, cyclic = 0, kind = 0, H = 0, NB = 0, K = 0;
    double scale = 1.0;
    double tap[5] = {0, 0, 0, 0, 0};
    double bcl[4][7] = {};
    double bcr[4][7] = {};
    double um = 0, umu = 0, uiw = 1;
    std::vector<double> m, iw, mu, pf, pb, phi, psi;
    std::vector<int> sstart, slen;
    // the tridiagonal matrix itself (for tests and the adjoint)
    std::vector<double> a, b, c;
    double corner = 0.0;  // periodic corner value (alpha)
};

// the boundary columns of B^T (or (A - B)^T) of a plan (plan.cpp); non-
// periodic lines need n >= 2 UB
OpTHost trans_table(const OpHost &h);

}  // namespace cs3d
