// plan.cpp -- host FP64 construction of compact-operator plans (see plan.hpp).
#include "plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace cs3d {

namespace {

// Lele 6th-order compact first derivative, interior and boundary closures
// (Derivatives.hpp:79-97).
struct D1c {
    static constexpr double alpha = 1.0 / 3.0, a = 14.0 / 9.0, b = 1.0 / 9.0;
    static constexpr double al11 = 5.0, a1 = -197.0 / 60.0, b1 = -5.0 / 12.0,
                            c1 = 5.0, d1 = -5.0 / 3.0, e1 = 5.0 / 12.0,
                            f1 = -1.0 / 20.0;
    static constexpr double al21 = 1.0 / 8.0, al22 = 3.0 / 4.0,
                            a2 = -43.0 / 96.0, b2 = -5.0 / 6.0, c2 = 9.0 / 8.0,
                            d2 = 1.0 / 6.0, e2 = -1.0 / 96.0;
};
// second derivative (Derivatives.hpp:100-120)
struct D2c {
    static constexpr double alpha = 2.0 / 11.0, a = 12.0 / 11.0, b = 3.0 / 11.0;
    static constexpr double al11 = 10.0, a1 = 145.0 / 12.0, b1 = -76.0 / 3.0,
                            c1 = 29.0 / 2.0, d1 = -4.0 / 3.0, e1 = 1.0 / 12.0;
    static constexpr double al21 = 2.0 / 11.0, al22 = -131.0 / 22.0,
                            a2 = 177.0 / 88.0, b2 = -507.0 / 44.0,
                            c2 = 783.0 / 44.0, d2 = -201.0 / 22.0,
                            e2 = 81.0 / 88.0, f2 = -3.0 / 44.0;
};

void fail(const std::string &msg) { throw std::invalid_argument(msg); }

}  // namespace

// Segmentation.  Uniform: every segment has M points (periodic lines with
// n % M == 0).  Otherwise every segment is full except one short segment,
// placed in the middle (P >= 3) or first (P == 2), so that the first and the
// last segments always hold the 7-point boundary closures inside their
// windows.
static void segment(int n, int M, OpHost &h, bool want_uniform)
{
    h.P = (n + M - 1) / M;
    h.uniform = (want_uniform && n % M == 0) ? 1 : 0;
    if (h.P > MAX_P) fail("line too long for the compiled segment length");
    h.sstart.resize(h.P);
    h.slen.resize(h.P);
    int shortg = h.P >= 3 ? h.P / 2 : 0;
    int rem = n - (h.P - 1) * M;  // length of the short segment (1..M)
    int s = 0;
    for (int g = 0; g < h.P; g++) {
        int len = (g == shortg) ? rem : M;
        h.sstart[g] = s;
        h.slen[g] = len;
        s += len;
    }
}

OpHost build_plan(int kind, int periodic, int n, double dh, double alpha_f,
                  int M, double tol, bool allow_uniform, bool transpose)
{
    OpHost h;
    h.n = n;
    h.M = M;
    h.kind = kind;
    h.cyclic = periodic ? 1 : 0;
    if (M < 1 || M > MAX_SEG) fail("segment length out of range");

    // ---- minimum line lengths (the reference's stencils index up to +-4
    // in the periodic filter and use 6/7-point closures otherwise)
    if (kind == OP_FILTER) {
        if (n < 8) fail("filter needs at least 8 points per line (Filter.cpp:5-96)");
        if (!(alpha_f > -0.5 && alpha_f < 0.5)) fail("alpha_f must lie in (-0.5, 0.5)");
    } else {
        if (periodic && n < 5) fail("periodic derivative needs at least 5 points");
        if (!periodic && n < 8) fail("non-periodic derivative needs at least 8 points");
    }

    // ---- LHS (Derivatives.hpp:133-161, Filter.hpp:83-104)
    double al = kind == OP_D1 ? D1c::alpha : kind == OP_D2 ? D2c::alpha : alpha_f;
    h.a.assign(n, al);
    h.b.assign(n, 1.0);
    h.c.assign(n, al);
    h.corner = periodic ? al : 0.0;
    if (!periodic) {
        if (kind == OP_D1 || kind == OP_D2) {
            double a11 = kind == OP_D1 ? D1c::al11 : D2c::al11;
            double a21 = kind == OP_D1 ? D1c::al21 : D2c::al21;
            double a22 = kind == OP_D1 ? D1c::al22 : D2c::al22;
            h.c[0] = a11;
            h.c[1] = a22;
            h.a[1] = a21;
            h.a[n - 1] = a11;
            h.a[n - 2] = a22;
            h.c[n - 2] = a21;
        } else {
            for (int r = 0; r < 3; r++) {
                h.a[r] = h.c[r] = 0.0;
                h.a[n - 1 - r] = h.c[n - 1 - r] = 0.0;
            }
        }
    }
    h.a[0] = periodic ? h.a[0] : 0.0;          // no sub-diagonal in row 0
    h.c[n - 1] = periodic ? h.c[n - 1] : 0.0;  // no super-diagonal in row n-1
    if (transpose && !periodic) {
        // the solve of the transposed operator runs on A^T: sub-diagonal
        // a'_i = c_{i-1}, super-diagonal c'_i = a_{i+1} (the periodic A is a
        // symmetric circulant: A^T = A)
        std::vector<double> at(n, 0.0), ct(n, 0.0);
        for (int i = 1; i < n; i++) at[i] = h.c[i - 1];
        for (int i = 0; i + 1 < n; i++) ct[i] = h.a[i + 1];
        h.a = at;
        h.c = ct;
    }

    // ---- explicit stencil B (Derivatives.cpp:119-247, Filter.cpp:5-96)
    if (kind == OP_D1) {
        h.H = 2;
        h.tap[0] = 0.0;
        h.tap[1] = (D1c::a / 2.0) / dh;
        h.tap[2] = (D1c::b / 4.0) / dh;
    } else if (kind == OP_D2) {
        h.H = 2;
        double d2 = dh * dh;
        h.tap[0] = -(D2c::b / 2.0 + 2.0 * D2c::a) / d2;
        h.tap[1] = D2c::a / d2;
        h.tap[2] = (D2c::b / 4.0) / d2;
    } else {
        h.H = 4;
        double af = alpha_f;
        h.tap[0] = (93.0 + 70.0 * af) / 128.0;
        h.tap[1] = ((7.0 + 18.0 * af) / 16.0) / 2.0;
        h.tap[2] = ((-7.0 + 14.0 * af) / 32.0) / 2.0;
        h.tap[3] = ((1.0 - 2.0 * af) / 16.0) / 2.0;
        h.tap[4] = ((-1.0 + 2.0 * af) / 128.0) / 2.0;
    }
    if (!periodic) {
        if (kind == OP_D1) {
            h.NB = 2;
            const double r0[7] = {D1c::a1, D1c::b1, D1c::c1, D1c::d1, D1c::e1, D1c::f1, 0};
            const double r1[7] = {D1c::a2, D1c::b2, D1c::c2, D1c::d2, D1c::e2, 0, 0};
            const double rn2[7] = {0, 0, -D1c::e2, -D1c::d2, -D1c::c2, -D1c::b2, -D1c::a2};
            const double rn1[7] = {0, -D1c::f1, -D1c::e1, -D1c::d1, -D1c::c1, -D1c::b1, -D1c::a1};
            for (int u = 0; u < 7; u++) {
                h.bcl[0][u] = r0[u] / dh;
                h.bcl[1][u] = r1[u] / dh;
                h.bcr[0][u] = rn2[u] / dh;
                h.bcr[1][u] = rn1[u] / dh;
            }
        } else if (kind == OP_D2) {
            h.NB = 2;
            double d2 = dh * dh;
            const double r0[7] = {D2c::a1, D2c::b1, D2c::c1, D2c::d1, D2c::e1, 0, 0};
            const double r1[7] = {D2c::a2, D2c::b2, D2c::c2, D2c::d2, D2c::e2, D2c::f2, 0};
            const double rn2[7] = {0, D2c::f2, D2c::e2, D2c::d2, D2c::c2, D2c::b2, D2c::a2};
            const double rn1[7] = {0, 0, D2c::e1, D2c::d1, D2c::c1, D2c::b1, D2c::a1};
            for (int u = 0; u < 7; u++) {
                h.bcl[0][u] = r0[u] / d2;
                h.bcl[1][u] = r1[u] / d2;
                h.bcr[0][u] = rn2[u] / d2;
                h.bcr[1][u] = rn1[u] / d2;
            }
        } else {
            // Filter.hpp:60-81: explicit one-sided rows 0-2, 6th-order row 3
            h.NB = 4;
            double af = alpha_f;
            double a0 = (11.0 + 10.0 * af) / 16.0, a1 = ((15.0 + 34.0 * af) / 32.0) / 2.0,
                   a2 = ((-3.0 + 6.0 * af) / 16.0) / 2.0, a3 = ((1.0 - 2.0 * af) / 32.0) / 2.0;
            const double A0[5] = {15.0 / 16.0, 4.0 / 16.0, -6.0 / 16.0, 4.0 / 16.0, -1.0 / 16.0};
            const double B0[5] = {1.0 / 16.0, 3.0 / 4.0, 6.0 / 16.0, -4.0 / 16.0, 1.0 / 16.0};
            const double C0[5] = {-1.0 / 16.0, 4.0 / 16.0, 5.0 / 8.0, 4.0 / 16.0, -1.0 / 16.0};
            const double six[7] = {a3, a2, a1, a0, a1, a2, a3};
            for (int u = 0; u < 7; u++) {
                h.bcl[0][u] = u < 5 ? A0[u] : 0.0;
                h.bcl[1][u] = u < 5 ? B0[u] : 0.0;
                h.bcl[2][u] = u < 5 ? C0[u] : 0.0;
                h.bcl[3][u] = six[u];
                // right end: row n-1-r uses f[n-1-v] with the row-r weights
                h.bcr[0][u] = six[u];
                h.bcr[1][u] = u >= 2 ? C0[6 - u] : 0.0;
                h.bcr[2][u] = u >= 2 ? B0[6 - u] : 0.0;
                h.bcr[3][u] = u >= 2 ? A0[6 - u] : 0.0;
            }
        }
    }

    // ---- the filter in difference form.  The kernels apply the filter as
    // out = f - A^-1 (A - B) f.  In the interior (A - B) is exactly
    // kappa * D8 with kappa = (1 - 2 alpha_f)/256 and D8 the 8th central
    // difference (1, -8, 28, -56, 70, -56, 28, -8, 1): the two operators
    // have the symbol (1 - 2 alpha_f)(1 - cos theta)^4 / 16.  Evaluated as
    // nested second differences this never rounds the low-wavenumber gain
    // (a direct B f in FP32 moves it by ~1e-8 per application, thousands of
    // times the filter's own effect on resolved scales).  Boundary rows keep
    // explicit (A - B) rows.
    if (kind == OP_FILTER) {
        const double al = alpha_f;
        h.tap[0] = (1.0 - 2.0 * al) / 256.0;
        for (int i = 1; i < 5; i++) h.tap[i] = 0.0;
        if (!periodic) {
            // A rows inside the 7-point closure windows: identity on rows
            // 0-2 and n-3..n-1, (alpha, 1, alpha) on rows 3 and n-4
            const double Al[4][7] = {{1, 0, 0, 0, 0, 0, 0}, {0, 1, 0, 0, 0, 0, 0},
                                     {0, 0, 1, 0, 0, 0, 0}, {0, 0, al, 1, al, 0, 0}};
            const double Ar[4][7] = {{0, 0, al, 1, al, 0, 0}, {0, 0, 0, 0, 1, 0, 0},
                                     {0, 0, 0, 0, 0, 1, 0}, {0, 0, 0, 0, 0, 0, 1}};
            for (int r = 0; r < 4; r++)
                for (int u = 0; u < 7; u++) {
                    h.bcl[r][u] = Al[r][u] - h.bcl[r][u];
                    h.bcr[r][u] = Ar[r][u] - h.bcr[r][u];
                }
        }
    }

    // ---- factorisation into the two recurrences
    h.m.assign(n, 0.0);
    h.iw.assign(n, 1.0);
    h.mu.assign(n, 0.0);
    double rho = 0.0;
    if (periodic) {
        // A = (I + rho S^-1)(I + rho S)/(1 + rho^2); rho = 2 al / (1 + sqrt(1 - 4 al^2))
        rho = 2.0 * al / (1.0 + std::sqrt(1.0 - 4.0 * al * al));
        h.scale = 1.0 + rho * rho;
        for (int i = 0; i < n; i++) {
            h.m[i] = rho;
            h.iw[i] = 1.0;
            h.mu[i] = -rho;
        }
        h.um = rho;
        h.umu = -rho;
        h.uiw = 1.0;
    } else {
        // Thomas LU once per plan (Utils.cpp:4-20 does it on every line)
        double w = h.b[0];
        h.iw[0] = 1.0 / w;
        h.m[0] = 0.0;
        for (int i = 1; i < n; i++) {
            double mi = h.a[i] / w;
            w = h.b[i] - mi * h.c[i - 1];
            if (w == 0.0) fail("singular compact matrix");
            h.m[i] = mi;
            h.iw[i] = 1.0 / w;
        }
        for (int i = 0; i < n - 1; i++) h.mu[i] = -h.c[i] * h.iw[i];
        h.mu[n - 1] = 0.0;
        h.scale = 1.0;
    }

    segment(n, M, h, allow_uniform && periodic != 0);
    int P = h.P;
    if (!periodic) {
        if (h.slen[0] < h.NB || h.slen[P - 1] != M)
            fail("line too short for the non-periodic closures at this segment length");
    }

    // ---- per-point fix-up products
    h.pf.assign(n, 0.0);
    h.pb.assign(n, 0.0);
    std::vector<double> Pf(P), Pb(P);
    for (int g = 0; g < P; g++) {
        int s = h.sstart[g], len = h.slen[g];
        double p = 1.0;
        for (int i = s; i < s + len; i++) {
            p *= -h.m[i];
            h.pf[i] = p;
        }
        Pf[g] = p;
        p = 1.0;
        for (int i = s + len - 1; i >= s; i--) {
            p *= h.mu[i];
            h.pb[i] = p;
        }
        Pb[g] = p;
    }

    // ---- cross-segment coefficients
    std::vector<double> phi((size_t)P * P, 0.0), psi((size_t)P * P, 0.0);
    if (periodic) {
        double allf = 1.0, allb = 1.0;
        for (int g = 0; g < P; g++) {
            allf *= Pf[g];
            allb *= Pb[g];
        }
        for (int g = 0; g < P; g++) {
            double p = 1.0;
            for (int k = 1; k <= P; k++) {
                // C_g = sum_k e_{g-k} prod_{l=g-k+1}^{g-1} Pf_l / (1 - allf)
                phi[(size_t)g * P + k - 1] = p / (1.0 - allf);
                p *= Pf[((g - k) % P + P) % P];
            }
            p = 1.0;
            for (int k = 1; k <= P; k++) {
                psi[(size_t)g * P + k - 1] = p / (1.0 - allb);
                p *= Pb[(g + k) % P];
            }
        }
    } else {
        for (int g = 0; g < P; g++) {
            double p = 1.0;
            for (int k = 1; k <= g; k++) {
                phi[(size_t)g * P + k - 1] = p;
                p *= Pf[g - k];
            }
            p = 1.0;
            for (int k = 1; g + k < P; k++) {
                psi[(size_t)g * P + k - 1] = p;
                p *= Pb[g + k];
            }
        }
    }
    int K = 0;
    for (int g = 0; g < P; g++)
        for (int k = 1; k <= P; k++) {
            if (std::fabs(phi[(size_t)g * P + k - 1]) > tol ||
                std::fabs(psi[(size_t)g * P + k - 1]) > tol)
                K = std::max(K, k);
        }
    if (K < 1) K = 1;
    if (K > KMAX) fail("too many cross-segment terms");
    h.K = K;
    h.phi.assign((size_t)P * K, 0.0);
    h.psi.assign((size_t)P * K, 0.0);
    for (int g = 0; g < P; g++)
        for (int k = 0; k < K; k++) {
            h.phi[(size_t)g * K + k] = phi[(size_t)g * P + k];
            h.psi[(size_t)g * K + k] = psi[(size_t)g * P + k];
        }
    return h;
}

// Entry (u, j) of the explicit matrix the kernels apply before the solve:
// B for the derivatives, A - B for the filter (rows as plan.hpp describes:
// closure rows bcl/bcr at the ends of non-periodic lines, the interior
// stencil elsewhere).
static double b_entry(const OpHost &h, int u, int j)
{
    const int n = h.n;
    if (u < h.NB) return j < 7 ? h.bcl[u][j] : 0.0;
    if (u >= n - h.NB) {
        const int jj = j - (n - 7);
        return jj >= 0 ? h.bcr[u - (n - h.NB)][jj] : 0.0;
    }
    const int e = j - u;
    if (e < -h.H || e > h.H) return 0.0;
    if (h.kind == OP_D1) return e > 0 ? h.tap[e] : e < 0 ? -h.tap[-e] : 0.0;
    if (h.kind == OP_D2) return h.tap[e < 0 ? -e : e];
    // kappa * D8: the 8th central difference (1, -8, 28, -56, 70, ...)
    static const double d8[9] = {1, -8, 28, -56, 70, -56, 28, -8, 1};
    return h.tap[0] * d8[e + 4];
}

OpTHost trans_table(const OpHost &h)
{
    OpTHost t;
    t.n = h.n;
    t.periodic = h.cyclic;
    t.kind = h.kind;
    if (h.cyclic) return t;  // no boundary columns
    t.JB = std::max(7, h.NB + h.H);
    t.UB = t.JB + h.H;
    if (t.JB > TJB || t.UB > TUB) fail("transposed boundary block too large");
    if (h.n < 2 * t.UB)
        fail("transposed operator needs at least " + std::to_string(2 * t.UB) +
             " points on a non-periodic line");
    for (int j = 0; j < t.JB; j++)
        for (int u = 0; u < t.UB; u++) {
            t.BTl[j][u] = b_entry(h, u, j);
            t.BTr[j][u] = b_entry(h, h.n - t.UB + u, h.n - t.JB + j);
        }
    return t;
}

}  // namespace cs3d
