// ops.hip -- plan upload and the standalone compact operator kernel
// (one field in, one field out, any direction, any layout).  This is the
// GPU body of Derivatives::calc1stDerivField / calc2ndDerivField
// (Derivatives.cpp:3-96) and Filter::filterField (Filter.cpp:129-171).
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "plan_dev.hpp"

namespace cs3d {

std::string hip_msg(hipError_t e, const char *what)
{
    return std::string(what) + ": " + hipGetErrorString(e);
}

static void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(hip_msg(e, what));
}

template <typename real>
void PlanDev<real>::release()
{
    if (dev) (void)hipFree(dev);
    if (tables) (void)hipFree(tables);
    if (tdev) (void)hipFree(tdev);
    dev = nullptr;
    tables = nullptr;
    tdev = nullptr;
}

OpHost plan_for(int kind, int periodic, int n, double dh, double alpha_f, bool fp32, int seg,
                bool transpose)
{
    // cross-segment terms below this weight are dropped: a sixth (FP32) and
    // a tenth (FP64) of the unit roundoff, relative to the carried value
    const double tol = fp32 ? 1e-8 : 1e-17;
    // uniform kernels only for periodic lines that split into equal
    // segments of the uniform length; everything else takes the table kernels
    const int MU = seg > 0 ? seg : kind == OP_FILTER ? SEG_FILTER : SEG_DERIV;
    if (periodic && n % MU == 0)
        return build_plan(kind, periodic, n, dh, alpha_f, MU, tol, true, transpose);
    return build_plan(kind, periodic, n, dh, alpha_f, SEG_TABLE, tol, false, transpose);
}

template <typename real>
void make_plan(PlanDev<real> &pl, int kind, int periodic, int n, double dh, double alpha_f, int seg,
               bool transpose)
{
    pl.release();
    OpHost h = plan_for(kind, periodic, n, dh, alpha_f, sizeof(real) == 4, seg, transpose);
    pl.host = h;
    OpDev<real> d;
    std::memset(&d, 0, sizeof(d));
    d.n = h.n;
    d.M = h.M;
    d.P = h.P;
    d.uniform = h.uniform;
    d.cyclic = h.cyclic;
    d.kind = h.kind;
    d.H = h.H;
    d.NB = h.NB;
    d.K = h.K;
    d.scale = (real)h.scale;
    for (int i = 0; i < 5; i++) d.tap[i] = (real)h.tap[i];
    for (int r = 0; r < 4; r++)
        for (int u = 0; u < 7; u++) {
            d.bcl[r][u] = (real)h.bcl[r][u];
            d.bcr[r][u] = (real)h.bcr[r][u];
        }
    d.um = (real)h.um;
    d.umu = (real)h.umu;
    d.uiw = (real)h.uiw;
    if (h.uniform) {
        for (int t = 0; t < h.M; t++) {
            d.upf[t] = (real)h.pf[t];
            d.upb[t] = (real)h.pb[t];
        }
        for (int k = 0; k < h.K; k++) {
            d.uphi[k] = (real)h.phi[k];
            d.upsi[k] = (real)h.psi[k];
        }
        if (h.cyclic) {
            // segment transfer factors: the products of -m and of mu over
            // one segment (plan.cpp's Pf, Pb)
            const double lam[2] = {h.pf[h.M - 1], h.pb[0]};
            for (int s = 0; s < 2; s++) {
                double p = lam[s];
                for (int i = 0; i < 4; i++) {
                    d.wlam[s][i] = (real)p;
                    p *= p;
                }
                double l64 = 1.0;
                for (int k = 0; k < 64; k++) l64 *= lam[s];
                double pk = 1.0;
                for (int k = 0; k <= 64; k++) {
                    d.wrow[s][k] = (real)(pk / (1.0 - l64));
                    pk *= lam[s];
                }
            }
        }
    }
    // tables, laid out per lane slot so that a lane reads its segment's
    // M constants at g * M + t whatever its length (zeros past the end of a
    // short segment: the kernels need no bounds test): m iw mu pf pb (P * M
    // each); phi psi (P rows of Ks >= K, zero past K, so that the kernels'
    // unrolled carry terms read zero weights); sstart slen (P)
    const int kcu = kind == OP_FILTER ? kc_filter<real>() : kc_deriv<real>();
    const int Ks = std::max(h.K, kcu);
    d.Ks = Ks;
    const size_t slots = (size_t)h.P * h.M;
    size_t nreal = 5 * slots + 2 * (size_t)h.P * Ks;
    size_t bytes = nreal * sizeof(real) + 2 * (size_t)h.P * sizeof(int);
    std::vector<char> buf(bytes, 0);
    real *rp = (real *)buf.data();
    auto put_slots = [&](const std::vector<double> &v) {
        for (int g = 0; g < h.P; g++)
            for (int t = 0; t < h.M; t++)
                rp[(size_t)g * h.M + t] = t < h.slen[g] ? (real)v[h.sstart[g] + t] : real(0);
        rp += slots;
    };
    auto put_rows = [&](const std::vector<double> &v) {
        for (int g = 0; g < h.P; g++)
            for (int k = 0; k < h.K; k++) rp[(size_t)g * Ks + k] = (real)v[(size_t)g * h.K + k];
        rp += (size_t)h.P * Ks;
    };
    put_slots(h.m);
    put_slots(h.iw);
    put_slots(h.mu);
    put_slots(h.pf);
    put_slots(h.pb);
    put_rows(h.phi);
    put_rows(h.psi);
    int *ip = (int *)rp;
    for (int g = 0; g < h.P; g++) ip[g] = h.sstart[g];
    for (int g = 0; g < h.P; g++) ip[h.P + g] = h.slen[g];
    hip_check(hipMalloc(&pl.tables, bytes), "hipMalloc(plan tables)");
    hip_check(hipMemcpy(pl.tables, buf.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy(plan)");
    real *dp = (real *)pl.tables;
    d.m = dp;
    d.iw = dp + slots;
    d.mu = dp + 2 * slots;
    d.pf = dp + 3 * slots;
    d.pb = dp + 4 * slots;
    d.phi = dp + 5 * slots;
    d.psi = d.phi + (size_t)h.P * Ks;
    d.sstart = (const int *)(d.psi + (size_t)h.P * Ks);
    d.slen = d.sstart + h.P;
    pl.hd = d;
    hip_check(hipMalloc(&pl.dev, sizeof(OpDev<real>)), "hipMalloc(plan)");
    hip_check(hipMemcpy(pl.dev, &d, sizeof(d), hipMemcpyHostToDevice), "hipMemcpy(plan)");
    // one workgroup: up to 1024 lanes, at most 64 lines
    const int kc = kind == OP_FILTER ? kc_filter<real>() : kc_deriv<real>();
    // (either lane layout: the ring of the larger)
    LaunchGeom g = line_geom(h.P, std::max(h.K, kc), 1024, 1, sizeof(real), MAP_ROW);
    if (g.LX >= 1) {
        const LaunchGeom gc = line_geom(h.P, std::max(h.K, kc), 1024, 1, sizeof(real), MAP_COL);
        g.shm = std::max(g.shm, gc.shm);
    }
    if (g.LX < 1) throw std::invalid_argument("too many segments per line");
    if (transpose) {
        // the solved lines (n + 1 values each) go to LDS after the ring
        auto bytes = [&](int LX) {
            return std::max(ring_bytes(h.P, std::max(h.K, kc), LX, 1, sizeof(real), false),
                            ring_bytes(h.P, std::max(h.K, kc), LX, 1, sizeof(real), true)) +
                   (size_t)LX * (h.n + 1) * sizeof(real);
        };
        while (g.LX > 0 && bytes(g.LX) > LDS_BYTES) g.LX--;
        if (g.LX < 1) throw std::invalid_argument("line too long for the transposed operator");
        g.threads = g.LX * h.P;
        g.shm = bytes(g.LX);
        pl.thost = trans_table(h);
        OpTDev<real> t;
        std::memset(&t, 0, sizeof(t));
        t.n = pl.thost.n;
        t.periodic = pl.thost.periodic;
        t.kind = pl.thost.kind;
        t.JB = pl.thost.JB;
        t.UB = pl.thost.UB;
        for (int j = 0; j < TJB; j++)
            for (int u = 0; u < TUB; u++) {
                t.BTl[j][u] = (real)pl.thost.BTl[j][u];
                t.BTr[j][u] = (real)pl.thost.BTr[j][u];
            }
        hip_check(hipMalloc(&pl.tdev, sizeof(t)), "hipMalloc(transposed plan)");
        hip_check(hipMemcpy(pl.tdev, &t, sizeof(t), hipMemcpyHostToDevice), "hipMemcpy(plan)");
    }
    pl.LX = g.LX;
    pl.threads = g.threads;
    pl.shm = g.shm;
}

LineMap xyz_lines(int dir, int Nx, int Ny, int Nz)
{
    LineMap m;
    long sxy = (long)Nx * Ny;
    if (dir == 0) {  // lines (j, k): L = j + Ny k
        m.n1 = Ny; m.s1 = Nx; m.s2 = sxy; m.sp = 1; m.nlines = Ny * Nz;
    } else if (dir == 1) {  // lines (i, k): L = i + Nx k
        m.n1 = Nx; m.s1 = 1; m.s2 = sxy; m.sp = Nx; m.nlines = Nx * Nz;
    } else {  // lines (i, j): L = i + Nx j
        m.n1 = Nx; m.s1 = 1; m.s2 = Nx; m.sp = sxy; m.nlines = Nx * Ny;
    }
    return m;
}

LineMap ref_lines(int dir, int Nx, int Ny, int Nz)
{
    if (dir == 0) return xyz_lines(0, Nx, Ny, Nz);
    LineMap m;
    if (dir == 1) {  // YZX: i*Nz*Ny + k*Ny + j; L = k + Nz i
        m.n1 = Nz; m.s1 = Ny; m.s2 = (long)Nz * Ny; m.sp = 1; m.nlines = Nx * Nz;
    } else {  // ZXY: j*Nx*Nz + i*Nz + k; L = i + Nx j
        m.n1 = Nx; m.s1 = Nz; m.s2 = (long)Nx * Nz; m.sp = 1; m.nlines = Nx * Ny;
    }
    return m;
}

template <int KIND, typename real, int M, bool UNI, int MAP>
__global__ __launch_bounds__(1024) void k_field_op(const OpDev<real> *__restrict__ op,
                                                   LineMap lm, const real *__restrict__ in,
                                                   real *__restrict__ out)
{
    extern __shared__ char smem_raw[];
    real *xb = (real *)smem_raw;
    constexpr int HW = KIND == OP_FILTER ? 4 : 2, W = M + 2 * HW;
    constexpr int KC = KIND == OP_FILTER ? kc_filter<real>()
                       : KIND == OP_D2   ? kc_d2<real>()
                                         : kc_deriv<real>();
    const int P = op->P;
    const int LX = blockDim.x / P;
    Lane<real, MAP> c = make_lane<real, MAP>(op, P, LX, op->K > KC ? op->K : KC, 1, xb);
    const int L = blockIdx.x * LX + c.line;
    const bool active = L < lm.nlines;
    Win<real, M, HW, vec_windows<real, M, MAP, UNI>()> win;
    win.init((unsigned)lm.base(active ? L : lm.nlines - 1), (unsigned)lm.sp, c.s, op->n,
             op->cyclic);
    real w[W], r[1][M];
    win.load(in, w);
    rhs<KIND, real, M, HW, UNI>(op, c, w, r[0]);
    solve_batch<1, KC, real, M, UNI>(op, c, r);
    if (KIND == OP_FILTER) {
        // the filter output is f - A^-1 (A - B) f
#pragma unroll
        for (int t = 0; t < M; t++) r[0][t] = w[HW + t] - r[0][t];
    }
    if (active) win.store(out, r[0], c.len);
}

// The transposed operator (plan.hpp): the line's own values are solved with
// A^T (op: the transposed plan), the solved line is staged in LDS after the
// exchange ring, and every lane applies B^T (or (A - B)^T) to its points:
// the interior stencil (negated for D1) from a window of the staged line,
// the dense boundary columns of tt on non-periodic lines.  The filter's
// transpose is g - (A - B)^T A^-T g.
template <int KIND, typename real, int M, bool UNI, int MAP>
__global__ __launch_bounds__(1024) void k_field_opT(const OpDev<real> *__restrict__ op,
                                                    const OpTDev<real> *__restrict__ tt,
                                                    LineMap lm, const real *__restrict__ in,
                                                    real *__restrict__ out, real coef,
                                                    int accumulate)
{
    extern __shared__ char smem_raw[];
    real *xb = (real *)smem_raw;
    constexpr int HW = KIND == OP_FILTER ? 4 : 2, W = M + 2 * HW;
    constexpr int KC = KIND == OP_FILTER ? kc_filter<real>()
                       : KIND == OP_D2   ? kc_d2<real>()
                                         : kc_deriv<real>();
    const int P = op->P, n = op->n;
    const int LX = blockDim.x / P;
    Lane<real, MAP> c = make_lane<real, MAP>(op, P, LX, op->K > KC ? op->K : KC, 1, xb);
    const int L = blockIdx.x * LX + c.line;
    const bool active = L < lm.nlines;
    Win<real, M, 0, false> win;
    win.init((unsigned)lm.base(active ? L : lm.nlines - 1), (unsigned)lm.sp, c.s, n, op->cyclic);
    real g[M], r[1][M];
#pragma unroll
    for (int t = 0; t < M; t++) {
        g[t] = t < c.len ? ldo(in, win.mid(t)) : real(0);
        r[0][t] = g[t];
    }
    solve_batch<1, KC, real, M, UNI>(op, c, r);
    // the solved line, n + 1 values per line (odd stride across the lines of
    // a column workgroup), after both parities of the ring
    real *st = xb + 2 * c.cap + c.line * (n + 1);
#pragma unroll
    for (int t = 0; t < M; t++)
        if (t < c.len) st[c.s + t] = r[0][t];
    __syncthreads();
    real w[W], res[M];
#pragma unroll
    for (int t = 0; t < W; t++) {
        int p = c.s - HW + t;
        if (op->cyclic) {
            p += p < 0 ? n : 0;
            p -= p >= n ? n : 0;
        } else {
            p = p < 0 ? 0 : p >= n ? n - 1 : p;
        }
        w[t] = st[p];
    }
    stencil<KIND, real, M, HW>(op, w, res);
    if (!op->cyclic) {
        const int JB = tt->JB, UB = tt->UB;
#pragma unroll
        for (int t = 0; t < M; t++) {
            const int j = c.s + t;
            if (t < c.len && j < JB) {
                real a = 0;
                for (int u = 0; u < UB; u++) a += tt->BTl[j][u] * st[u];
                res[t] = KIND == OP_D1 ? -a : a;
            } else if (t < c.len && j >= n - JB) {
                real a = 0;
                for (int u = 0; u < UB; u++) a += tt->BTr[j - (n - JB)][u] * st[n - UB + u];
                res[t] = KIND == OP_D1 ? -a : a;
            }
        }
    }
#pragma unroll
    for (int t = 0; t < M; t++) {
        // the D1 stencil is antisymmetric: its transpose is its negative
        real v = KIND == OP_D1 ? -res[t] : KIND == OP_FILTER ? g[t] - res[t] : res[t];
        v *= coef;
        if (active && t < c.len) {
            real *o = (real *)((char *)out + win.mid(t));
            *o = accumulate ? *o + v : v;
        }
    }
}

template <int KIND, typename real, int M, bool UNI>
static hipError_t launch_kmuT(const PlanDev<real> &pl, int map, const LineMap &lm,
                              const real *in, real *out, real coef, int acc, hipStream_t st)
{
    if (pl.LX < 1 || pl.threads < 1 || pl.threads > 1024 || !pl.tdev)
        return hipErrorInvalidConfiguration;
    dim3 grid((lm.nlines + pl.LX - 1) / pl.LX), block(pl.threads);
    if (map == MAP_ROW)
        hipLaunchKernelGGL((k_field_opT<KIND, real, M, UNI, MAP_ROW>), grid, block, pl.shm, st,
                           pl.dev, pl.tdev, lm, in, out, coef, acc);
    else
        hipLaunchKernelGGL((k_field_opT<KIND, real, M, UNI, MAP_COL>), grid, block, pl.shm, st,
                           pl.dev, pl.tdev, lm, in, out, coef, acc);
    return hipGetLastError();
}

template <int KIND, typename real>
static hipError_t launch_kT(const PlanDev<real> &pl, int map, const LineMap &lm, const real *in,
                            real *out, real coef, int acc, hipStream_t st)
{
    constexpr int MU = KIND == OP_FILTER ? SEG_FILTER : SEG_DERIV;
    if (pl.host.uniform) {
        if (pl.host.M != MU) return hipErrorInvalidConfiguration;
        return launch_kmuT<KIND, real, MU, true>(pl, map, lm, in, out, coef, acc, st);
    }
    if (pl.host.M != SEG_TABLE) return hipErrorInvalidConfiguration;
    return launch_kmuT<KIND, real, SEG_TABLE, false>(pl, map, lm, in, out, coef, acc, st);
}

template <typename real>
hipError_t launch_field_opT(const PlanDev<real> &pl, int map, const LineMap &lm, const real *in,
                            real *out, real coef, int accumulate, hipStream_t st)
{
    if (lm.nlines <= 0) return hipSuccess;
    switch (pl.host.kind) {
    case OP_D1: return launch_kT<OP_D1, real>(pl, map, lm, in, out, coef, accumulate, st);
    case OP_D2: return launch_kT<OP_D2, real>(pl, map, lm, in, out, coef, accumulate, st);
    default: return launch_kT<OP_FILTER, real>(pl, map, lm, in, out, coef, accumulate, st);
    }
}

template <int KIND, typename real, int M, bool UNI>
static hipError_t launch_kmu(const PlanDev<real> &pl, int map, const LineMap &lm,
                             const real *in, real *out, hipStream_t st)
{
    if (pl.LX < 1 || pl.threads < 1 || pl.threads > 1024) return hipErrorInvalidConfiguration;
    dim3 grid((lm.nlines + pl.LX - 1) / pl.LX), block(pl.threads);
    if (map == MAP_ROW)
        hipLaunchKernelGGL((k_field_op<KIND, real, M, UNI, MAP_ROW>), grid, block, pl.shm, st,
                           pl.dev, lm, in, out);
    else
        hipLaunchKernelGGL((k_field_op<KIND, real, M, UNI, MAP_COL>), grid, block, pl.shm, st,
                           pl.dev, lm, in, out);
    return hipGetLastError();
}

template <int KIND, typename real>
static hipError_t launch_k(const PlanDev<real> &pl, int map, const LineMap &lm,
                           const real *in, real *out, hipStream_t st)
{
    constexpr int MU = KIND == OP_FILTER ? SEG_FILTER : SEG_DERIV;
    if (pl.host.uniform) {
        if (pl.host.M != MU) return hipErrorInvalidConfiguration;
        return launch_kmu<KIND, real, MU, true>(pl, map, lm, in, out, st);
    }
    if (pl.host.M != SEG_TABLE) return hipErrorInvalidConfiguration;
    return launch_kmu<KIND, real, SEG_TABLE, false>(pl, map, lm, in, out, st);
}

template <typename real>
hipError_t launch_field_op(const PlanDev<real> &pl, int map, const LineMap &lm,
                           const real *in, real *out, hipStream_t st)
{
    if (lm.nlines <= 0) return hipSuccess;
    switch (pl.host.kind) {
    case OP_D1: return launch_k<OP_D1, real>(pl, map, lm, in, out, st);
    case OP_D2: return launch_k<OP_D2, real>(pl, map, lm, in, out, st);
    default: return launch_k<OP_FILTER, real>(pl, map, lm, in, out, st);
    }
}

template struct PlanDev<float>;
template struct PlanDev<double>;
template void make_plan<float>(PlanDev<float> &, int, int, int, double, double, int, bool);
template void make_plan<double>(PlanDev<double> &, int, int, int, double, double, int, bool);
template hipError_t launch_field_opT<float>(const PlanDev<float> &, int, const LineMap &,
                                            const float *, float *, float, int, hipStream_t);
template hipError_t launch_field_opT<double>(const PlanDev<double> &, int, const LineMap &,
                                             const double *, double *, double, int, hipStream_t);
template hipError_t launch_field_op<float>(const PlanDev<float> &, int, const LineMap &,
                                           const float *, float *, hipStream_t);
template hipError_t launch_field_op<double>(const PlanDev<double> &, int, const LineMap &,
                                            const double *, double *, hipStream_t);

}  // namespace cs3d
