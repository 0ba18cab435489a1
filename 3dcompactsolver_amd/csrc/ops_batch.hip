// ops_batch.hip -- batched standalone operators along one direction: one
// launch applies the first-derivative plan to N1 fields and the second-
// derivative plan to N2, in one solve batch (two barriers for the batch, each
// field read and written once).  The adjoint's general path forms its
// derivative fields and its transposed terms this way: per stage product 3
// forward launches instead of 30 single-field ones and 6 transposed launches
// instead of 45.  The per-field arithmetic is that of k_field_op and
// k_field_opT (ops.hip).
#include <algorithm>
#include <cstdlib>

#include "stage_common.hpp"

namespace cs3d {

// most lanes per workgroup: column lines read rows of LX adjacent lines,
// so the widest workgroups the registers allow; row workgroups hold one
// line of at most 256 segments (opb_geom)
template <typename real, int MAP>
constexpr int opb_threads()
{
    return MAP == MAP_ROW ? 256 : sizeof(real) == 4 ? 1024 : 512;
}

// solved lines staged in LDS at once by the transposed kernels: batches of
// more than five go in two groups, which halves the staging area (long
// lines keep more lines per workgroup within the LDS)
constexpr int opb_stage_group(int N)
{
    return N <= 5 ? N : (N + 1) / 2;
}

// out[i] = D1(in[i]) for i < N1, out[N1 + i] = D2(in[i]) for i < N2 (N2 <= N1)
template <typename real, int M, bool UNI, int MAP, int N1, int N2>
__global__ __launch_bounds__((opb_threads<real, MAP>())) void k_ops_fwd(
    const OpDev<real> *__restrict__ d1, const OpDev<real> *__restrict__ d2, LineMap lm,
    OpBatch<real> b)
{
    static_assert(N2 <= N1, "second derivatives of the first inputs");
    extern __shared__ char smem_raw[];
    real *xb = (real *)smem_raw;
    constexpr int HW = 2, W = M + 2 * HW, N = N1 + N2;
    const int P = d1->P;
    const int LX = blockDim.x / P;
    Lane<real, MAP> c = make_lane<real, MAP>(d1, P, LX, sweep_kr(d1, d2), N, xb);
    const int L = blockIdx.x * LX + c.line;
    const bool active = L < lm.nlines;
    Win<real, M, HW, vec_windows<real, M, MAP, UNI>()> win;
    win.init((unsigned)lm.base(active ? L : lm.nlines - 1), (unsigned)lm.sp, c.s, d1->n,
             d1->cyclic);
    real r[N][M];
#pragma unroll
    for (int i = 0; i < N1; i++) {
        real w[W];
        win.load(b.in[i], w);
        rhs<OP_D1, real, M, HW, UNI>(d1, c, w, r[i]);
        if (i < N2) rhs<OP_D2, real, M, HW, UNI>(d2, c, w, r[N1 + (i < N2 ? i : 0)]);
    }
    solve_batch2<N1, N2, kc_deriv<real>(), kc_d2<real>(), real, M, UNI, MAP>(d1, d2, c, r);
    if (active) {
#pragma unroll
        for (int i = 0; i < N; i++) win.store(b.out[i], r[i], c.len);
    }
}

// Transposed, PAIR (N1 == N2): out[f] (+)= coef[f] D1^T(in[f])
// + coef[N1 + f] D2^T(in[N1 + f]); LIST (N2 == 0): out[i] (+)= coef[i]
// D1^T(in[i]).  acc[f] selects accumulate or overwrite per output.  The
// solved lines are staged in LDS over the exchange ring (opb_stage_group
// at a time) for the B^T stencils.
template <typename real, int M, bool UNI, int MAP, int N1, int N2>
__global__ __launch_bounds__((opb_threads<real, MAP>())) void k_ops_T(
    const OpDev<real> *__restrict__ d1, const OpDev<real> *__restrict__ d2,
    const OpTDev<real> *__restrict__ t1, const OpTDev<real> *__restrict__ t2, LineMap lm,
    OpBatch<real> b)
{
    static_assert(N2 == 0 || N2 == N1, "a list of D1 terms or pairs of D1 and D2 terms");
    extern __shared__ char smem_raw[];
    real *xb = (real *)smem_raw;
    constexpr int HW = 2, W = M + 2 * HW, N = N1 + N2;
    const int P = d1->P, n = d1->n;
    const int LX = blockDim.x / P;
    Lane<real, MAP> c = make_lane<real, MAP>(d1, P, LX, sweep_kr(d1, d2), N, xb);
    const int L = blockIdx.x * LX + c.line;
    const bool active = L < lm.nlines;
    Win<real, M, 0, false> win;
    win.init((unsigned)lm.base(active ? L : lm.nlines - 1), (unsigned)lm.sp, c.s, n, d1->cyclic);
    real r[N][M];
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int t = 0; t < M; t++) r[i][t] = t < c.len ? ldo(b.in[i], win.mid(t)) : real(0);
    // the first NP outputs, where accumulated, are read with the inputs, so
    // that their loads overlap the first ones instead of following the
    // solve (any later accumulated output is read at the end)
    constexpr int NO = N2 ? N1 : N;
    constexpr int NP = NO < 4 ? NO : 4;
    real ob[NP][M];
#pragma unroll
    for (int f = 0; f < NP; f++)
#pragma unroll
        for (int t = 0; t < M; t++)
            ob[f][t] = b.acc[f] && active && t < c.len ? ldo(b.out[f], win.mid(t)) : real(0);
    solve_batch2<N1, N2, kc_deriv<real>(), kc_d2<real>(), real, M, UNI, MAP>(d1, d2, c, r);
    // B^T of every solve into r[i] (D1 negated: its interior stencil is
    // antisymmetric), weighted.  The solved lines go through LDS in groups
    // of SG (opb_stage_group), each group staged over the ring after a
    // barrier (every lane past the last ring read, or the previous group's
    // window reads).
    constexpr int SG = opb_stage_group(N);
    const int ls = n + 1;  // odd stride across the lines of a column workgroup
#pragma unroll
    for (int g0 = 0; g0 < N; g0 += SG) {
        __syncthreads();
#pragma unroll
        for (int i = g0; i < g0 + SG && i < N; i++) {
            real *st = xb + ((i - g0) * LX + c.line) * ls;
#pragma unroll
            for (int t = 0; t < M; t++)
                if (t < c.len) st[c.s + t] = r[i][t];
        }
        __syncthreads();
#pragma unroll
        for (int i = g0; i < g0 + SG && i < N; i++) {
            const bool first = i < N1;
            const OpDev<real> *__restrict__ op = first ? d1 : d2;
            const OpTDev<real> *__restrict__ tt = first ? t1 : t2;
            const real *st = xb + ((i - g0) * LX + c.line) * ls;
            real w[W], res[M];
#pragma unroll
            for (int t = 0; t < W; t++) {
                int p = c.s - HW + t;
                if (d1->cyclic) {
                    p += p < 0 ? n : 0;
                    p -= p >= n ? n : 0;
                } else {
                    p = p < 0 ? 0 : p >= n ? n - 1 : p;
                }
                w[t] = st[p];
            }
            if (first)
                stencil<OP_D1, real, M, HW>(op, w, res);
            else
                stencil<OP_D2, real, M, HW>(op, w, res);
            if (!d1->cyclic) {
                const int JB = tt->JB, UB = tt->UB;
#pragma unroll
                for (int t = 0; t < M; t++) {
                    const int j = c.s + t;
                    // (the dense columns hold B^T itself: pre-negated for
                    // D1, whose weight below negates the interior stencil)
                    if (t < c.len && j < JB) {
                        real a = 0;
                        for (int u = 0; u < UB; u++) a += tt->BTl[j][u] * st[u];
                        res[t] = first ? -a : a;
                    } else if (t < c.len && j >= n - JB) {
                        real a = 0;
                        for (int u = 0; u < UB; u++)
                            a += tt->BTr[j - (n - JB)][u] * st[n - UB + u];
                        res[t] = first ? -a : a;
                    }
                }
            }
            const real cf = first ? -b.coef[i] : b.coef[i];
#pragma unroll
            for (int t = 0; t < M; t++) r[i][t] = cf * res[t];
        }
    }
#pragma unroll
    for (int f = 0; f < NO; f++) {
        real *out = b.out[f];
        const int acc = b.acc[f];
#pragma unroll
        for (int t = 0; t < M; t++) {
            if (active && t < c.len) {
                real *o = (real *)((char *)out + win.mid(t));
                const real v = N2 ? r[f][t] + r[N1 + (N2 ? f : 0)][t] : r[f][t];
                if (f < NP)
                    *o = ob[f < NP ? f : 0][t] + v;
                else
                    *o = acc ? *o + v : v;
            }
        }
    }
}

// lines per workgroup (a power of two on column lines) and LDS bytes: the
// ring for N planes, and for the transposed kernels the N solved lines
// staged over it
template <typename real>
static LaunchGeom opb_geom(const OpHost &h1, const OpHost &h2, int N, bool trans, int map)
{
    const int kr = std::max({h1.K, h2.K, kc_deriv<real>()});
    const int P = h1.P;
    auto bytes = [&](int LX) {
        const size_t ring = ring_bytes(P, kr, LX, N, sizeof(real), map == MAP_COL);
        const size_t stage =
            trans ? (size_t)opb_stage_group(N) * LX * (h1.n + 1) * sizeof(real) : 0;
        return ring > stage ? ring : stage;
    };
    LaunchGeom g;
    // row lines (contiguous): one line per workgroup where a line fills a
    // wavefront, so that workgroups on a CU run out of phase, as the x sweep
    int col = opb_threads<real, MAP_COL>();
    if (const char *v = std::getenv("CS3D_OPS_BATCH_THREADS")) {
        const int x = std::atoi(v);
        if (x >= 64 && x <= col) col = x;
    }
    const int cap = map == MAP_ROW ? std::min(opb_threads<real, MAP_ROW>(), std::max(64, P)) : col;
    g.LX = P > 0 ? std::min(64, cap / P) : 0;
    if (map == MAP_COL)
        while (g.LX & (g.LX - 1)) g.LX &= g.LX - 1;
    while (g.LX > 0 && bytes(g.LX) > LDS_BYTES) g.LX = map == MAP_COL ? g.LX / 2 : g.LX - 1;
    g.threads = g.LX * P;
    g.shm = g.LX > 0 ? bytes(g.LX) : 0;
    return g;
}

template <typename real, int M, bool UNI, int N1, int N2>
static hipError_t launch_fwd(const PlanDev<real> &d1, const PlanDev<real> &d2, int map,
                             const LineMap &lm, const OpBatch<real> &b, hipStream_t st)
{
    const LaunchGeom g = opb_geom<real>(d1.host, d2.host, N1 + N2, false, map);
    if (g.LX < 1) return hipErrorInvalidConfiguration;
    dim3 grid((lm.nlines + g.LX - 1) / g.LX), block(g.threads);
    if (map == MAP_ROW)
        hipLaunchKernelGGL((k_ops_fwd<real, M, UNI, MAP_ROW, N1, N2>), grid, block, g.shm, st,
                           d1.dev, d2.dev, lm, b);
    else
        hipLaunchKernelGGL((k_ops_fwd<real, M, UNI, MAP_COL, N1, N2>), grid, block, g.shm, st,
                           d1.dev, d2.dev, lm, b);
    return hipGetLastError();
}

template <typename real, int M, bool UNI, int N1, int N2>
static hipError_t launch_T(const PlanDev<real> &d1, const PlanDev<real> &d2, int map,
                           const LineMap &lm, const OpBatch<real> &b, hipStream_t st)
{
    const LaunchGeom g = opb_geom<real>(d1.host, d2.host, N1 + N2, true, map);
    if (g.LX < 1) return hipErrorInvalidConfiguration;
    dim3 grid((lm.nlines + g.LX - 1) / g.LX), block(g.threads);
    if (map == MAP_ROW)
        hipLaunchKernelGGL((k_ops_T<real, M, UNI, MAP_ROW, N1, N2>), grid, block, g.shm, st,
                           d1.dev, d2.dev, d1.tdev, d2.tdev, lm, b);
    else
        hipLaunchKernelGGL((k_ops_T<real, M, UNI, MAP_COL, N1, N2>), grid, block, g.shm, st,
                           d1.dev, d2.dev, d1.tdev, d2.tdev, lm, b);
    return hipGetLastError();
}

// the plans of one batch: one direction, the same segmentation, the
// derivative kinds in order, transposed or not as the launch asks
template <typename real>
static bool opb_plans_ok(const PlanDev<real> &d1, const PlanDev<real> &d2, bool trans)
{
    const OpHost &a = d1.host, &c = d2.host;
    if (!d1.dev || !d2.dev || (trans && (!d1.tdev || !d2.tdev))) return false;
    return a.kind == OP_D1 && c.kind == OP_D2 && a.n == c.n && a.P == c.P && a.M == c.M &&
           a.uniform == c.uniform && a.cyclic == c.cyclic &&
           (a.uniform ? a.M == SEG_DERIV || a.M == SEG_DERIV3 : a.M == SEG_TABLE);
}

template <typename real>
hipError_t launch_ops_fwd(const PlanDev<real> &d1, const PlanDev<real> &d2, int map,
                          const LineMap &lm, int n1, int n2, const OpBatch<real> &b,
                          hipStream_t st)
{
    if (lm.nlines <= 0) return hipSuccess;
    if (!opb_plans_ok(d1, d2, false)) return hipErrorInvalidValue;
    const bool uni = d1.host.uniform != 0;
    if (n2 != 4) return hipErrorInvalidValue;
    // 3-point segments: the z derivatives (8 + 4) only
    if (uni && d1.host.M == SEG_DERIV3)
        return n1 == 8 ? launch_fwd<real, SEG_DERIV3, true, 8, 4>(d1, d2, map, lm, b, st)
                       : hipErrorInvalidValue;
    switch (n1) {
    case 4:
        return uni ? launch_fwd<real, SEG_DERIV, true, 4, 4>(d1, d2, map, lm, b, st)
                   : launch_fwd<real, SEG_TABLE, false, 4, 4>(d1, d2, map, lm, b, st);
    case 6:
        return uni ? launch_fwd<real, SEG_DERIV, true, 6, 4>(d1, d2, map, lm, b, st)
                   : launch_fwd<real, SEG_TABLE, false, 6, 4>(d1, d2, map, lm, b, st);
    case 8:
        return uni ? launch_fwd<real, SEG_DERIV, true, 8, 4>(d1, d2, map, lm, b, st)
                   : launch_fwd<real, SEG_TABLE, false, 8, 4>(d1, d2, map, lm, b, st);
    default: return hipErrorInvalidValue;
    }
}

template <typename real>
hipError_t launch_ops_T(const PlanDev<real> &d1, const PlanDev<real> &d2, int map,
                        const LineMap &lm, int n1, int n2, const OpBatch<real> &b,
                        hipStream_t st)
{
    if (lm.nlines <= 0) return hipSuccess;
    if (!opb_plans_ok(d1, d2, true)) return hipErrorInvalidValue;
    const bool uni = d1.host.uniform != 0;
    // 3-point segments: the z terms (the pair and the list of nine) only
    if (uni && d1.host.M == SEG_DERIV3) {
        if (n1 == 4 && n2 == 4) return launch_T<real, SEG_DERIV3, true, 4, 4>(d1, d2, map, lm, b, st);
        if (n1 == 9 && n2 == 0) return launch_T<real, SEG_DERIV3, true, 9, 0>(d1, d2, map, lm, b, st);
        return hipErrorInvalidValue;
    }
    if (n2 == 4 && n1 == 4)
        return uni ? launch_T<real, SEG_DERIV, true, 4, 4>(d1, d2, map, lm, b, st)
                   : launch_T<real, SEG_TABLE, false, 4, 4>(d1, d2, map, lm, b, st);
    if (n2 != 0) return hipErrorInvalidValue;
    switch (n1) {
    case 5:
        return uni ? launch_T<real, SEG_DERIV, true, 5, 0>(d1, d2, map, lm, b, st)
                   : launch_T<real, SEG_TABLE, false, 5, 0>(d1, d2, map, lm, b, st);
    case 7:
        return uni ? launch_T<real, SEG_DERIV, true, 7, 0>(d1, d2, map, lm, b, st)
                   : launch_T<real, SEG_TABLE, false, 7, 0>(d1, d2, map, lm, b, st);
    case 9:
        return uni ? launch_T<real, SEG_DERIV, true, 9, 0>(d1, d2, map, lm, b, st)
                   : launch_T<real, SEG_TABLE, false, 9, 0>(d1, d2, map, lm, b, st);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_ops_fwd<float>(const PlanDev<float> &, const PlanDev<float> &, int,
                                          const LineMap &, int, int, const OpBatch<float> &,
                                          hipStream_t);
template hipError_t launch_ops_fwd<double>(const PlanDev<double> &, const PlanDev<double> &, int,
                                           const LineMap &, int, int, const OpBatch<double> &,
                                           hipStream_t);
template hipError_t launch_ops_T<float>(const PlanDev<float> &, const PlanDev<float> &, int,
                                        const LineMap &, int, int, const OpBatch<float> &,
                                        hipStream_t);
template hipError_t launch_ops_T<double>(const PlanDev<double> &, const PlanDev<double> &, int,
                                         const LineMap &, int, int, const OpBatch<double> &,
                                         hipStream_t);

}  // namespace cs3d
