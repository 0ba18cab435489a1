// line_engine.hpp -- device side of the segmented compact line solve.
//
// A workgroup holds LX lines x P segments; each lane owns one segment of at
// most M points of one line, in registers.  The two first-order recurrences of
// the factorised compact matrix (plan.hpp) run segment-locally, and the
// carries between segments go through a small LDS exchange buffer (one
// barrier per recurrence, double-buffered so that the next write never races
// the previous reads).
//
// Lane layouts:
//   MAP_COL: tid = g * LX + line.  Consecutive lanes are consecutive lines;
//            used where lines are strided in memory (Y and Z sweeps of an XYZ
//            field), so every load instruction of a wavefront reads LX
//            consecutive addresses.
//   MAP_ROW: tid = line * P + g.  Consecutive lanes are consecutive segments
//            of one line; used where lines are contiguous (X sweeps).
#pragma once

#include <hip/hip_runtime.h>

#include "plan.hpp"

namespace cs3d {

enum { MAP_COL = 0, MAP_ROW = 1 };

// Maps line L of a batch to its first element; point p of the line is at
// base(L) + p * sp.  base(L) = (L % n1) * s1 + (L / n1) * s2.
struct LineMap {
    int n1;
    int nlines;
    long s1, s2, sp;
    __device__ __forceinline__ long base(int L) const
    {
        return (long)(L % n1) * s1 + (long)(L / n1) * s2;
    }
};

template <typename real, int MAP>
struct Lane {
    int g, line, P, LX;
    int s, len;  // first point and length of this lane's segment
    int par;     // exchange buffer parity
    real *xb;    // LDS exchange buffer, 2 * P * LX entries

    __device__ __forceinline__ int slot(int gg, int ln) const
    {
        return MAP == MAP_COL ? gg * LX + ln : ln * P + gg;
    }
};

template <typename real, int MAP>
__device__ __forceinline__ Lane<real, MAP> make_lane(const OpDev<real> *__restrict__ op,
                                                     int P, int LX, real *xb)
{
    Lane<real, MAP> c;
    int tid = threadIdx.x;
    if (MAP == MAP_COL) {
        c.g = tid / LX;
        c.line = tid - c.g * LX;
    } else {
        c.line = tid / P;
        c.g = tid - c.line * P;
    }
    c.P = P;
    c.LX = LX;
    c.par = 0;
    c.xb = xb;
    if (op->uniform) {
        c.s = c.g * op->M;
        c.len = op->M;
    } else {
        c.s = op->sstart[c.g];
        c.len = op->slen[c.g];
    }
    return c;
}

// Publish v for this lane's segment; after the barrier return
// sum_{k=1..K} w[k-1] * v(segment g - dir*k), wrapping when cyclic.
template <typename real, int MAP>
__device__ __forceinline__ real exchange(Lane<real, MAP> &c, real v, const real *__restrict__ w,
                                         int K, int dir, int cyclic)
{
    real *b = c.xb + c.par * c.P * c.LX;
    c.par ^= 1;
    b[c.slot(c.g, c.line)] = v;
    __syncthreads();
    real acc = 0;
    for (int k = 1; k <= K; k++) {
        int gg = c.g - dir * k;
        if (cyclic) {
            gg += gg < 0 ? c.P : 0;
            gg -= gg >= c.P ? c.P : 0;
        } else if (gg < 0 || gg >= c.P) {
            break;
        }
        acc += w[k - 1] * b[c.slot(gg, c.line)];
    }
    return acc;
}

// Load the window f[s-HW .. s+M+HW) of one line; periodic lines wrap,
// non-periodic ones clamp (clamped values are never used by a stencil).
template <typename real, int M, int HW, typename Src>
__device__ __forceinline__ void load_window(Src src, int s, int n, int periodic,
                                            real (&w)[M + 2 * HW])
{
#pragma unroll
    for (int t = 0; t < M + 2 * HW; t++) {
        int p = s - HW + t;
        if (periodic) {
            p += p < 0 ? n : 0;
            p -= p >= n ? n : 0;
        } else {
            p = p < 0 ? 0 : p;
            p = p >= n ? n - 1 : p;
        }
        w[t] = src(p);
    }
}

// Interior stencil B f for the M points of the segment.  The window w holds
// f[s-HW+t]; the stencil needs HW >= its own half width.
template <int KIND, typename real, int M, int HW>
__device__ __forceinline__ void stencil(const OpDev<real> *__restrict__ op,
                                        const real (&w)[M + 2 * HW], real (&r)[M])
{
    const real t0 = op->tap[0], t1 = op->tap[1], t2 = op->tap[2];
#pragma unroll
    for (int t = 0; t < M; t++) {
        const int o = HW + t;
        if (KIND == OP_D1) {
            static_assert(HW >= 2, "window too narrow");
            r[t] = t1 * (w[o + 1] - w[o - 1]) + t2 * (w[o + 2] - w[o - 2]);
        } else if (KIND == OP_D2) {
            static_assert(HW >= 2, "window too narrow");
            r[t] = t0 * w[o] + t1 * (w[o - 1] + w[o + 1]) + t2 * (w[o - 2] + w[o + 2]);
        } else {
            static_assert(KIND != OP_FILTER || HW >= 4, "window too narrow");
            if constexpr (HW >= 4) {
                const real t3 = op->tap[3], t4 = op->tap[4];
                r[t] = t0 * w[o] + t1 * (w[o - 1] + w[o + 1]) + t2 * (w[o - 2] + w[o + 2]) +
                       t3 * (w[o - 3] + w[o + 3]) + t4 * (w[o - 4] + w[o + 4]);
            }
        }
    }
}

// Boundary-closure rows of non-periodic lines: rows < NB and >= n - NB use
// their own 7-point one-sided stencils (NB = 2 for derivatives, 4 for the
// filter).  The first segment starts at 0 and the last one is full
// (plan.cpp), so every window index is a compile-time constant.
template <int KIND, typename real, int M, int HW>
__device__ __forceinline__ void boundary_rows(const OpDev<real> *__restrict__ op, int s,
                                              int len, const real (&w)[M + 2 * HW],
                                              real (&r)[M])
{
    constexpr int NB = KIND == OP_FILTER ? 4 : 2;
    static_assert(M >= NB && M + HW >= 6 + (KIND == OP_FILTER ? 1 : 0), "window too short");
    if (op->NB == 0) return;
    if (s == 0) {
#pragma unroll
        for (int t = 0; t < NB; t++) {
            real a = 0;
#pragma unroll
            for (int u = 0; u < 7; u++)
                if (HW + u < M + 2 * HW) a += op->bcl[t][u] * w[HW + u];
            r[t] = a;
        }
    }
    if (s + len == op->n) {
        // window index of f[n-7] is HW + (n - 7 - s) = HW + M - 7
#pragma unroll
        for (int rr = 0; rr < NB; rr++) {
            real a = 0;
#pragma unroll
            for (int u = 0; u < 7; u++) {
                constexpr int o = HW + M - 7;
                if (o + u >= 0) a += op->bcr[rr][u] * w[o + u < 0 ? 0 : o + u];
            }
            r[M - NB + rr] = a;
        }
    }
}

// Solve A x = r in place for the segment (r holds B f on entry).
template <typename real, int M, bool UNI, int MAP>
__device__ __forceinline__ void tri_solve(const OpDev<real> *__restrict__ op,
                                          Lane<real, MAP> &c, real (&r)[M])
{
    const int K = op->K, cyc = op->cyclic;
    if (UNI) {
        const real m = op->um, mu = op->umu;
#pragma unroll
        for (int t = 1; t < M; t++) r[t] = r[t] - m * r[t - 1];
        real C = exchange(c, r[M - 1], op->uphi, K, 1, cyc);
#pragma unroll
        for (int t = 0; t < M; t++) r[t] = r[t] + op->upf[t] * C;
#pragma unroll
        for (int t = M - 2; t >= 0; t--) r[t] = r[t] + mu * r[t + 1];
        real D = exchange(c, r[0], op->upsi, K, -1, cyc);
        const real sc = op->scale;
#pragma unroll
        for (int t = 0; t < M; t++) r[t] = sc * (r[t] + op->upb[t] * D);
    } else {
        const int s = c.s, len = c.len;
        const real *__restrict__ tm = op->m + s;
        const real *__restrict__ tiw = op->iw + s;
        const real *__restrict__ tmu = op->mu + s;
        const real *__restrict__ tpf = op->pf + s;
        const real *__restrict__ tpb = op->pb + s;
        real e = r[0];
#pragma unroll
        for (int t = 1; t < M; t++) {
            if (t < len) {
                r[t] = r[t] - tm[t] * r[t - 1];
                e = r[t];
            }
        }
        real C = exchange(c, e, op->phi + (long)c.g * K, K, 1, cyc);
        real carry = 0;
#pragma unroll
        for (int t = M - 1; t >= 0; t--) {
            if (t < len) {
                real y = r[t] + tpf[t] * C;
                carry = tiw[t] * y + tmu[t] * carry;
                r[t] = carry;
            }
        }
        real D = exchange(c, r[0], op->psi + (long)c.g * K, K, -1, cyc);
        const real sc = op->scale;
#pragma unroll
        for (int t = 0; t < M; t++)
            if (t < len) r[t] = sc * (r[t] + tpb[t] * D);
    }
}

// The whole operator on one segment: window -> out.
template <int KIND, typename real, int M, int HW, bool UNI, int MAP>
__device__ __forceinline__ void compact_op(const OpDev<real> *__restrict__ op,
                                           Lane<real, MAP> &c, const real (&w)[M + 2 * HW],
                                           real (&r)[M])
{
    stencil<KIND, real, M, HW>(op, w, r);
    if (!UNI) boundary_rows<KIND, real, M, HW>(op, c.s, c.len, w, r);
    tri_solve<real, M, UNI, MAP>(op, c, r);
}

}  // namespace cs3d
