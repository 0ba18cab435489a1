// line_engine.hpp -- device side of the segmented compact line solve.
//
// A workgroup holds LX lines x P segments; each lane owns one segment of at
// most M points of one line, in registers.  The two first-order recurrences of
// the factorised compact matrix (plan.hpp) run segment-locally, and the
// carries between segments go through an LDS exchange ring: one barrier per
// recurrence, shared by every solve of a batch (solve_batch), double-buffered
// so that the next write never races the previous reads.
//
// Lane layouts:
//   MAP_COL: tid = g * LX + line.  Consecutive lanes are consecutive lines;
//            used where lines are strided in memory (Y and Z sweeps of an XYZ
//            field), so every load instruction of a wavefront reads LX
//            consecutive addresses.
//   MAP_ROW: tid = line * P + g.  Consecutive lanes are consecutive segments
//            of one line; used where lines are contiguous (X sweeps).
//
// Addressing: every field of a sweep has the same layout, so a lane computes
// the byte offsets of its window once (window_offsets) and every load and
// store after that is one instruction with a 32-bit offset from the field's
// base address.  Fields must therefore stay below 4 GiB each (checked on the
// host).
//
// Exchange ring: per line, a row of P + KR positions (KR >= the K of every
// operator using the ring, row padded to spread the lines of a wavefront over
// the LDS banks); each position holds the values of the whole batch side by
// side.  The forward carry of segment g
// reads ring[KR + g - k] for k = 1..K: the primary values sit at KR + g and
// the last KR segments also write a copy at KR + g - P (the cyclic wrap; zero
// for non-periodic lines).  The backward carry reads ring[g + k]: primaries
// at g, and the first KR segments copy to P + g.  No index ever wraps at run
// time, so the reads are plain LDS loads at constant offsets.
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
    __host__ __device__ __forceinline__ long base(int L) const
    {
        return (long)(L % n1) * s1 + (long)(L / n1) * s2;
    }
};

// Multiplicative inverse of an odd number modulo 32.
__host__ __device__ constexpr int inv_mod32(int s)
{
    int x = 1;
    for (int i = 0; i < 32; i++) {
        if (((s * x) & 31) == 1) return x;
        x += 2;
    }
    return 1;
}

// Ring row length (in positions) for P segments, margin KR, LX lines per
// workgroup, an odd position stride S and esize-byte elements: at least
// P + KR, and chosen so that row * S = B / LX (mod B).  The LDS has 32 banks
// of 4 bytes and serves a 4-byte read in groups of B = 32 lanes, an 8-byte
// read in groups of B = 16 (two banks each); the lanes of one group then hit
// distinct banks: in MAP_COL they are B / LX consecutive segments of LX
// lines, at offsets line * row * S + g * S; in MAP_ROW consecutive segments
// of one line, at offsets g * S.  (FP64 with the 4-byte rule, row * S =
// 32 / LX mod 32, put lines l and l + 4 of an 8-line group on one bank pair:
// the FP64 y and z sweeps ran 13% slower.)  A workgroup of one line needs
// no padding (its lanes are consecutive segments at offsets g * S either
// way), and in a MAP_COL workgroup of 32 or more lines a group holds one
// segment of B lines, for which any odd row * S will do (col: row odd, at
// most one position of padding instead of up to B - 1).
__host__ __device__ inline int ring_row(int P, int KR, int LX, int S, bool col, int esize)
{
    if (LX == 1) return P + KR;
    if (col && LX >= 32) return (P + KR) | 1;
    const int bk = esize == 8 ? 16 : 32;
    const int want = LX >= bk ? 1 : bk / LX;
    const int r = P + KR;
    const int target = (want * inv_mod32(S)) & (bk - 1);  // S^-1 mod 32 is S^-1 mod 16 too
    return r + (((target - r) % bk) + bk) % bk;
}

// Elements of one parity of the exchange buffer, for batches of up to
// `planes` values per position.
__host__ __device__ inline int ring_cap(int P, int KR, int LX, int planes, bool col, int esize)
{
    int best = 0;
    for (int S = 1; S <= (planes | 1); S += 2) {
        const int e = ring_row(P, KR, LX, S, col, esize) * S;
        best = e > best ? e : best;
    }
    return LX * best;
}

// LDS bytes of a workgroup's exchange buffers (two parities).
__host__ __device__ inline size_t ring_bytes(int P, int KR, int LX, int planes, int esize,
                                             bool col)
{
    return 2 * (size_t)ring_cap(P, KR, LX, planes, col, esize) * esize;
}

// Workgroup shape of a line-solve launch: LX lines x P segments per
// workgroup, and the LDS bytes of its exchange ring.
struct LaunchGeom {
    int threads;
    int LX;
    size_t shm;
};

// LDS of one workgroup on gfx950
constexpr size_t LDS_BYTES = 160 * 1024;

// LX = max_threads / P lines (at most 64, fewer when the exchange ring would
// not fit the LDS) per workgroup, ring margin kr, `planes` values per ring
// position, and room for stage_per_lane window values per lane
// (load_windows_lds) in the same LDS; LX = 0 when a line does not fit.
// map: the lane layout (MAP_COL or MAP_ROW) of the kernel launched.
// pow2: LX rounded down to a power of two (column sweeps: a workgroup's
// lines are adjacent x points, and 8 or 16 of them fill aligned 32- or
// 64-byte pieces of a row where 12 or 15 straddle them).
inline LaunchGeom line_geom(int P, int kr, int max_threads, int planes, int esize, int map,
                            int stage_per_lane = 0, bool pow2 = false)
{
    auto bytes = [&](int LX) {
        const size_t ring = ring_bytes(P, kr, LX, planes, esize, map == MAP_COL);
        const size_t stage = (size_t)stage_per_lane * LX * P * esize;
        return ring > stage ? ring : stage;
    };
    LaunchGeom g;
    g.LX = P > 0 ? max_threads / P : 0;
    if (g.LX > 64) g.LX = 64;
    if (pow2)
        while (g.LX & (g.LX - 1)) g.LX &= g.LX - 1;
    while (g.LX > 0 && bytes(g.LX) > LDS_BYTES) g.LX = pow2 ? g.LX / 2 : g.LX - 1;
    g.threads = g.LX * P;
    g.shm = g.LX > 0 ? bytes(g.LX) : 0;
    return g;
}

template <typename real, int MAP>
struct Lane {
    int g, line, P, LX;
    int s, len;   // first point and length of this lane's segment
    int KR;       // ring margin (>= every K read through the ring)
    int cap;      // elements of one parity of the exchange buffer
    int par;      // exchange buffer parity
    real *xb;     // LDS exchange buffers
};

template <typename real, int MAP>
__device__ __forceinline__ Lane<real, MAP> make_lane(const OpDev<real> *__restrict__ op, int P,
                                                     int LX, int KR, int planes, real *xb)
{
    Lane<real, MAP> c;
    const int tid = threadIdx.x;
    if (MAP == MAP_COL) {
        c.g = tid / LX;
        c.line = tid - c.g * LX;
    } else {
        c.line = tid / P;
        c.g = tid - c.line * P;
    }
    c.P = P;
    c.LX = LX;
    c.KR = KR;
    c.cap = ring_cap(P, KR, LX, planes, MAP == MAP_COL, (int)sizeof(real));
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

// ------------------------------------------------------------ addressing

template <typename real>
__device__ __forceinline__ real ldo(const real *__restrict__ f, unsigned boff)
{
    return *(const real *)((const char *)f + boff);
}

template <typename real>
__device__ __forceinline__ void sto(real *__restrict__ f, unsigned boff, real v)
{
    *(real *)((char *)f + boff) = v;
}

// a store, non-temporal when NT
template <bool NT, typename T>
__device__ __forceinline__ void st_hint(T v, T *p)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

// A lane's window f[s-HW .. s+M+HW) of one line.
//
// Scalar form: the byte offset of every window point (periodic lines wrap,
// non-periodic ones clamp; clamped entries are never used by a stencil).
//
// VEC form (contiguous lines, uniform periodic plans, M a multiple of the
// 16-byte vector width VW): the window, widened to whole VW-aligned chunks,
// is read with one 16-byte load per chunk, and the segment is stored the same
// way.  With n a multiple of VW no chunk straddles the periodic wrap.
//
// Interleaved fields: a pair field holds two values per point (a, b) and a
// quad field four, at point index * 2 (* 4) elements, so the offsets above
// scale by 2 (4).  Column lanes then move 2 (4) values per point with one
// load or store, which widens the contiguous run of a wavefront row from
// 16 lines * 4 bytes to 128 (256) bytes.
template <typename real, int M, int HW, bool VEC>
struct Win {
    static constexpr int W = M + 2 * HW;
    static constexpr int VW = 16 / (int)sizeof(real);
    static constexpr int LO = VEC ? ((HW + VW - 1) / VW) * VW : HW;  // left extent
    static constexpr int D = LO - HW;                                // chunk -> window shift
    static constexpr int NC = VEC ? (M + HW + LO + VW - 1) / VW : W;
    static_assert(!VEC || M % VW == 0, "vector windows need whole chunks per segment");
    typedef real vec __attribute__((ext_vector_type(16 / sizeof(real))));
    unsigned off[NC];
    unsigned spb;  // byte stride between consecutive points of the line

    __device__ __forceinline__ void init(unsigned base, unsigned sp, int s, int n, int periodic)
    {
        spb = sp * (unsigned)sizeof(real);
#pragma unroll
        for (int c = 0; c < NC; c++) {
            int p = VEC ? s - LO + c * VW : s - HW + c;
            if (VEC || periodic) {
                p += p < 0 ? n : 0;
                p -= p >= n ? n : 0;
            } else {
                p = p < 0 ? 0 : p;
                p = p >= n ? n - 1 : p;
            }
            off[c] = (base + (unsigned)p * sp) * (unsigned)sizeof(real);
        }
    }
    // byte offset of the segment's own point t (0 <= t < len): these never
    // wrap, so only the first one needs keeping
    __device__ __forceinline__ unsigned mid(int t) const
    {
        if (VEC) return at(HW + t);
        return off[HW] + (unsigned)t * spb;
    }
    // byte offset of window point t
    __device__ __forceinline__ unsigned at(int t) const
    {
        if (VEC) return off[(t + D) / VW] + (unsigned)(((t + D) % VW) * sizeof(real));
        return off[t];
    }
    __device__ __forceinline__ void load(const real *__restrict__ f, real (&w)[W]) const
    {
        if (VEC) {
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const vec v = *(const vec *)((const char *)f + off[c]);
#pragma unroll
                for (int e = 0; e < VW; e++) {
                    const int t = c * VW + e - D;
                    if (t >= 0 && t < W) w[t] = v[e];
                }
            }
        } else {
#pragma unroll
            for (int t = 0; t < W; t++) w[t] = ldo(f, off[t]);
        }
    }
    // the segment's own M values
    __device__ __forceinline__ void load_seg(const real *__restrict__ f, real (&r)[M]) const
    {
        if (VEC) {
#pragma unroll
            for (int c = 0; c < M / VW; c++) {
                const vec v = *(const vec *)((const char *)f + off[LO / VW + c]);
#pragma unroll
                for (int e = 0; e < VW; e++) r[c * VW + e] = v[e];
            }
        } else {
#pragma unroll
            for (int t = 0; t < M; t++) r[t] = ldo(f, off[HW + t]);
        }
    }
    // store the segment's M values (the first len of them on table plans).
    // NT: non-temporal stores, for results the next kernels read from HBM
    // anyway (the sweeps' outputs: 4% faster FP32 steps, 3% FP64), so that
    // they do not displace the windows being read from the caches (with
    // ordinary stores in the y sweep alone, the z sweep after it runs 11%
    // slower)
    template <bool NT = false>
    __device__ __forceinline__ void store(real *__restrict__ f, const real (&r)[M], int len) const
    {
        if (VEC) {
#pragma unroll
            for (int c = 0; c < M / VW; c++) {
                vec v;
#pragma unroll
                for (int e = 0; e < VW; e++) v[e] = r[c * VW + e];
                st_hint<NT>(v, (vec *)((char *)f + off[LO / VW + c]));
            }
        } else {
#pragma unroll
            for (int t = 0; t < M; t++)
                if (t < len) st_hint<NT>(r[t], (real *)((char *)f + off[HW + t]));
        }
    }

    // ---- pair fields
    typedef real vec2 __attribute__((ext_vector_type(2)));
    typedef real vec4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ const char *pair_at(const real *f, unsigned o)
    {
        return (const char *)f + 2ull * o;
    }
    // the window of both values
    __device__ __forceinline__ void load_pair(const real *__restrict__ f, real (&wa)[W],
                                              real (&wb)[W]) const
    {
        if (VEC) {
            // a chunk of VW points is 2 * VW values: two 16-byte loads
#pragma unroll
            for (int c = 0; c < NC; c++) {
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const vec v = *(const vec *)(pair_at(f, off[c]) + 16 * h);
#pragma unroll
                    for (int e = 0; e < VW; e++) {
                        const int t = c * VW + h * (VW / 2) + e / 2 - D;
                        if (t >= 0 && t < W) {
                            if (e % 2)
                                wb[t] = v[e];
                            else
                                wa[t] = v[e];
                        }
                    }
                }
            }
        } else {
#pragma unroll
            for (int t = 0; t < W; t++) {
                const vec2 v = *(const vec2 *)pair_at(f, off[t]);
                wa[t] = v[0];
                wb[t] = v[1];
            }
        }
    }
    // the segment's own M points of both values
    __device__ __forceinline__ void load_seg_pair(const real *__restrict__ f, real (&ra)[M],
                                                  real (&rb)[M]) const
    {
        if (VEC) {
#pragma unroll
            for (int c = 0; c < M / VW; c++) {
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const vec v = *(const vec *)(pair_at(f, off[LO / VW + c]) + 16 * h);
#pragma unroll
                    for (int e = 0; e < VW; e++) {
                        const int t = c * VW + h * (VW / 2) + e / 2;
                        if (e % 2)
                            rb[t] = v[e];
                        else
                            ra[t] = v[e];
                    }
                }
            }
        } else {
#pragma unroll
            for (int t = 0; t < M; t++) {
                const vec2 v = *(const vec2 *)pair_at(f, off[HW + t]);
                ra[t] = v[0];
                rb[t] = v[1];
            }
        }
    }
    template <bool NT = false>
    __device__ __forceinline__ void store_pair(real *__restrict__ f, const real (&ra)[M],
                                               const real (&rb)[M], int len) const
    {
        if (VEC) {
#pragma unroll
            for (int c = 0; c < M / VW; c++) {
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    vec v;
#pragma unroll
                    for (int e = 0; e < VW; e++) {
                        const int t = c * VW + h * (VW / 2) + e / 2;
                        v[e] = e % 2 ? rb[t] : ra[t];
                    }
                    st_hint<NT>(v, (vec *)(pair_at(f, off[LO / VW + c]) + 16 * h));
                }
            }
        } else {
#pragma unroll
            for (int t = 0; t < M; t++)
                if (t < len) {
                    vec2 v;
                    v[0] = ra[t];
                    v[1] = rb[t];
                    st_hint<NT>(v, (vec2 *)pair_at(f, off[HW + t]));
                }
        }
    }

    // ---- quad fields (one point = 4 values: one 16-byte load in FP32)
    static __device__ __forceinline__ const char *quad_at(const real *f, unsigned o)
    {
        return (const char *)f + 4ull * o;
    }
    // the window of a quad field: one vector load per window point
    __device__ __forceinline__ void load_quad(const real *__restrict__ f, real (&w0)[W],
                                              real (&w1)[W], real (&w2)[W], real (&w3)[W]) const
    {
#pragma unroll
        for (int t = 0; t < W; t++) {
            const vec4 v = *(const vec4 *)quad_at(f, at(t));
            w0[t] = v[0];
            w1[t] = v[1];
            w2[t] = v[2];
            w3[t] = v[3];
        }
    }
    __device__ __forceinline__ void load_seg_quad(const real *__restrict__ f, real (&r0)[M],
                                                  real (&r1)[M], real (&r2)[M],
                                                  real (&r3)[M]) const
    {
#pragma unroll
        for (int t = 0; t < M; t++) {
            const vec4 v = *(const vec4 *)quad_at(f, mid(t));
            r0[t] = v[0];
            r1[t] = v[1];
            r2[t] = v[2];
            r3[t] = v[3];
        }
    }
    // one point's quad (t < len on table plans: the caller checks)
    template <bool NT = false>
    __device__ __forceinline__ void store_quad_at(real *__restrict__ f, int t, real r0, real r1,
                                                  real r2, real r3) const
    {
        vec4 v;
        v[0] = r0;
        v[1] = r1;
        v[2] = r2;
        v[3] = r3;
        st_hint<NT>(v, (vec4 *)quad_at(f, mid(t)));
    }
    template <bool NT = false>
    __device__ __forceinline__ void store_quad(real *__restrict__ f, const real (&r0)[M],
                                               const real (&r1)[M], const real (&r2)[M],
                                               const real (&r3)[M], int len) const
    {
#pragma unroll
        for (int t = 0; t < M; t++)
            if (VEC || t < len) {
                vec4 v;
                v[0] = r0[t];
                v[1] = r1[t];
                v[2] = r2[t];
                v[3] = r3[t];
                st_hint<NT>(v, (vec4 *)quad_at(f, mid(t)));
            }
    }
};

// one value of an interleaved field at the point of single-field byte
// offset o: component k of N
template <int N, typename real>
__device__ __forceinline__ real ldo_part(const real *__restrict__ f, unsigned o, int k)
{
    return *(const real *)((const char *)f + (unsigned long long)N * o + (unsigned)k * sizeof(real));
}

// Windows of NF fields for lanes of uniform periodic MAP_COL plans, without
// re-reading the neighbours' points: each lane loads its own M points of
// every field (src(f, offset) returns field f at a byte offset), publishes
// them in LDS, and takes its HW halo points on each side from the
// neighbouring segments' slots (cyclic).  The staging area is the exchange
// buffer, NF * M * lanes elements; the function starts and ends with a
// barrier, so it may be called between solve batches.
// Core of load_windows_lds: pt(t, v) loads the NF values of own point t.
template <int NF, typename real, int M, int HW, typename Pt>
__device__ __forceinline__ void load_windows_lds_pt(const Lane<real, MAP_COL> &c, Pt pt,
                                                    real (&w)[NF][M + 2 * HW])
{
    static_assert(HW <= M, "halo wider than a segment");
    real *st = c.xb;
    const int T = c.P * c.LX;
    const int me = c.g * c.LX + c.line;
    __syncthreads();
#pragma unroll
    for (int t = 0; t < M; t++) {
        real v[NF];
        pt(t, v);
#pragma unroll
        for (int f = 0; f < NF; f++) {
            w[f][HW + t] = v[f];
            st[(f * M + t) * T + me] = v[f];
        }
    }
    __syncthreads();
    const int gl = c.g == 0 ? c.P - 1 : c.g - 1;
    const int gr = c.g == c.P - 1 ? 0 : c.g + 1;
    const int ml = gl * c.LX + c.line, mr = gr * c.LX + c.line;
#pragma unroll
    for (int f = 0; f < NF; f++) {
#pragma unroll
        for (int h = 0; h < HW; h++) {
            w[f][h] = st[(f * M + M - HW + h) * T + ml];
            w[f][HW + M + h] = st[(f * M + h) * T + mr];
        }
    }
    __syncthreads();
}

template <int NF, typename real, int M, int HW, bool VEC, typename Src>
__device__ __forceinline__ void load_windows_lds(const Lane<real, MAP_COL> &c,
                                                 const Win<real, M, HW, VEC> &win, Src src,
                                                 real (&w)[NF][M + 2 * HW])
{
    load_windows_lds_pt<NF, real, M, HW>(c, [&](int t, real (&v)[NF]) {
#pragma unroll
        for (int f = 0; f < NF; f++) v[f] = src(f, win.mid(t));
    }, w);
}

// the windows of both values of a pair field, one load per own point
template <typename real, int M, int HW, bool VEC>
__device__ __forceinline__ void load_windows_lds_pair(const Lane<real, MAP_COL> &c,
                                                      const Win<real, M, HW, VEC> &win,
                                                      const real *__restrict__ f,
                                                      real (&w)[2][M + 2 * HW])
{
    typedef typename Win<real, M, HW, VEC>::vec2 vec2;
    load_windows_lds_pt<2, real, M, HW>(c, [&](int t, real (&v)[2]) {
        const vec2 x = *(const vec2 *)Win<real, M, HW, VEC>::pair_at(f, win.mid(t));
        v[0] = x[0];
        v[1] = x[1];
    }, w);
}

// LDS elements per lane that load_windows_lds stages for NF fields
constexpr int stage_per_lane(int NF, int M) { return NF * M; }

// whether a kernel instantiation reads its windows as 16-byte chunks
template <typename real, int M, int MAP, bool UNI>
constexpr bool vec_windows()
{
    return MAP == MAP_ROW && UNI && M % (16 / (int)sizeof(real)) == 0;
}

// ------------------------------------------------------------ stencils

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
        }
    }
    if (KIND == OP_FILTER) {
        // (A - B) f = kappa * D8 f, D8 as four nested second differences
        static_assert(KIND != OP_FILTER || HW >= 4, "window too narrow");
        constexpr int W = M + 2 * HW;
        real e[W];
#pragma unroll
        for (int j = 0; j < W; j++) e[j] = w[j];
#pragma unroll
        for (int lev = 1; lev <= 4; lev++) {
#pragma unroll
            for (int j = 0; j + 2 * lev < W; j++) e[j] = (e[j] + e[j + 2]) - real(2) * e[j + 1];
        }
        // e[j] is now D8 f at window index j + 4
#pragma unroll
        for (int t = 0; t < M; t++) r[t] = t0 * e[HW - 4 + t];
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

// B f on the segment, boundary rows included.
template <int KIND, typename real, int M, int HW, bool UNI, int MAP>
__device__ __forceinline__ void rhs(const OpDev<real> *__restrict__ op, const Lane<real, MAP> &c,
                                    const real (&w)[M + 2 * HW], real (&r)[M])
{
    stencil<KIND, real, M, HW>(op, w, r);
    if constexpr (!UNI) boundary_rows<KIND, real, M, HW>(op, c.s, c.len, w, r);
}

// ------------------------------------------------------------ recurrences

// forward recurrence y_t = r_t - m_t y_{t-1} with zero carry; returns the
// segment's last value.  Table plans read the lane's constants at its slot
// g * M (plan.hpp) and carry the last value through the points past a
// short segment's end with selects, not branches.
template <typename real, int M, bool UNI, int MAP>
__device__ __forceinline__ real fwd_local(const OpDev<real> *__restrict__ op,
                                          const Lane<real, MAP> &c, real (&r)[M])
{
    if (UNI) {
        const real m = op->um;
#pragma unroll
        for (int t = 1; t < M; t++) r[t] = r[t] - m * r[t - 1];
        return r[M - 1];
    } else {
        const real *__restrict__ tm = op->m + c.g * M;
#pragma unroll
        for (int t = 1; t < M; t++) {
            const real y = r[t] - tm[t] * r[t - 1];
            r[t] = t < c.len ? y : r[t - 1];
        }
        return r[M - 1];
    }
}

// forward fix-up y_t += pf_t C, then the backward recurrence
// x_t = iw_t y_t + mu_t x_{t+1} with zero carry; returns the first value
template <typename real, int M, bool UNI, int MAP>
__device__ __forceinline__ real bwd_local(const OpDev<real> *__restrict__ op,
                                          const Lane<real, MAP> &c, real (&r)[M], real C)
{
    if (UNI) {
        // periodic factorisation: iw = 1
        const real mu = op->umu;
#pragma unroll
        for (int t = 0; t < M; t++) r[t] = r[t] + op->upf[t] * C;
#pragma unroll
        for (int t = M - 2; t >= 0; t--) r[t] = r[t] + mu * r[t + 1];
        return r[0];
    } else {
        const real *__restrict__ tiw = op->iw + c.g * M;
        const real *__restrict__ tmu = op->mu + c.g * M;
        const real *__restrict__ tpf = op->pf + c.g * M;
        real carry = 0;
#pragma unroll
        for (int t = M - 1; t >= 0; t--) {
            const real y = r[t] + tpf[t] * C;
            const real x = tiw[t] * y + tmu[t] * carry;
            carry = t < c.len ? x : real(0);
            r[t] = carry;
        }
        return r[0];
    }
}

// backward fix-up and output scale: out_t = scale (x_t + pb_t D) (points
// past a short segment's end are computed and never stored)
template <typename real, int M, bool UNI, int MAP>
__device__ __forceinline__ void finish(const OpDev<real> *__restrict__ op,
                                       const Lane<real, MAP> &c, real (&r)[M], real D)
{
    const real sc = op->scale;
    if (UNI) {
#pragma unroll
        for (int t = 0; t < M; t++) r[t] = sc * (r[t] + op->upb[t] * D);
    } else {
        const real *__restrict__ tpb = op->pb + c.g * M;
#pragma unroll
        for (int t = 0; t < M; t++) r[t] = sc * (r[t] + tpb[t] * D);
    }
}

// ------------------------------------------------------------ exchange

// Carry weights of one operator for this lane: uniform plans share one row,
// table plans hold one row of Ks per segment.
template <typename real, bool UNI, bool FWD, int MAP>
__device__ __forceinline__ const real *carry_weights(const OpDev<real> *__restrict__ op,
                                                     const Lane<real, MAP> &c)
{
    if (UNI) return FWD ? op->uphi : op->upsi;
    return (FWD ? op->phi : op->psi) + (long)c.g * op->Ks;
}

// Publish v[0..N1+N2) for this lane's segment; after the barrier,
// out[i] = sum_{k=1..K} w[k-1] v_i(segment g -+ k), with the weights and K of
// operator 1 for i < N1 and of operator 2 after that.  The values of one
// ring position sit side by side at an odd stride S (entry (q, i) at
// q * S + i), so every read of a lane is at a constant offset from one
// address.  Both plan kinds store zero weights past K (at least KC of them),
// so on lines of at least KR >= KC segments the first KC terms run unrolled
// without branches; terms past KC and shorter lines run in a plain loop.
template <int N1, int N2, bool FWD, int KC1, int KC2, bool UNI, typename real, int MAP>
__device__ __forceinline__ void exchange(Lane<real, MAP> &c, const real (&v)[N1 + N2],
                                         real (&out)[N1 + N2], const real *__restrict__ w1,
                                         int K1, const real *__restrict__ w2, int K2, int cyc)
{
    constexpr int N = N1 + N2, S = N | 1;
    const int row = ring_row(c.P, c.KR, c.LX, S, MAP == MAP_COL, (int)sizeof(real));
    real *b = c.xb + c.par * c.cap;
    c.par ^= 1;
    const int prim = (c.line * row + (FWD ? c.KR + c.g : c.g)) * S;
    const bool copy = FWD ? c.g >= c.P - c.KR : c.g < c.KR;
    const int cpos = FWD ? prim - c.P * S : prim + c.P * S;
#pragma unroll
    for (int i = 0; i < N; i++) {
        b[prim + i] = v[i];
        if (copy) b[cpos + i] = cyc ? v[i] : real(0);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < N; i++) out[i] = 0;
    constexpr int dir = FWD ? -S : S;
    const real *p = b + prim;
    // the unrolled reads reach KC1 (KC2) positions back/forward; every one
    // of them was written only when the line has at least KR >= KC segments
    const bool unrolled = c.P >= c.KR;
    const int k1 = unrolled ? KC1 : 0, k2 = unrolled ? KC2 : 0;
    if (unrolled) {
#pragma unroll
        for (int k = 1; k <= KC1; k++) {
            const real wk = w1[k - 1];
#pragma unroll
            for (int i = 0; i < N1; i++) out[i] += wk * p[dir * k + i];
        }
        if constexpr (N2 > 0) {
#pragma unroll
            for (int k = 1; k <= KC2; k++) {
                const real wk = w2[k - 1];
#pragma unroll
                for (int i = N1; i < N; i++) out[i] += wk * p[dir * k + i];
            }
        }
    }
    for (int k = k1 + 1; k <= K1; k++) {
        const real wk = w1[k - 1];
#pragma unroll
        for (int i = 0; i < N1; i++) out[i] += wk * p[dir * k + i];
    }
    if constexpr (N2 > 0) {
        for (int k = k2 + 1; k <= K2; k++) {
            const real wk = w2[k - 1];
#pragma unroll
            for (int i = N1; i < N; i++) out[i] += wk * p[dir * k + i];
        }
    }
}

// N1 solves with operator 1 and N2 with operator 2, r holding B f on entry
// and A^-1 B f on exit; two barriers for the whole batch.  KC1, KC2: carry
// terms unrolled on uniform plans for each operator (the lane's ring margin
// must be at least both).
template <int N1, int N2, int KC1, int KC2, typename real, int M, bool UNI, int MAP>
__device__ __forceinline__ void solve_batch2(const OpDev<real> *__restrict__ op1,
                                             const OpDev<real> *__restrict__ op2,
                                             Lane<real, MAP> &c, real (&r)[N1 + N2][M])
{
    constexpr int N = N1 + N2;
    real e[N], C[N];
#pragma unroll
    for (int i = 0; i < N; i++)
        e[i] = fwd_local<real, M, UNI, MAP>(i < N1 ? op1 : op2, c, r[i]);
    exchange<N1, N2, true, KC1, KC2, UNI>(c, e, C, carry_weights<real, UNI, true>(op1, c), op1->K,
                           carry_weights<real, UNI, true>(op2, c), op2->K, op1->cyclic);
#pragma unroll
    for (int i = 0; i < N; i++)
        e[i] = bwd_local<real, M, UNI, MAP>(i < N1 ? op1 : op2, c, r[i], C[i]);
    exchange<N1, N2, false, KC1, KC2, UNI>(c, e, C, carry_weights<real, UNI, false>(op1, c), op1->K,
                            carry_weights<real, UNI, false>(op2, c), op2->K, op1->cyclic);
#pragma unroll
    for (int i = 0; i < N; i++) finish<real, M, UNI, MAP>(i < N1 ? op1 : op2, c, r[i], C[i]);
}

// ------------------------------------------------------------ wavefront carries
//
// Lines of exactly 64 uniform periodic segments with MAP_ROW lanes: one line
// is one wavefront, segment g on lane g.  The forward carry into segment g,
// sum_{k>=1} uphi[k-1] e_{g-k} (cyclic), is then the cyclic weighted scan
// sum_{i>=0} lam^i x_{g-i} of x_g = e_{g-1}: a weighted inclusive scan inside
// each row of 16 lanes (four DPP row shifts, zero past the row's edge), the
// four row totals (readlane), and each row total once per pass round the
// line, weighted lam^(distance) / (1 - lam^64).  Backward the same mirrored
// (x_g = b_{g+1}, row shifts to the left, totals at each row's first lane).
// No LDS, no barrier, and the sum is not truncated.

template <int CTRL>
__device__ __forceinline__ float dpp_move(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float read_lane(float x, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double read_lane(double x, int l)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// this lane's weights of the four row totals, forward (w[0]) and backward
template <typename real>
__device__ __forceinline__ void wave_row_weights(const OpDev<real> *__restrict__ op, int lane,
                                                 real (&w)[2][4])
{
    const int r = lane >> 4, j = lane & 15;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        w[0][q] = op->wrow[0][j + 1 + 16 * ((r - q - 1) & 3)];
        w[1][q] = op->wrow[1][16 - j + 16 * ((q - r - 1) & 3)];
    }
}

template <bool FWD, typename real>
__device__ __forceinline__ real wave_carry(real v, const real *__restrict__ wl,
                                           const real (&wq)[4], int lane)
{
    real x = __shfl(v, FWD ? (lane + 63) & 63 : (lane + 1) & 63, 64);
    constexpr int B = FWD ? 0x110 : 0x100;  // DPP row_shr:n / row_shl:n
    x += wl[0] * dpp_move<B + 1>(x);
    x += wl[1] * dpp_move<B + 2>(x);
    x += wl[2] * dpp_move<B + 4>(x);
    x += wl[3] * dpp_move<B + 8>(x);
    real out = x;
#pragma unroll
    for (int q = 0; q < 4; q++) out += wq[q] * read_lane(x, FWD ? 16 * q + 15 : 16 * q);
    return out;
}

// solve_batch2 on lines of 64 uniform periodic segments, one per wavefront
// (every lane of the wavefront must take part)
template <int N1, int N2, typename real, int M, int MAP>
__device__ __forceinline__ void solve_batch_wave(const OpDev<real> *__restrict__ op1,
                                                 const OpDev<real> *__restrict__ op2,
                                                 Lane<real, MAP> &c, real (&r)[N1 + N2][M])
{
    static_assert(MAP == MAP_ROW, "segment g must sit on lane g");
    constexpr int N = N1 + N2;
    const int lane = threadIdx.x & 63;
    real w1[2][4], w2[2][4];
    wave_row_weights(op1, lane, w1);
    if constexpr (N2 > 0) wave_row_weights(op2, lane, w2);
#pragma unroll
    for (int i = 0; i < N; i++) {
        const OpDev<real> *op = i < N1 ? op1 : op2;
        const real e = fwd_local<real, M, true, MAP>(op, c, r[i]);
        const real C = wave_carry<true>(e, op->wlam[0], i < N1 ? w1[0] : w2[0], lane);
        const real b = bwd_local<real, M, true, MAP>(op, c, r[i], C);
        const real D = wave_carry<false>(b, op->wlam[1], i < N1 ? w1[1] : w2[1], lane);
        finish<real, M, true, MAP>(op, c, r[i], D);
    }
}

template <int N, int KC, typename real, int M, bool UNI, int MAP>
__device__ __forceinline__ void solve_batch(const OpDev<real> *__restrict__ op,
                                            Lane<real, MAP> &c, real (&r)[N][M])
{
    solve_batch2<N, 0, KC, KC, real, M, UNI, MAP>(op, op, c, r);
}

}  // namespace cs3d
