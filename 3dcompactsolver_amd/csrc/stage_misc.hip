// stage_misc.hip -- the step-level kernels around the stage sweeps: the
// filter sweeps, the FP64 reductions (CFL maximum, diagnostics), the dt
// update and the sponge update; and the sweep dispatcher.
#include "stage_common.hpp"

namespace cs3d {

// -------------------------------------------------------------- filter
// One filter sweep over the five conserved fields (Filter.cpp:129-171 as
// called by filterConservedData, CSolver_AWS.cpp:1682-1893).

// workgroups of at most 512 lanes: the five fields of a segment stay in
// registers through the batched solve
// FP32 column filters fit 1024 lanes (32 lines of 32 segments: 128-byte
// rows); FP64 keeps 512 for its registers
// (the compensated variants, with the low parts and FP64 sums in
// registers, need more than 128 VGPRs: 512 lanes)
template <typename real, bool CMP = false>
constexpr int filter_threads() { return (sizeof(real) == 4 && !CMP) ? 1024 : 512; }

constexpr int RED_THREADS = 256;
constexpr int RED_SLOTS = RED_SLOTS_HOST;

// the CFL wave-speed sum of calcDtFromCFL (CSolver_AWS.cpp:513-525) at one
// point of the state
template <typename real>
__device__ __forceinline__ double cfl_speed(const Phys<real> &ph, real rp, real ru, real rv,
                                            real rw, real re)
{
    real irho, U, V, W, pp, Tp;
    prim(ph, rp, ru, rv, rw, re, irho, U, V, W, pp, Tp);
    const double rho = (double)ph.rho_ref + (double)rp;
    const double p = (double)ph.p_ref + (double)pp;
    const double cs = sqrt((double)ph.gamma * p / rho);
    return (fabs((double)U) + cs) * (double)ph.idx + (fabs((double)V) + cs) * (double)ph.idy +
           (fabs((double)W) + cs) * (double)ph.idz;
}

// RED: the last sweep of the filter also leaves, per workgroup, the CFL
// maximum and the NaN count of its output (two doubles, for k_reduce_cfl)
// SET: the five fields are a conserved set (quads + rhoE'), else separate
// CMP: compensated FP32 mode: the correction A^-1 (A - B) f is formed from
// the high parts and subtracted from hi + lo in FP64
template <typename real, int M, bool UNI, int MAP, bool RED, bool SET, bool CMP = false>
__global__ __launch_bounds__((filter_threads<real, CMP>())) void k_filter5(const OpDev<real> *__restrict__ fo, LineMap lm,
                                                  FiltIO<real> io, Phys<real> ph,
                                                  double *__restrict__ part)
{
    constexpr int HW = 4, W = M + 2 * HW;
    CS3D_LANE_SETUP(MAP, fo, fo->K > kc_filter<real>() ? fo->K : kc_filter<real>(), 5)
    real r[5][M], fc[5][M];
    // B f of one window, and the segment's own values, for
    // out = f - A^-1 (A - B) f (held in registers: read again after the
    // solve they come from HBM)
    auto one = [&](int v, const real(&w)[W]) {
        rhs<OP_FILTER, real, M, HW, UNI>(fo, c, w, r[v]);
#pragma unroll
        for (int t = 0; t < M; t++) fc[v][t] = w[HW + t];
    };
    if constexpr (SET) {
        real w0[W], w1[W], w2[W], w3[W], w4[W];
        win.load_quad(io.qin.m, w0, w1, w2, w3);
        one(0, w0);
        one(1, w1);
        one(2, w2);
        one(3, w3);
        win.load(io.qin.e, w4);
        one(4, w4);
    } else {
#pragma unroll
        for (int v = 0; v < 5; v++) {
            real w[W];
            win.load(io.sep.in[v], w);
            one(v, w);
        }
    }
    // table plans unroll fewer carry terms (all of the filter's would not
    // fit the registers of a 1024-lane workgroup) and loop over the rest
    solve_batch<5, UNI ? kc_filter<real>() : kc_deriv<real>(), real, M, UNI>(fo, c, r);
    static_assert(!CMP || (SET && sizeof(real) == 4), "compensated sets are FP32 conserved sets");
    if constexpr (CMP) {
        if (active) {
            // the correction folded into hi + lo here; r is spent
            real lo[5][M];
            win.load_seg_quad(io.qin_lo.m, lo[0], lo[1], lo[2], lo[3]);
            win.load_seg(io.qin_lo.e, lo[4]);
#pragma unroll
            for (int v = 0; v < 5; v++)
#pragma unroll
                for (int t = 0; t < M; t++) {
                    split_part(two_part(fc[v][t], lo[v][t]) - (double)r[v][t], fc[v][t], lo[v][t]);
                    r[v][t] = real(0);
                }
            win.store_quad(io.qout_lo.m, lo[0], lo[1], lo[2], lo[3], c.len);
            win.store(io.qout_lo.e, lo[4], c.len);
        }
    }
    if (active) {
#pragma unroll
        for (int v = 0; v < 5; v++)
#pragma unroll
            for (int t = 0; t < M; t++) fc[v][t] -= r[v][t];
        if constexpr (SET) {
            win.store_quad(io.qout.m, fc[0], fc[1], fc[2], fc[3], c.len);
            win.store(io.qout.e, fc[4], c.len);
        } else {
#pragma unroll
            for (int v = 0; v < 5; v++) win.store(io.sep.out[v], fc[v], c.len);
        }
    }
    if constexpr (RED) {
        double umax = 0.0, nan = 0.0;
        if (active) {
#pragma unroll
            for (int t = 0; t < M; t++) {
                if (!UNI && t >= c.len) continue;
                const double u = cfl_speed(ph, fc[0][t], fc[1][t], fc[2][t], fc[3][t], fc[4][t]);
                if (u != u)
                    nan += 1.0;
                else
                    umax = fmax(umax, u);
            }
        }
        // tree over the workgroup in the exchange ring's LDS, free once every
        // lane is past the solve (the workgroup size LX * P need not be a
        // multiple of the wavefront; the host checks the ring holds
        // 2 * blockDim doubles)
        extern __shared__ double red[];
        const int nt = blockDim.x, tid = threadIdx.x;
        if (nt % 64 == 0) {
            // whole wavefronts: across each wavefront by lane shuffles (no
            // barrier), then one value per wavefront through LDS
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                umax = fmax(umax, __shfl_xor(umax, o));
                nan += __shfl_xor(nan, o);
            }
            const int nw = nt / 64;
            __syncthreads();
            if ((tid & 63) == 0) {
                red[tid / 64] = umax;
                red[nw + tid / 64] = nan;
            }
            __syncthreads();
            if (tid == 0) {
                double a = red[0], b = red[nw];
                for (int i = 1; i < nw; i++) {
                    a = fmax(a, red[i]);
                    b += red[nw + i];
                }
                part[2 * (long)blockIdx.x] = a;
                part[2 * (long)blockIdx.x + 1] = b;
            }
            return;
        }
        __syncthreads();
        red[tid] = umax;
        red[nt + tid] = nan;
        __syncthreads();
        int h = 1;
        while (2 * h < nt) h *= 2;
        for (; h > 0; h /= 2) {
            if (tid < h && tid + h < nt) {
                red[tid] = fmax(red[tid], red[tid + h]);
                red[nt + tid] += red[nt + tid + h];
            }
            __syncthreads();
        }
        if (tid == 0) {
            part[2 * (long)blockIdx.x] = red[0];
            part[2 * (long)blockIdx.x + 1] = red[nt];
        }
    }
}

// -------------------------------------------------------------- reductions
// Per-workgroup FP64 partial results over the state after the step:
// [0] max of the CFL wave-speed sum (calcDtFromCFL, CSolver_AWS.cpp:513-525)
// and, when diag is set, the sums of calcTaylorGreenQuantities
// (CSolver_AWS.cpp:2188-2228): [1] sum rho|u|^2/2 [2] sum 2 mu S:grad u
// [3] sum rho|w|^2/2 [4] sum mu rho |w|^2 [5] sum div^2 [6] sum mu; [7]
// counts NaNs; and for calcTurbulenceQuantities (CSolver_AWS.cpp:2040-2095)
// [8] sum rho' [9] sum rho'^2 [10] sum |u|^2 [11] sum (Ux^2+Vy^2+Wz^2)/3
// [12] sum of the sound speed [13] sum |w|^2.

template <typename real>
__global__ __launch_bounds__(RED_THREADS) void k_reduce_state(Fields<real> F, Phys<real> ph,
                                                               long N, int diag,
                                                               double *__restrict__ part)
{
    __shared__ double sh[RED_SLOTS][RED_THREADS];
    double acc[RED_SLOTS] = {};
    for (long i = (long)blockIdx.x * RED_THREADS + threadIdx.x; i < N;
         i += (long)gridDim.x * RED_THREADS) {
        real irho, U, V, W, pp, Tp, q5[5];
        qset_load(F.q, i, q5);
        prim(ph, q5[0], q5[1], q5[2], q5[3], q5[4], irho, U, V, W, pp, Tp);
        double rho = (double)ph.rho_ref + (double)q5[0];
        double p = (double)ph.p_ref + (double)pp;
        double cs = sqrt((double)ph.gamma * p / rho);
        double u = (fabs((double)U) + cs) * (double)ph.idx +
                   (fabs((double)V) + cs) * (double)ph.idy +
                   (fabs((double)W) + cs) * (double)ph.idz;
        if (u != u)
            acc[7] += 1.0;  // NaN met: counted, reported as non-finite
        else
            acc[0] = fmax(acc[0], u);
        if (diag) {
            // diag 2: kinetic energy and mean viscosity only (no stage inputs)
            real mu, amu;
            visc(ph, ph.T0s + Tp, mu, amu);
            double m = mu;
            double uv = (double)U * U + (double)V * V + (double)W * W;
            acc[1] += rho * uv / 2.0;
            acc[6] += m;
            if (diag == 1) {
                double td = F.diag[0][i], w2 = F.diag[1][i], dv = F.diag[2][i];
                acc[2] += 2.0 * m * td;
                acc[3] += 0.5 * rho * w2;
                acc[4] += m * rho * w2;
                acc[5] += dv * dv;
                const double rp = q5[0];
                acc[8] += rp;
                acc[9] += rp * rp;
                acc[10] += uv;
                acc[11] += (double)F.diag[3][i];
                acc[12] += cs;
                acc[13] += w2;
            }
        }
    }
    for (int s = 0; s < RED_SLOTS; s++) sh[s][threadIdx.x] = acc[s];
    __syncthreads();
    for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh[0][threadIdx.x] = fmax(sh[0][threadIdx.x], sh[0][threadIdx.x + w]);
            for (int s = 1; s < RED_SLOTS; s++) sh[s][threadIdx.x] += sh[s][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < RED_SLOTS) part[(long)blockIdx.x * RED_SLOTS + threadIdx.x] = sh[threadIdx.x][0];
}

// fixed-order reduction of the partials into out[RED_SLOTS]
__global__ __launch_bounds__(RED_THREADS) void k_reduce_final(const double *__restrict__ part,
                                                               int nb, double *__restrict__ out,
                                                               StepScalars *sc)
{
    __shared__ double sh[RED_SLOTS][RED_THREADS];
    double acc[RED_SLOTS] = {};
    for (int b = threadIdx.x; b < nb; b += RED_THREADS) {
        acc[0] = fmax(acc[0], part[(long)b * RED_SLOTS]);
        for (int s = 1; s < RED_SLOTS; s++) acc[s] += part[(long)b * RED_SLOTS + s];
    }
    for (int s = 0; s < RED_SLOTS; s++) sh[s][threadIdx.x] = acc[s];
    __syncthreads();
    for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh[0][threadIdx.x] = fmax(sh[0][threadIdx.x], sh[0][threadIdx.x + w]);
            for (int s = 1; s < RED_SLOTS; s++) sh[s][threadIdx.x] += sh[s][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int s = 0; s < RED_SLOTS; s++) out[s] = sh[s][0];
        sc->cfl_max = sh[0][0];
        if (sh[7][0] > 0.0 || sh[0][0] > 1e300) sc->nonfinite = 1;
    }
}

// The filter's CFL partials (maximum, NaN count per workgroup) into the
// k_reduce_final outputs: slot 0 the maximum, slot 7 the NaN count, the
// other slots zero.
constexpr int CFL_THREADS = 1024;
__global__ __launch_bounds__(CFL_THREADS) void k_reduce_cfl(const double *__restrict__ part, int nb,
                                                            double *__restrict__ out,
                                                            StepScalars *sc)
{
    __shared__ double sm[2][CFL_THREADS];
    double mx = 0.0, nan = 0.0;
    for (int b = threadIdx.x; b < nb; b += CFL_THREADS) {
        mx = fmax(mx, part[2 * (long)b]);
        nan += part[2 * (long)b + 1];
    }
    sm[0][threadIdx.x] = mx;
    sm[1][threadIdx.x] = nan;
    __syncthreads();
    for (int w = CFL_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sm[0][threadIdx.x] = fmax(sm[0][threadIdx.x], sm[0][threadIdx.x + w]);
            sm[1][threadIdx.x] += sm[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if ((int)threadIdx.x < RED_SLOTS)
        out[threadIdx.x] = threadIdx.x == 0 ? sm[0][0] : threadIdx.x == 7 ? sm[1][0] : 0.0;
    if (threadIdx.x == 0) {
        sc->cfl_max = sm[0][0];
        if (sm[1][0] > 0.0 || sm[0][0] > 1e300) sc->nonfinite = 1;
    }
}

// calcDtFromCFL (CSolver_AWS.cpp:530-543), on the device so that the time
// loop never waits for the host.  Q1: the CFL branch is the one taken.
__global__ void k_dt(StepScalars *sc, double cfl, double dt_fixed)
{
    double dt = dt_fixed > 0.0 ? dt_fixed : cfl / sc->cfl_max;
    if (!(dt > 0.0) || dt != dt || dt > 1e300) sc->nonfinite = 1;
    sc->dt = dt;
    if (sc->timeStep == 0) {
        sc->timeStep++;
        sc->time = 0.0;
    } else {
        sc->timeStep++;
        sc->time += dt;
    }
}

// updateSponge (CSolver_AWS.cpp:1937-2031): running averages, then the
// sponge face planes of the state are reset to them.
template <typename real>
__global__ void k_sponge(Fields<real> F, Phys<real> ph, StageArgs sa, const StepScalars *sc,
                         double avgT, double epsP, double spongeP)
{
    const long N = (long)sa.Nx * sa.Ny * sa.Nz;
    const double eps = 1.0 / (avgT / sc->dt + 1.0);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        real a[5], qv[5];
        qset_load(F.avg, i, a);
        qset_load(F.q, i, qv);
#pragma unroll
        for (int v = 0; v < 5; v++) a[v] += (real)eps * (qv[v] - a[v]);
        const real rho = ph.rho_ref + a[0];
        const real target = (real)(spongeP / (double)ph.gm1) - ph.rhoE_ref +
                            real(0.5) * (a[1] * a[1] + a[2] * a[2] + a[3] * a[3]) / rho;
        a[4] = (real)epsP * a[4] + (real)(1.0 - epsP) * target;
        qset_store(F.avg, i, a);
        const int ii = (int)(i % sa.Nx), jj = (int)((i / sa.Nx) % sa.Ny),
                  kk = (int)(i / ((long)sa.Nx * sa.Ny));
        const bool plane = (sa.sponge_face[0] && ii == 0) || (sa.sponge_face[1] && ii == sa.Nx - 1) ||
                           (sa.sponge_face[2] && jj == 0) || (sa.sponge_face[3] && jj == sa.Ny - 1) ||
                           (sa.sponge_face[4] && kk == 0) || (sa.sponge_face[5] && kk == sa.Nz - 1);
        if (plane) {
            qset_store(F.qstale, i, qv);
            qset_store(F.q, i, a);
            if (F.qlo.m) {
                const real z[5] = {0, 0, 0, 0, 0};
                qset_store(F.qlo, i, z);
            }
        }
    }
}

// ---------------------------------------------------------- host launchers

template <typename real>
hipError_t launch_sweep(int which, const OpDev<real> *d1, const OpDev<real> *d2,
                        const OpDev<real> *fx, int uni, int M,
                        const LaunchGeom &lg, const LineMap &lm, const Fields<real> &F,
                        const Phys<real> &ph, const StageArgs &sa, const StepScalars *sc,
                        hipStream_t st)
{
    if (M != SEG_DERIV && !(M == SEG_DERIV3 && which == 2 && uni)) return hipErrorInvalidValue;
    if (lg.LX < 1 || lg.threads < 1 || lg.threads > sweep_threads<real>())
        return hipErrorInvalidConfiguration;
    int use_k = sa.rkStep > 1;
    // sweep order of a stage: y (first), z (middle), x (last, with the
    // pointwise assembly and the stage update)
    if (which == 0)
        return uni ? launch_xpass<real, true>(d1, d2, fx, lg, lm, F, ph, sa, sc, st)
                   : launch_xpass<real, false>(d1, d2, fx, lg, lm, F, ph, sa, sc, st);
    if (which == 1)
        return uni ? launch_ypass<real, true>(d1, d2, lg, lm, F, ph, use_k, sa, st)
                   : launch_ypass<real, false>(d1, d2, lg, lm, F, ph, use_k, sa, st);
    return uni ? launch_zpass<real, true>(d1, d2, M, lg, lm, F, ph, use_k, sa, st)
               : launch_zpass<real, false>(d1, d2, M, lg, lm, F, ph, use_k, sa, st);
}

template <typename real, bool RED, bool SET, bool CMP = false>
static hipError_t launch_f5(const OpDev<real> *fo, int uni, int M, const LaunchGeom &lg, int map,
                            const LineMap &lm, const FiltIO<real> &io, const Phys<real> &ph,
                            double *part, hipStream_t st)
{
    dim3 grid((lm.nlines + lg.LX - 1) / lg.LX), block(lg.threads);
    const size_t shm = lg.shm;
    if (uni) {
        if (M != SEG_FILTER) return hipErrorInvalidValue;
        if (map == MAP_ROW)
            hipLaunchKernelGGL((k_filter5<real, SEG_FILTER, true, MAP_ROW, RED, SET, CMP>), grid, block, shm, st, fo, lm, io, ph, part);
        else
            hipLaunchKernelGGL((k_filter5<real, SEG_FILTER, true, MAP_COL, RED, SET, CMP>), grid, block, shm, st, fo, lm, io, ph, part);
    } else {
        if (M != SEG_TABLE) return hipErrorInvalidValue;
        if (map == MAP_ROW)
            hipLaunchKernelGGL((k_filter5<real, SEG_TABLE, false, MAP_ROW, RED, SET, CMP>), grid, block, shm, st, fo, lm, io, ph, part);
        else
            hipLaunchKernelGGL((k_filter5<real, SEG_TABLE, false, MAP_COL, RED, SET, CMP>), grid, block, shm, st, fo, lm, io, ph, part);
    }
    return hipGetLastError();
}

template <typename real>
hipError_t launch_filter5(const OpDev<real> *fo, int uni, int M, const LaunchGeom &lg, int map,
                          const LineMap &lm, const FiltIO<real> &io, const Phys<real> &ph,
                          double *part, hipStream_t st)
{
    if (lg.LX < 1 || lg.threads < 1 || lg.threads > filter_threads<real>())
        return hipErrorInvalidConfiguration;
    // the CFL partials come from the state: a conserved set only
    if (part && !io.set) return hipErrorInvalidValue;
    if (!io.set) return launch_f5<real, false, false>(fo, uni, M, lg, map, lm, io, ph, part, st);
    if (io.qin_lo.m) {
        if (lg.threads > filter_threads<real, true>()) return hipErrorInvalidConfiguration;
        if constexpr (sizeof(real) == 4)
            return part ? launch_f5<real, true, true, true>(fo, uni, M, lg, map, lm, io, ph, part, st)
                        : launch_f5<real, false, true, true>(fo, uni, M, lg, map, lm, io, ph, part, st);
        return hipErrorInvalidValue;
    }
    return part ? launch_f5<real, true, true>(fo, uni, M, lg, map, lm, io, ph, part, st)
                : launch_f5<real, false, true>(fo, uni, M, lg, map, lm, io, ph, part, st);
}

hipError_t launch_reduce_cfl(const double *part, int nb, double *out, StepScalars *sc,
                             hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce_cfl, dim3(1), dim3(CFL_THREADS), 0, st, part, nb, out, sc);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_reduce(const Fields<real> &F, const Phys<real> &ph, long N, int diag,
                         double *part, int nb, double *out, StepScalars *sc, hipStream_t st)
{
    hipLaunchKernelGGL((k_reduce_state<real>), dim3(nb), dim3(RED_THREADS), 0, st, F, ph, N, diag,
                       part);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_THREADS), 0, st, part, nb, out, sc);
    return hipGetLastError();
}

hipError_t launch_dt(StepScalars *sc, double cfl, double dt_fixed, hipStream_t st)
{
    hipLaunchKernelGGL(k_dt, dim3(1), dim3(1), 0, st, sc, cfl, dt_fixed);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_sponge(const Fields<real> &F, const Phys<real> &ph, const StageArgs &sa,
                         const StepScalars *sc, double avgT, double epsP, double spongeP,
                         hipStream_t st)
{
    long N = (long)sa.Nx * sa.Ny * sa.Nz;
    int nb = (int)((N + 255) / 256);
    if (nb > 65535) nb = 65535;
    hipLaunchKernelGGL((k_sponge<real>), dim3(nb), dim3(256), 0, st, F, ph, sa, sc, avgT, epsP,
                       spongeP);
    return hipGetLastError();
}

#define CS3D_INST(real)                                                                          \
    template hipError_t launch_sweep<real>(int, const OpDev<real> *, const OpDev<real> *,         \
                                           const OpDev<real> *, int, int,                        \
                                           const LaunchGeom &, const LineMap &,                 \
                                           const Fields<real> &, const Phys<real> &,            \
                                           const StageArgs &, const StepScalars *, hipStream_t); \
    template hipError_t launch_filter5<real>(const OpDev<real> *, int, int, const LaunchGeom &,  \
                                             int, const LineMap &, const FiltIO<real> &,       \
                                             const Phys<real> &, double *, hipStream_t);       \
    template hipError_t launch_reduce<real>(const Fields<real> &, const Phys<real> &, long, int, \
                                            double *, int, double *, StepScalars *, hipStream_t); \
    template hipError_t launch_sponge<real>(const Fields<real> &, const Phys<real> &,            \
                                            const StageArgs &, const StepScalars *, double,     \
                                            double, double, hipStream_t);
CS3D_INST(float)
CS3D_INST(double)

}  // namespace cs3d
