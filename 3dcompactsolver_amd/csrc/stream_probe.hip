// stream_probe.hip -- the HBM rate this GPU sustains on plain streams,
// measured with hand-written kernels: a copy (b = a), a triad
// (a = b + s c) and the column sweeps' mix (5 arrays read, 9 written per
// element) over 16-byte vectors.  No reference counterpart: it gives
// the sweeps' roofline fractions a measured ceiling beside the 8 TB/s peak,
// and a copy of known size calibrates the HBM traffic counters.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/cs3d.h"
#include "capi_common.hpp"

namespace cs3d {

typedef float f4 __attribute__((ext_vector_type(4)));

// Two access shapes, U vectors per thread in flight, all loads issued
// before any store.  GS: passes grid-strided over the array (a fixed grid);
// BLK: one pass, workgroup b owns the 256 * U vectors from b * 256 * U
// (a grid covering the array).  n4 vectors per array.
template <int U, bool GS>
__device__ __forceinline__ long first_index(long &step)
{
    if (GS) {
        step = (long)gridDim.x * blockDim.x;
        return (long)blockIdx.x * blockDim.x + threadIdx.x;
    }
    step = blockDim.x;
    return (long)blockIdx.x * blockDim.x * U + threadIdx.x;
}

template <int U, bool GS>
__global__ __launch_bounds__(256) void k_copy(const f4 *__restrict__ a, f4 *__restrict__ b,
                                              long n4)
{
    long step;
    long i = first_index<U, GS>(step);
    for (; i + (U - 1) * step < n4; i += U * step) {
        f4 v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = a[i + u * step];
#pragma unroll
        for (int u = 0; u < U; u++) __builtin_nontemporal_store(v[u], &b[i + u * step]);
        if (!GS) return;
    }
    for (; i < n4; i += step) b[i] = a[i];
}

template <int U, bool GS>
__global__ __launch_bounds__(256) void k_triad(f4 *__restrict__ a, const f4 *__restrict__ b,
                                               const f4 *__restrict__ c, float s, long n4)
{
    long step;
    long i = first_index<U, GS>(step);
    for (; i + (U - 1) * step < n4; i += U * step) {
        f4 x[U], y[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            x[u] = b[i + u * step];
            y[u] = c[i + u * step];
        }
#pragma unroll
        for (int u = 0; u < U; u++) __builtin_nontemporal_store(x[u] + s * y[u], &a[i + u * step]);
        if (!GS) return;
    }
    for (; i < n4; i += step) a[i] = b[i] + s * c[i];
}

// the mix of the column sweeps: every element reads 5 arrays and writes 9
constexpr int MIX_R = 5, MIX_W = 9;
struct MixPtrs {
    const f4 *in[MIX_R];
    f4 *out[MIX_W];
};

template <int U, bool GS>
__global__ __launch_bounds__(256) void k_mix(MixPtrs p, long n4)
{
    long step;
    long i = first_index<U, GS>(step);
    for (; i + (U - 1) * step < n4; i += U * step) {
        f4 x[U];
#pragma unroll
        for (int u = 0; u < U; u++) x[u] = p.in[0][i + u * step];
#pragma unroll
        for (int k = 1; k < MIX_R; k++)
#pragma unroll
            for (int u = 0; u < U; u++) x[u] += p.in[k][i + u * step];
#pragma unroll
        for (int k = 0; k < MIX_W; k++)
#pragma unroll
            for (int u = 0; u < U; u++) __builtin_nontemporal_store(x[u] * (float)(k + 1), &p.out[k][i + u * step]);
        if (!GS) return;
    }
    for (; i < n4; i += step) {
        f4 x = p.in[0][i];
        for (int k = 1; k < MIX_R; k++) x += p.in[k][i];
        for (int k = 0; k < MIX_W; k++) p.out[k][i] = x * (float)(k + 1);
    }
}

__global__ void k_fill(f4 *__restrict__ a, long n4, float v)
{
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
        a[i] = f4{v, v, v, v};
}

// the launch shapes tried: grid-strided with 8 or 16 workgroups per CU
// and 4 or 8 vectors in flight, and one-pass with 4, 8 or 16
struct ProbeShape {
    int U;
    bool gs;
    int per_cu;
};
static const ProbeShape PROBE_SHAPES[] = {{4, true, 8}, {8, true, 8}, {4, true, 16},
                                          {4, false, 0}, {8, false, 0}, {16, false, 0}};

template <int U, bool GS>
static void launch_probe(int kind, dim3 grid, f4 *const *buf, long n4)
{
    if (kind == 0) {
        hipLaunchKernelGGL((k_copy<U, GS>), grid, dim3(256), 0, nullptr, buf[0], buf[1], n4);
    } else if (kind == 1) {
        hipLaunchKernelGGL((k_triad<U, GS>), grid, dim3(256), 0, nullptr, buf[0], buf[1], buf[2],
                           0.5f, n4);
    } else {
        MixPtrs p;
        for (int k = 0; k < MIX_R; k++) p.in[k] = buf[k];
        for (int k = 0; k < MIX_W; k++) p.out[k] = buf[MIX_R + k];
        hipLaunchKernelGGL((k_mix<U, GS>), grid, dim3(256), 0, nullptr, p, n4);
    }
}

static void launch_shape(const ProbeShape &p, int kind, int cus, f4 *const *buf, long n4)
{
    const long per = 256L * p.U;
    const dim3 grid(p.gs ? (unsigned)(cus * p.per_cu) : (unsigned)((n4 + per - 1) / per));
    if (p.gs)
        p.U == 8 ? launch_probe<8, true>(kind, grid, buf, n4)
                 : launch_probe<4, true>(kind, grid, buf, n4);
    else
        p.U == 16 ? launch_probe<16, false>(kind, grid, buf, n4)
        : p.U == 8 ? launch_probe<8, false>(kind, grid, buf, n4)
                   : launch_probe<4, false>(kind, grid, buf, n4);
}

}  // namespace cs3d

using namespace cs3d;

extern "C" int cs3d_stream_probe(int kind, size_t bytes, int reps, double *best_gbps,
                                 double *mean_gbps)
{
    if (kind < 0 || kind > 2 || bytes < (1u << 20) || bytes % 16 || reps < 1)
        return set_error(CS3D_ERR_ARG, "cs3d_stream_probe: kind 0/1/2, bytes >= 1 MiB and a "
                                       "multiple of 16, reps >= 1");
    if (bytes / 16 / 256 > 0x7fffffffL)
        return set_error(CS3D_ERR_ARG, "cs3d_stream_probe: array too large");
    const long n4 = (long)(bytes / 16);
    const int narr = kind == 0 ? 2 : kind == 1 ? 3 : MIX_R + MIX_W;
    f4 *buf[MIX_R + MIX_W] = {};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::string err;
    auto ok = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && err.empty()) err = std::string(what) + ": " + hipGetErrorString(e);
        return e == hipSuccess;
    };
    int dev = 0, cus = 256;
    if (ok(hipGetDevice(&dev), "hipGetDevice"))
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    for (int a = 0; a < narr && err.empty(); a++)
        if (ok(hipMalloc(&buf[a], bytes), "hipMalloc"))
            hipLaunchKernelGGL(k_fill, dim3((unsigned)cus * 8), dim3(256), 0, nullptr, buf[a], n4,
                               1.0f + a);
    // every launch shape: 2 untimed launches, then reps timed; the result
    // is the shape with the best single launch, and the mean of its reps
    double best = 0.0, best_mean = 0.0;
    if (err.empty() && ok(hipEventCreate(&e0), "hipEventCreate") &&
        ok(hipEventCreate(&e1), "hipEventCreate")) {
        const double moved = (double)bytes * narr;
        for (const ProbeShape &p : PROBE_SHAPES) {
            double sb = 0.0, sum = 0.0;
            for (int r = -2; r < reps && err.empty(); r++) {
                ok(hipEventRecord(e0, nullptr), "hipEventRecord");
                launch_shape(p, kind, cus, buf, n4);
                ok(hipGetLastError(), "launch");
                ok(hipEventRecord(e1, nullptr), "hipEventRecord");
                float ms = 0.0f;
                if (ok(hipEventSynchronize(e1), "hipEventSynchronize") &&
                    ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime") && r >= 0) {
                    const double gbps = moved / (ms * 1e-3) / 1e9;
                    sb = gbps > sb ? gbps : sb;
                    sum += gbps;
                }
            }
            if (sb > best) {
                best = sb;
                best_mean = sum / reps;
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (int a = 0; a < narr; a++)
        if (buf[a]) (void)hipFree(buf[a]);
    if (!err.empty()) return set_error(CS3D_ERR_HIP, "cs3d_stream_probe: " + err);
    if (best_gbps) *best_gbps = best;
    if (mean_gbps) *mean_gbps = best_mean;
    return CS3D_OK;
}
