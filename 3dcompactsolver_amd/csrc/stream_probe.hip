// stream_probe.hip -- the HBM rate this GPU sustains on plain streams,
// measured with hand-written kernels: a copy (b = a) and a triad
// (a = b + s c) over 16-byte vectors.  No reference counterpart: it gives
// the sweeps' roofline fractions a measured ceiling beside the 8 TB/s peak,
// and a copy of known size calibrates the HBM traffic counters.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/cs3d.h"
#include "capi_common.hpp"

namespace cs3d {

typedef float f4 __attribute__((ext_vector_type(4)));

// each thread moves U vectors per pass, all loads issued before any store,
// passes grid-strided; n4 vectors per array
template <int U>
__global__ __launch_bounds__(256) void k_copy(const f4 *__restrict__ a, f4 *__restrict__ b,
                                              long n4)
{
    const long stride = (long)gridDim.x * blockDim.x;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (U - 1) * stride < n4; i += U * stride) {
        f4 v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = a[i + u * stride];
#pragma unroll
        for (int u = 0; u < U; u++) __builtin_nontemporal_store(v[u], &b[i + u * stride]);
    }
    for (; i < n4; i += stride) b[i] = a[i];
}

template <int U>
__global__ __launch_bounds__(256) void k_triad(f4 *__restrict__ a, const f4 *__restrict__ b,
                                               const f4 *__restrict__ c, float s, long n4)
{
    const long stride = (long)gridDim.x * blockDim.x;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (U - 1) * stride < n4; i += U * stride) {
        f4 x[U], y[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            x[u] = b[i + u * stride];
            y[u] = c[i + u * stride];
        }
#pragma unroll
        for (int u = 0; u < U; u++) __builtin_nontemporal_store(x[u] + s * y[u], &a[i + u * stride]);
    }
    for (; i < n4; i += stride) a[i] = b[i] + s * c[i];
}

__global__ void k_fill(f4 *__restrict__ a, long n4, float v)
{
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
        a[i] = f4{v, v, v, v};
}

}  // namespace cs3d

using namespace cs3d;

extern "C" int cs3d_stream_probe(int kind, size_t bytes, int reps, double *best_gbps,
                                 double *mean_gbps)
{
    if ((kind != 0 && kind != 1) || bytes < (1u << 20) || bytes % 16 || reps < 1)
        return set_error(CS3D_ERR_ARG, "cs3d_stream_probe: kind 0/1, bytes >= 1 MiB and a "
                                       "multiple of 16, reps >= 1");
    const long n4 = (long)(bytes / 16);
    const int narr = kind == 0 ? 2 : 3;
    f4 *buf[3] = {nullptr, nullptr, nullptr};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::string err;
    auto ok = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && err.empty()) err = std::string(what) + ": " + hipGetErrorString(e);
        return e == hipSuccess;
    };
    int dev = 0, cus = 256;
    if (ok(hipGetDevice(&dev), "hipGetDevice"))
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    // 8 workgroups of 256 lanes per CU, grid-strided: every CU keeps 32
    // waves of 16-byte loads in flight
    const dim3 grid((unsigned)cus * 8), block(256);
    double best = 0.0, sum = 0.0;
    for (int a = 0; a < narr && err.empty(); a++)
        if (ok(hipMalloc(&buf[a], bytes), "hipMalloc"))
            hipLaunchKernelGGL(k_fill, grid, block, 0, nullptr, buf[a], n4, 1.0f + a);
    if (err.empty() && ok(hipEventCreate(&e0), "hipEventCreate") &&
        ok(hipEventCreate(&e1), "hipEventCreate")) {
        const double moved = (double)bytes * narr;
        for (int r = -2; r < reps && err.empty(); r++) {
            ok(hipEventRecord(e0, nullptr), "hipEventRecord");
            if (kind == 0)
                hipLaunchKernelGGL((k_copy<4>), grid, block, 0, nullptr, buf[0], buf[1], n4);
            else
                hipLaunchKernelGGL((k_triad<4>), grid, block, 0, nullptr, buf[0], buf[1], buf[2],
                                   0.5f, n4);
            ok(hipGetLastError(), "launch");
            ok(hipEventRecord(e1, nullptr), "hipEventRecord");
            float ms = 0.0f;
            if (ok(hipEventSynchronize(e1), "hipEventSynchronize") &&
                ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime") && r >= 0) {
                const double gbps = moved / (ms * 1e-3) / 1e9;
                best = gbps > best ? gbps : best;
                sum += gbps;
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (int a = 0; a < narr; a++)
        if (buf[a]) (void)hipFree(buf[a]);
    if (!err.empty()) return set_error(CS3D_ERR_HIP, "cs3d_stream_probe: " + err);
    if (best_gbps) *best_gbps = best;
    if (mean_gbps) *mean_gbps = sum / reps;
    return CS3D_OK;
}
