// capi_ops.cpp -- C ABI of the standalone compact operators
// (Derivatives / Filter drop-in, include/cs3d.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <stdexcept>
#include <cstdint>
#include <string>

#include "../../include/cs3d.h"
#include "capi_common.hpp"
#include "plan_dev.hpp"

using namespace cs3d;

struct cs3d_ops {
    int Nx, Ny, Nz, dir, precision;
    int periodic = 1;
    double dh = 1.0, alpha_f = 0.0;
    hipStream_t stream = nullptr;
    PlanDev<float> f[6];   // indexed by kind: filter, D1, D2, then their
    PlanDev<double> d[6];  // transposes (built on first use)
    void *dbuf_in = nullptr, *dbuf_out = nullptr;
    size_t nbytes = 0;
};

extern "C" {

int cs3d_ops_create(int Nx, int Ny, int Nz, double Lx, double Ly, double Lz, int bc_type,
                    int dir, double alpha_f, int precision, cs3d_ops **out)
{
    if (!out) return set_error(CS3D_ERR_ARG, "cs3d_ops_create: out is NULL");
    *out = nullptr;
    if (Nx < 1 || Ny < 1 || Nz < 1 || dir < 0 || dir > 2 ||
        (bc_type != CS3D_PERIODIC_SOLVE && bc_type != CS3D_DIRICHLET_SOLVE) ||
        (precision != CS3D_FP64 && precision != CS3D_FP32))
        return set_error(CS3D_ERR_ARG, "cs3d_ops_create: bad argument");
    cs3d_ops *o = new cs3d_ops();
    o->Nx = Nx;
    o->Ny = Ny;
    o->Nz = Nz;
    o->dir = dir;
    o->precision = precision;
    int n = dir == 0 ? Nx : dir == 1 ? Ny : Nz;
    double L = dir == 0 ? Lx : dir == 1 ? Ly : Lz;
    // Domain.hpp:30-45: x_i = i*L/(N-1), dx = x_1 - x_0
    double dh = (1.0 / ((double)n - 1.0)) * L;
    int periodic = bc_type == CS3D_PERIODIC_SOLVE;
    o->periodic = periodic;
    o->dh = dh;
    o->alpha_f = alpha_f;
    try {
        for (int kind = 0; kind < 3; kind++) {
            // filter plans only need alpha_f; skip them when it is unusable
            if (kind == OP_FILTER && !(alpha_f > -0.5 && alpha_f < 0.5)) continue;
            if (precision == CS3D_FP32)
                make_plan(o->f[kind], kind, periodic, n, dh, alpha_f);
            else
                make_plan(o->d[kind], kind, periodic, n, dh, alpha_f);
        }
        hipError_t e = hipStreamCreateWithFlags(&o->stream, hipStreamNonBlocking);
        if (e != hipSuccess) throw std::runtime_error(hip_msg(e, "hipStreamCreate"));
    } catch (const std::invalid_argument &ex) {
        cs3d_ops_destroy(o);
        return set_error(CS3D_ERR_ARG, std::string("cs3d_ops_create: ") + ex.what());
    } catch (const std::exception &ex) {
        cs3d_ops_destroy(o);
        return set_error(CS3D_ERR_HIP, std::string("cs3d_ops_create: ") + ex.what());
    }
    *out = o;
    return CS3D_OK;
}

int cs3d_ops_apply(cs3d_ops *o, int kind, int layout, const void *in, void *out, int ptr_kind)
{
    if (!o || !in || !out) return set_error(CS3D_ERR_ARG, "cs3d_ops_apply: NULL argument");
    if (kind < 0 || kind > 5) return set_error(CS3D_ERR_ARG, "cs3d_ops_apply: bad kind");
    bool fp32 = o->precision == CS3D_FP32;
    const bool transposed = kind >= CS3D_FILTER_T;
    if (transposed && in == out)
        return set_error(CS3D_ERR_ARG, "cs3d_ops_apply: a transposed operator needs out != in");
    if (transposed && ((fp32 && !o->f[kind].dev) || (!fp32 && !o->d[kind].dev))) {
        const int base = kind - CS3D_FILTER_T;
        const int n = o->dir == 0 ? o->Nx : o->dir == 1 ? o->Ny : o->Nz;
        if (base == OP_FILTER && !(o->alpha_f > -0.5 && o->alpha_f < 0.5))
            return set_error(CS3D_ERR_ARG, "cs3d_ops_apply: operator not available (alpha_f?)");
        try {
            if (fp32)
                make_plan(o->f[kind], base, o->periodic, n, o->dh, o->alpha_f, 0, true);
            else
                make_plan(o->d[kind], base, o->periodic, n, o->dh, o->alpha_f, 0, true);
        } catch (const std::invalid_argument &ex) {
            return set_error(CS3D_ERR_ARG, std::string("cs3d_ops_apply: ") + ex.what());
        } catch (const std::exception &ex) {
            return set_error(CS3D_ERR_HIP, std::string("cs3d_ops_apply: ") + ex.what());
        }
    }
    if ((fp32 && !o->f[kind].dev) || (!fp32 && !o->d[kind].dev))
        return set_error(CS3D_ERR_ARG, "cs3d_ops_apply: operator not available (alpha_f?)");
    LineMap lm = layout == CS3D_LAYOUT_REFERENCE ? ref_lines(o->dir, o->Nx, o->Ny, o->Nz)
                                                 : xyz_lines(o->dir, o->Nx, o->Ny, o->Nz);
    int map = lm.sp == 1 ? MAP_ROW : MAP_COL;
    size_t N = (size_t)o->Nx * o->Ny * o->Nz;
    size_t bytes = N * (fp32 ? sizeof(float) : sizeof(double));
    if (bytes > 0xFFFFFFFFull)
        return set_error(CS3D_ERR_ARG, "cs3d_ops_apply: field above 4 GiB (32-bit offsets)");
    const void *din = in;
    void *dout = out;
    hipError_t e;
    if (ptr_kind == CS3D_HOST) {
        if (o->nbytes < bytes) {
            if (o->dbuf_in) (void)hipFree(o->dbuf_in);
            if (o->dbuf_out) (void)hipFree(o->dbuf_out);
            o->dbuf_in = o->dbuf_out = nullptr;
            o->nbytes = 0;
            if ((e = hipMalloc(&o->dbuf_in, bytes)) != hipSuccess ||
                (e = hipMalloc(&o->dbuf_out, bytes)) != hipSuccess)
                return set_error(CS3D_ERR_HIP, hip_msg(e, "cs3d_ops_apply: hipMalloc"));
            o->nbytes = bytes;
        }
        e = hipMemcpyAsync(o->dbuf_in, in, bytes, hipMemcpyHostToDevice, o->stream);
        if (e != hipSuccess) return set_error(CS3D_ERR_HIP, hip_msg(e, "cs3d_ops_apply: H2D"));
        din = o->dbuf_in;
        dout = o->dbuf_out;
    } else if (ptr_kind == CS3D_DEVICE) {
        // contiguous lines are read and written as 16-byte chunks
        if (((uintptr_t)in | (uintptr_t)out) % 16 != 0)
            return set_error(CS3D_ERR_ARG, "cs3d_ops_apply: device pointers must be 16-byte aligned");
    } else {
        return set_error(CS3D_ERR_ARG, "cs3d_ops_apply: bad ptr_kind");
    }
    // an earlier asynchronous failure is reported as such, not as this launch's
    e = hipGetLastError();
    if (e != hipSuccess)
        return set_error(CS3D_ERR_HIP, hip_msg(e, "cs3d_ops_apply: earlier HIP failure"));
    if (transposed && fp32)
        e = launch_field_opT<float>(o->f[kind], map, lm, (const float *)din, (float *)dout, 1.0f,
                                    0, o->stream);
    else if (transposed)
        e = launch_field_opT<double>(o->d[kind], map, lm, (const double *)din, (double *)dout,
                                     1.0, 0, o->stream);
    else if (fp32)
        e = launch_field_op<float>(o->f[kind], map, lm, (const float *)din, (float *)dout, o->stream);
    else
        e = launch_field_op<double>(o->d[kind], map, lm, (const double *)din, (double *)dout,
                                    o->stream);
    if (e != hipSuccess)
        return set_error(CS3D_ERR_HIP,
                         hip_msg(e, "cs3d_ops_apply: launch") + " (n=" +
                             std::to_string(fp32 ? o->f[kind].host.n : o->d[kind].host.n) +
                             " P=" + std::to_string(fp32 ? o->f[kind].host.P : o->d[kind].host.P) +
                             " threads=" +
                             std::to_string(fp32 ? o->f[kind].threads : o->d[kind].threads) + ")");
    if (ptr_kind == CS3D_HOST) {
        e = hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, o->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(o->stream);
        if (e != hipSuccess) return set_error(CS3D_ERR_HIP, hip_msg(e, "cs3d_ops_apply: D2H"));
    }
    return CS3D_OK;
}

int cs3d_ops_synchronize(cs3d_ops *o)
{
    if (!o) return set_error(CS3D_ERR_ARG, "cs3d_ops_synchronize: NULL");
    hipError_t e = hipStreamSynchronize(o->stream);
    if (e != hipSuccess) return set_error(CS3D_ERR_HIP, hip_msg(e, "cs3d_ops_synchronize"));
    return CS3D_OK;
}

int cs3d_plan_info(int kind, int bc_type, int n, double dh, double alpha_f, int precision,
                   int ints[8], double *reals, size_t nreals, int *segs, size_t nsegs)
{
    if (!ints) return set_error(CS3D_ERR_ARG, "cs3d_plan_info: ints is NULL");
    if (kind < 0 || kind > 2) return set_error(CS3D_ERR_ARG, "cs3d_plan_info: bad kind");
    OpHost h;
    try {
        h = plan_for(kind, bc_type == CS3D_PERIODIC_SOLVE, n, dh, alpha_f, precision == CS3D_FP32);
    } catch (const std::exception &ex) {
        return set_error(CS3D_ERR_ARG, std::string("cs3d_plan_info: ") + ex.what());
    }
    size_t need = 1 + 5 + 56 + 5 * (size_t)h.n + 2 * (size_t)h.P * h.K;
    ints[0] = h.n;
    ints[1] = h.M;
    ints[2] = h.P;
    ints[3] = h.K;
    ints[4] = h.uniform;
    ints[5] = h.cyclic;
    ints[6] = h.NB;
    ints[7] = (int)need;
    if (!reals) return CS3D_OK;
    if (nreals < need || !segs || nsegs < 2 * (size_t)h.P)
        return set_error(CS3D_ERR_ARG, "cs3d_plan_info: buffers too small");
    double *r = reals;
    *r++ = h.scale;
    for (int i = 0; i < 5; i++) *r++ = h.tap[i];
    for (int a = 0; a < 4; a++)
        for (int u = 0; u < 7; u++) *r++ = h.bcl[a][u];
    for (int a = 0; a < 4; a++)
        for (int u = 0; u < 7; u++) *r++ = h.bcr[a][u];
    for (const std::vector<double> *v : {&h.m, &h.iw, &h.mu, &h.pf, &h.pb, &h.phi, &h.psi})
        for (double x : *v) *r++ = x;
    for (int g = 0; g < h.P; g++) {
        segs[g] = h.sstart[g];
        segs[h.P + g] = h.slen[g];
    }
    return CS3D_OK;
}

void cs3d_ops_destroy(cs3d_ops *o)
{
    if (!o) return;
    for (int k = 0; k < 6; k++) {
        o->f[k].release();
        o->d[k].release();
    }
    if (o->dbuf_in) (void)hipFree(o->dbuf_in);
    if (o->dbuf_out) (void)hipFree(o->dbuf_out);
    if (o->stream) (void)hipStreamDestroy(o->stream);
    delete o;
}

}  // extern "C"
