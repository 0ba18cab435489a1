// walls.hip -- adiabatic and moving walls before the sweeps of a stage
// (preStepBCHandling, CSolver_AWS.cpp:553-826).
//
// Every point of a wall face plane of the stage state (*1 at the first
// stage, *k after) gets the wall velocity in its momenta (zero, or the
// moving wall's tangential components, CSolver_AWS.hpp:105-107) and the
// energy of p = rho R T_w, T_w = calcNeumann (Utils.hpp:52) of the
// temperatures of the six points inward: the value that makes the one-sided
// 7-point normal derivative vanish.  The density is kept.  The sweeps then
// form the wall primitives from these conserved values, as the reference
// differentiates its wall-corrected primitive arrays.
//
// One launch per face, in the reference's order (adiabatic faces X0..Z1,
// then moving faces X0..Z1): a point on an edge reads the values an earlier
// face left and keeps the last face's.
//
// At the first stage after a sponge update the reference's temperature
// array on the sponge face planes still belongs to the state before the
// update (see stage_x.hip); the neighbour temperatures are taken from the
// pre-update copy there, and a wall point on such a plane writes its new
// values into that copy as well, since the reference overwrites its
// primitive arrays at wall points.
#include "adjoint.hpp"

namespace cs3d {

__device__ __forceinline__ bool sponge_plane(const StageArgs &sa, int i, int j, int k)
{
    return (sa.sponge_face[0] && i == 0) || (sa.sponge_face[1] && i == sa.Nx - 1) ||
           (sa.sponge_face[2] && j == 0) || (sa.sponge_face[3] && j == sa.Ny - 1) ||
           (sa.sponge_face[4] && k == 0) || (sa.sponge_face[5] && k == sa.Nz - 1);
}

template <typename real>
__global__ __launch_bounds__(256) void k_wall_face(Fields<real> F, Phys<real> ph, StageArgs sa,
                                                   WallFace w)
{
    const int Nx = sa.Nx, Ny = sa.Ny, Nz = sa.Nz;
    const int na = w.dir == 0 ? Ny : Nx, nb = w.dir == 2 ? Ny : Nz;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= na * nb) return;
    const int a = t % na, b = t / na;
    int ijk[3];
    if (w.dir == 0) {
        ijk[0] = w.hi ? Nx - 1 : 0;
        ijk[1] = a;
        ijk[2] = b;
    } else if (w.dir == 1) {
        ijk[0] = a;
        ijk[1] = w.hi ? Ny - 1 : 0;
        ijk[2] = b;
    } else {
        ijk[0] = a;
        ijk[1] = b;
        ijk[2] = w.hi ? Nz - 1 : 0;
    }
    const long sxy = (long)Nx * Ny;
    const long ip = (long)ijk[2] * sxy + (long)ijk[1] * Nx + ijk[0];
    const long stride = w.dir == 0 ? 1 : w.dir == 1 ? (long)Nx : sxy;
    const long in = w.hi ? -stride : stride;
    const QSet<real> q = sa.rkStep > 1 ? F.qk : F.q;
    // temperatures (perturbation form T - T0s) of the six points inward;
    // the wall direction holds at least 7 points
    real Tn[6];
#pragma unroll
    for (int m = 0; m < 6; m++) {
        const long o = ip + (long)(m + 1) * in;
        int nijk[3] = {ijk[0], ijk[1], ijk[2]};
        nijk[w.dir] += w.hi ? -(m + 1) : (m + 1);
        const bool st = sa.stale && sponge_plane(sa, nijk[0], nijk[1], nijk[2]);
        real ir, U, V, Wv, pp, qo[5];
        qset_load(st ? F.qstale : q, o, qo);
        prim(ph, qo[0], qo[1], qo[2], qo[3], qo[4], ir, U, V, Wv, pp, Tn[m]);
    }
    // calcNeumann: its weights sum to one, so it acts on T - T0s alike
    const real Tw = (Tn[0] * real(360) - Tn[1] * real(450) + Tn[2] * real(400) -
                     Tn[3] * real(225) + Tn[4] * real(72) - Tn[5] * real(10)) /
                    real(147);
    const real rp = q.m[4 * ip];
    const real rho = ph.rho_ref + rp;
    const real u = (real)w.vel[0], v = (real)w.vel[1], ww = (real)w.vel[2];
    // p - p_ref = rho R T_w - rho_ref R T0s
    const real pp = rp * ph.RT0s + rho * (real)w.R * Tw;
    // rhoE - rhoE_ref (p_ref = (gamma - 1) rhoE_ref)
    const real re = pp / ph.gm1 + real(0.5) * rho * (u * u + v * v + ww * ww);
    const real m1 = rho * u, m2 = rho * v, m3 = rho * ww;
    const real qw[5] = {rp, m1, m2, m3, re};
    qset_store(q, ip, qw);
    // compensated mode: the wall values are set, not accumulated
    const QSet<real> qlo = sa.rkStep > 1 ? F.qklo : F.qlo;
    if (qlo.m) {
        const real z[5] = {0, 0, 0, 0, 0};
        qset_store(qlo, ip, z);
    }
    if (sa.stale && sponge_plane(sa, ijk[0], ijk[1], ijk[2])) qset_store(F.qstale, ip, qw);
}

// Reverse of k_wall_face on the cotangent lam, given the state st before
// the face: at a face point the momenta m = rho u_w and the energy
// E = rho R T_w/(gamma - 1) + rho |u_w|^2 / 2 were set, so their cotangents
// pass to the density and, through T_w = sum_m c_m T_m (the calcNeumann
// weights), to the six inward points' states by T = p/(rho R),
// p = (gamma - 1)(E - |m|^2/(2 rho)).  At the first stage after a sponge
// update (stS set) the face also set the pre-reset copy on the sponge
// planes, and the inward temperatures on them came from that copy: the
// copy's cotangent lamS joins lam at a face point on a sponge plane, and
// the inward terms on sponge planes go to lamS, with stS's values.  One
// thread per face point: a face point's inward points belong to it alone.
template <typename real>
__global__ __launch_bounds__(256) void k_wall_face_vjp(QSet<real> st, QSet<real> lam,
                                                       QSet<real> stS, QSet<real> lamS,
                                                       Phys<real> ph, StageArgs sa, WallFace w)
{
    const int Nx = sa.Nx, Ny = sa.Ny, Nz = sa.Nz;
    const int na = w.dir == 0 ? Ny : Nx, nb = w.dir == 2 ? Ny : Nz;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= na * nb) return;
    const int a = t % na, b = t / na;
    int ijk[3];
    if (w.dir == 0) {
        ijk[0] = w.hi ? Nx - 1 : 0;
        ijk[1] = a;
        ijk[2] = b;
    } else if (w.dir == 1) {
        ijk[0] = a;
        ijk[1] = w.hi ? Ny - 1 : 0;
        ijk[2] = b;
    } else {
        ijk[0] = a;
        ijk[1] = b;
        ijk[2] = w.hi ? Nz - 1 : 0;
    }
    const bool stale = stS.m != nullptr;
    const long sxy = (long)Nx * Ny;
    const long ip = (long)ijk[2] * sxy + (long)ijk[1] * Nx + ijk[0];
    const long stride = w.dir == 0 ? 1 : w.dir == 1 ? (long)Nx : sxy;
    const long in = w.hi ? -stride : stride;
    const real cn[6] = {real(360) / real(147), real(-450) / real(147), real(400) / real(147),
                        real(-225) / real(147), real(72) / real(147), real(-10) / real(147)};
    // whether inward point m + 1 read the pre-reset copy
    auto inward_stale = [&](int m) {
        int nijk[3] = {ijk[0], ijk[1], ijk[2]};
        nijk[w.dir] += w.hi ? -(m + 1) : (m + 1);
        return stale && sponge_plane(sa, nijk[0], nijk[1], nijk[2]);
    };
    real Tw = 0;
#pragma unroll
    for (int m = 0; m < 6; m++) {
        real qo[5], ir, U, V, Wv, pp, Tp;
        qset_load(inward_stale(m) ? stS : st, ip + (long)(m + 1) * in, qo);
        prim(ph, qo[0], qo[1], qo[2], qo[3], qo[4], ir, U, V, Wv, pp, Tp);
        Tw += cn[m] * (ph.T0s + Tp);
    }
    real qf[5], l[5];
    qset_load(st, ip, qf);
    qset_load(lam, ip, l);
    if (stale && sponge_plane(sa, ijk[0], ijk[1], ijk[2])) {
        real ls[5];
        qset_load(lamS, ip, ls);
#pragma unroll
        for (int k = 0; k < 5; k++) {
            l[k] += ls[k];
            ls[k] = 0;
        }
        qset_store(lamS, ip, ls);
    }
    const real rho = ph.rho_ref + qf[0];
    const real u[3] = {(real)w.vel[0], (real)w.vel[1], (real)w.vel[2]};
    const real R = (real)w.R;
    l[0] += l[1] * u[0] + l[2] * u[1] + l[3] * u[2] +
            l[4] * (R * Tw / ph.gm1 + real(0.5) * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
    const real gT = l[4] * rho * R / ph.gm1;
    l[1] = l[2] = l[3] = l[4] = 0;
    qset_store(lam, ip, l);
#pragma unroll
    for (int m = 0; m < 6; m++) {
        const long o = ip + (long)(m + 1) * in;
        const bool sp = inward_stale(m);
        const QSet<real> &so = sp ? stS : st;
        const QSet<real> &lt = sp ? lamS : lam;
        real qo[5], lo[5], ir, U, V, Wv, pp, Tp;
        qset_load(so, o, qo);
        qset_load(lt, o, lo);
        prim(ph, qo[0], qo[1], qo[2], qo[3], qo[4], ir, U, V, Wv, pp, Tp);
        const real r = ph.rho_ref + qo[0], irr = real(1) / (r * R);
        const real Tb = gT * cn[m];
        lo[4] += Tb * ph.gm1 * irr;
        lo[1] -= Tb * ph.gm1 * U * irr;
        lo[2] -= Tb * ph.gm1 * V * irr;
        lo[3] -= Tb * ph.gm1 * Wv * irr;
        lo[0] += Tb * (ph.gm1 * real(0.5) * (U * U + V * V + Wv * Wv) * irr -
                       (ph.T0s + Tp) / r);
        qset_store(lt, o, lo);
    }
}

template <typename real>
hipError_t launch_wall_face_vjp(const QSet<real> &st, const QSet<real> &lam, const QSet<real> &stS,
                                const QSet<real> &lamS, const Phys<real> &ph, const StageArgs &sa,
                                const WallFace &w, hipStream_t s)
{
    if ((stS.m == nullptr) != (lamS.m == nullptr)) return hipErrorInvalidValue;
    if (w.dir < 0 || w.dir > 2) return hipErrorInvalidValue;
    const int n = w.dir == 0 ? sa.Nx : w.dir == 1 ? sa.Ny : sa.Nz;
    if (n < 7) return hipErrorInvalidValue;
    const long pts = w.dir == 0 ? (long)sa.Ny * sa.Nz
                                : w.dir == 1 ? (long)sa.Nx * sa.Nz : (long)sa.Nx * sa.Ny;
    hipLaunchKernelGGL((k_wall_face_vjp<real>), dim3((unsigned)((pts + 255) / 256)), dim3(256), 0,
                       s, st, lam, stS, lamS, ph, sa, w);
    return hipGetLastError();
}

template hipError_t launch_wall_face_vjp<float>(const QSet<float> &, const QSet<float> &,
                                                const QSet<float> &, const QSet<float> &,
                                                const Phys<float> &, const StageArgs &,
                                                const WallFace &, hipStream_t);
template hipError_t launch_wall_face_vjp<double>(const QSet<double> &, const QSet<double> &,
                                                 const QSet<double> &, const QSet<double> &,
                                                 const Phys<double> &, const StageArgs &,
                                                 const WallFace &, hipStream_t);

template <typename real>
hipError_t launch_wall_face(const Fields<real> &F, const Phys<real> &ph, const StageArgs &sa,
                            const WallFace &w, hipStream_t st)
{
    if (w.dir < 0 || w.dir > 2) return hipErrorInvalidValue;
    const int n = w.dir == 0 ? sa.Nx : w.dir == 1 ? sa.Ny : sa.Nz;
    if (n < 7) return hipErrorInvalidValue;
    const long pts = w.dir == 0 ? (long)sa.Ny * sa.Nz
                                : w.dir == 1 ? (long)sa.Nx * sa.Nz : (long)sa.Nx * sa.Ny;
    const int nb = (int)((pts + 255) / 256);
    hipLaunchKernelGGL((k_wall_face<real>), dim3(nb), dim3(256), 0, st, F, ph, sa, w);
    return hipGetLastError();
}

template hipError_t launch_wall_face<float>(const Fields<float> &, const Phys<float> &,
                                            const StageArgs &, const WallFace &, hipStream_t);
template hipError_t launch_wall_face<double>(const Fields<double> &, const Phys<double> &,
                                             const StageArgs &, const WallFace &, hipStream_t);

}  // namespace cs3d
