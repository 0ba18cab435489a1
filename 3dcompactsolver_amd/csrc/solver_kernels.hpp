// solver_kernels.hpp -- data shared by the fused stage kernels and the host
// solver: field pointers, physical constants, per-stage arguments.
#pragma once

#include <hip/hip_runtime.h>

#include "line_engine.hpp"

namespace cs3d {

// Conserved variables are stored as perturbations of a constant reference
// state for density and total energy (rho' = rho - rho_ref,
// rhoE' = rhoE - rhoE_ref); momenta are stored as they are.  Compact
// derivatives of constants vanish and the filter preserves constants, so the
// scheme is unchanged; the point is that FP32 keeps the pressure and
// temperature fluctuations to full relative precision.
//
// A set of the five conserved fields (rho', rhoU, rhoV, rhoW, rhoE') lives
// in HBM as quads of the first four per point, m[4 i + v] for v = 0..3 (one
// 16-byte FP32 or 32-byte FP64 access per point), and rhoE' apart, e[i].
// A column sweep's wavefront row of 16 lines then reads the set as 256 + 64
// contiguous bytes instead of five 64-byte pieces (measured on the y and z
// sweeps' access pattern alone at 256^3 FP32: 26% and 30% less time).
template <typename real>
struct QSet {
    real *m;  // quads (rho', rhoU, rhoV, rhoW)
    real *e;  // rhoE'
};

// the five values of a set at point i
template <typename real>
__device__ __forceinline__ void qset_load(const QSet<real> &s, long i, real (&q)[5])
{
    typedef real vec4 __attribute__((ext_vector_type(4)));
    const vec4 v = *(const vec4 *)(s.m + 4 * i);
    q[0] = v[0];
    q[1] = v[1];
    q[2] = v[2];
    q[3] = v[3];
    q[4] = s.e[i];
}
template <typename real>
__device__ __forceinline__ void qset_store(const QSet<real> &s, long i, const real (&q)[5])
{
    typedef real vec4 __attribute__((ext_vector_type(4)));
    vec4 v;
    v[0] = q[0];
    v[1] = q[1];
    v[2] = q[2];
    v[3] = q[3];
    *(vec4 *)(s.m + 4 * i) = v;
    s.e[i] = q[4];
}
// the five values of a set at the point of single-field byte offset o
template <typename real>
__device__ __forceinline__ void qset_load_at(const QSet<real> &s, unsigned o, real (&q)[5])
{
    typedef real vec4 __attribute__((ext_vector_type(4)));
    const vec4 v = *(const vec4 *)((const char *)s.m + 4ull * o);
    q[0] = v[0];
    q[1] = v[1];
    q[2] = v[2];
    q[3] = v[3];
    q[4] = *(const real *)((const char *)s.e + o);
}

template <typename real>
struct Fields {
    QSet<real> q;   // state (*1)
    QSet<real> qk;  // stage state (*k)
    QSet<real> qkw; // where the last sweep writes the next stage state: qk
                    // itself, or a stage-state store of the adjoint
    QSet<real> acc; // RK accumulator (*2)
    // compensated FP32 mode: the rounding residuals (low parts) of the
    // state, the RK4 accumulator and TVDRK3's stage state, so that each
    // value is hi + lo to about twice the working precision; null when off
    QSet<real> qlo, acclo, qklo;
    // intermediates of a stage, interleaved (pairs and quads, see Win):
    // column sweeps write 2 or 4 values per point with one store
    real *dYut;     // (Uy, Ty) pairs of the stage (first sweep)
    real *dYvw;     // (Vy, Wy) pairs
    real *dZuw;     // (Uz, Wz) pairs (middle sweep)
    real *dZvt;     // (Vz, Tz) pairs
    real *mix[4];   // stale-primitive D_y U, D_y V, D_z U, D_z W (sponge runs only)
    real *Ry4;      // partial right-hand side of the first (y) sweep: quads of
                    // the continuity and momentum parts
    real *RyE;      // its energy part
    real *Rz4;      // the same for the middle (z) sweep
    real *RzE;
    real *k2[5];    // dt*RHS (only written in RHS-only mode)
    real *sigma;    // sponge strength (null without sponge)
    QSet<real> avg;    // sponge running averages (perturbation form)
    real *diag[4];     // per-point diagnostic inputs of the last stage
    QSet<real> qstale; // state on the sponge planes before the sponge reset
};

template <typename real>
struct Phys {
    real gm1, gamma;
    real iR;           // 1 / R_gas
    real rho_ref, rhoE_ref, p_ref;
    real T0s, RT0s;    // reference temperature p_ref/(rho_ref R), R*T0s
    real h_ref;        // rhoE_ref + p_ref
    real mu_ref, iTref_ig;  // IdealGas mu_ref and 1/T_ref (T_ref = 1)
    real cpPr;         // cp / Pr
    real idx, idy, idz;
};

// device-resident scalars of the time loop
struct StepScalars {
    double dt;
    double time;
    double cfl_max;    // max over the state of sum_i (|u_i| + c)/dx_i
    int timeStep;
    int filterTimeStep;
    int nonfinite;     // set when a reduction meets NaN/Inf
    int pad_;
};

struct StageArgs {
    int rk;            // 4 or 3
    int rkStep;        // 1..4
    int mode;          // 0: RK update, 1: write dt*RHS into k2 only
    int diag;          // store per-point diagnostic inputs
    int sponge_face[6];// 1 where a face is a sponge (its plane's RHS is zeroed)
    int Nx, Ny, Nz;
    int compensated;
    int stale;         // sweeps take primitives on sponge planes from qstale
    int wall_face[6];  // 1 where a face is a wall (adiabatic or moving)
};

// one wall face for k_wall_face (walls.hip)
struct WallFace {
    int dir, hi;       // normal direction, 0: the i/j/k = 0 face, 1: the last
    double vel[3];     // wall velocity, normal component zero
    double R;          // gas constant
};

// XCD-aware workgroup order: the hardware hands consecutive workgroups to
// the 8 XCDs in turn; remapping makes workgroups that share 128-byte lines
// (neighbouring line batches) run on the same XCD and share its L2.  A grid
// that is not a multiple of 8 keeps the remap on its first 8 * (nb / 8)
// workgroups (batches of LX lines whose rows straddle cache lines, as with
// 15 lines of 68 segments, share lines with both neighbours) and runs the
// last nb % 8 in place.
__device__ __forceinline__ int xcd_block(int b, int nb)
{
    const int nx = 8;
    const int n8 = nb - nb % nx;
    if (n8 < nx || b >= n8) return b;
    return (b % nx) * (n8 / nx) + b / nx;
}

// most solves batched through one exchange in a sweep kernel
constexpr int SWEEP_PLANES = 13;
// the FP64 z sweep on uniform lines solves its two mixed derivatives in the
// same batch as the other 13 (stage_z.hip)
template <typename real>
constexpr bool z_merged(bool uni, bool stale)
{
    return sizeof(real) == 8 && uni && !stale;
}
template <typename real>
constexpr int sweep_planes(int dir)
{
    return dir == 2 && z_merged<real>(true, false) ? SWEEP_PLANES + 2 : SWEEP_PLANES;
}
// most fields whose windows a sweep kernel stages through LDS at once
constexpr int SWEEP_STAGE_FIELDS = 5;

// workgroup size limit of the sweeps: FP32 fits 4 waves/SIMD (128 VGPRs),
// FP64 holds twice the registers and runs 2 waves/SIMD
template <typename real>
constexpr int sweep_threads() { return sizeof(real) == 4 ? 1024 : 512; }

// Column sweep variants that carry more live values than the registers of a
// full workgroup allow run half-size workgroups (twice the registers per
// lane): in FP32 (128 VGPRs at 1024 lanes) the z sweep for table plans
// (ragged or non-periodic z lines) and for the stale stage, the y sweep for
// table plans; in FP64 (256 VGPRs at 512 lanes) the z sweep for table
// plans and the y sweep's stale stage on them.  At full size they spilled
// 12-80 bytes per lane to scratch.  The host launches them with a matching
// geometry.
template <typename real>
constexpr bool col_small(int dir, bool uni, bool stale)
{
    return sizeof(real) == 4 ? (dir == 2 ? (!uni || stale) : !uni) : (!uni && (stale || dir == 2));
}
template <typename real, int DIR, bool UNI, bool STALE>
constexpr int col_threads()
{
    return col_small<real>(DIR, UNI, STALE) ? sweep_threads<real>() / 2 : sweep_threads<real>();
}

template <typename real>
struct Ptr5 {
    const real *in[5];
    real *out[5];
};

// what a filter sweep reads and writes: a conserved set (set = 1) or five
// separate fields (the adjoint's cotangents)
template <typename real>
struct FiltIO {
    int set;
    QSet<real> qin, qout;
    QSet<real> qin_lo, qout_lo;  // compensated mode: the sets' low parts
    Ptr5<real> sep;
};

// hi + lo of a compensated value, and its split back into FP32 hi and lo
__device__ __forceinline__ double two_part(float hi, float lo) { return (double)hi + (double)lo; }
__device__ __forceinline__ void split_part(double v, float &hi, float &lo)
{
    hi = (float)v;
    lo = (float)(v - (double)hi);
}

// launchers (stage_*.hip)
template <typename real>
hipError_t launch_sweep(int which, const OpDev<real> *d1, const OpDev<real> *d2,
                        const OpDev<real> *fx, int uni, int M,
                        const LaunchGeom &lg, const LineMap &lm, const Fields<real> &F,
                        const Phys<real> &ph, const StageArgs &sa, const StepScalars *sc,
                        hipStream_t st);
// part non-null: the sweep also leaves per-workgroup CFL partials there
// (maximum and NaN count, two doubles per workgroup) for launch_reduce_cfl
template <typename real>
hipError_t launch_filter5(const OpDev<real> *fo, int uni, int M, const LaunchGeom &lg, int map,
                          const LineMap &lm, const FiltIO<real> &io, const Phys<real> &ph,
                          double *part, hipStream_t st);
constexpr int RED_SLOTS_HOST = 16;
// the filter's CFL partials into the reduction outputs (slot 0 the CFL
// maximum, slot 7 the NaN count) and StepScalars::cfl_max
hipError_t launch_reduce_cfl(const double *part, int nb, double *out, StepScalars *sc,
                             hipStream_t st);
template <typename real>
hipError_t launch_reduce(const Fields<real> &F, const Phys<real> &ph, long N, int diag,
                         double *part, int nb, double *out, StepScalars *sc, hipStream_t st);
hipError_t launch_dt(StepScalars *sc, double cfl, double dt_fixed, hipStream_t st);
// per-(j, k) row sums of the plane statistics (profiles.hip): 9 doubles a
// row; pass 0 means (6 used), pass 1 second moments about `mean` (6 per j)
template <typename real>
hipError_t launch_plane_rows(const Fields<real> &F, const Phys<real> &ph, int Nx, int Ny, int Nz,
                             int pass, const double *mean, double *rows, hipStream_t st);
// The shear-layer profiles without any allocation (profiles.hip): the same
// row sums into a file-scope device array, reduced over the planes on the
// device, the 11 rows of cs3d_diag_shear_layer copied into host_out (11 Ny
// doubles).  Grids with Ny * Nz above SHEAR_MAX_ROWS or Ny above
// SHEAR_MAX_NY are refused (hipErrorInvalidValue); calls that overlap on
// one GPU share its array and race.
constexpr long SHEAR_MAX_ROWS = 131072;
constexpr int SHEAR_MAX_NY = 8192;
template <typename real>
hipError_t shear_profiles_static(const Fields<real> &F, const Phys<real> &ph, int Nx, int Ny,
                                 int Nz, int fix_q12, double *host_out, hipStream_t st);
// preStepBCHandling of one wall face on the stage state (walls.hip)
template <typename real>
hipError_t launch_wall_face(const Fields<real> &F, const Phys<real> &ph, const StageArgs &sa,
                            const WallFace &w, hipStream_t st);
template <typename real>
hipError_t launch_sponge(const Fields<real> &F, const Phys<real> &ph, const StageArgs &sa,
                         const StepScalars *sc, double avgT, double epsP, double spongeP,
                         hipStream_t st);

}  // namespace cs3d
