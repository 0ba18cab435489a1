// plan_dev.hpp -- device residency of operator plans and the launchers of the
// standalone per-field operator (the Derivatives / Filter drop-in path).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "line_engine.hpp"
#include "plan.hpp"

namespace cs3d {

// allow_uniform = false forces the per-point table path even when the
// segments come out equal (the caller picked M for the table kernels).
// transpose: the solve runs on A^T (the transposed operator's first step)
OpHost build_plan(int kind, int periodic, int n, double dh, double alpha_f, int M, double tol,
                  bool allow_uniform, bool transpose = false);
// the plan the library uses for an operator in a precision (segment length
// and truncation chosen as the kernels expect)
// seg > 0: the uniform segment length (default: SEG_FILTER for the filter,
// SEG_DERIV for the derivatives)
OpHost plan_for(int kind, int periodic, int n, double dh, double alpha_f, bool fp32, int seg = 0,
                bool transpose = false);

// Points per lane.  Derivatives keep 4 points per lane in both precisions
// (the fused sweeps hold many windows at once); the periodic filter keeps 8;
// non-periodic and ragged lines (table path) keep 4.
constexpr int SEG_DERIV = 4;
constexpr int SEG_FILTER = 8;
constexpr int SEG_TABLE = 4;
// uniform 3-point segments for the forward z sweep on periodic z lines whose
// length divides by 3 but not by 4 (42 in the shear layer), which would
// otherwise take the table kernels
constexpr int SEG_DERIV3 = 3;

// Carry terms unrolled by the kernels on uniform plans (any further terms
// run in a loop), one count per operator: the K its periodic 4-point plan
// keeps at the truncation tolerance of plan_for, 5 (FP32) / 11 (FP64) for
// the first derivative, 3 / 6 for the second; the filter's grow as alpha_f
// approaches 1/2 (12 / 25 at 0.49 with 8-point segments).  A lane's ring
// margin is max(K, KC) over its operators.
template <typename real>
constexpr int kc_deriv() { return sizeof(real) == 4 ? 5 : 11; }
template <typename real>
constexpr int kc_d2() { return sizeof(real) == 4 ? 3 : 6; }
template <typename real>
constexpr int kc_filter() { return sizeof(real) == 4 ? 12 : 25; }
// the filter on 4-point segments (fused into the x sweep): 23 / 49 at 0.49
template <typename real>
constexpr int kc_filter_x() { return sizeof(real) == 4 ? 23 : 49; }

template <typename real>
struct PlanDev {
    OpHost host;
    OpDev<real> hd{};           // host copy holding device table pointers
    OpDev<real> *dev = nullptr; // device copy read by kernels
    void *tables = nullptr;
    int LX = 0;                 // lines per workgroup (standalone operator)
    int threads = 0;            // workgroup size = P * LX
    size_t shm = 0;             // LDS bytes of the exchange ring
    // transposed plans (make_plan with transpose): the boundary columns of
    // B^T, and LDS for the solved lines after the ring (shm includes it)
    OpTDev<real> *tdev = nullptr;
    OpTHost thost;

    void release();
};

// Builds (FP64 host factorisation) and uploads a plan.  Throws
// std::invalid_argument for unsupported sizes, std::runtime_error on HIP
// failures.
// transpose: the plan of the transposed operator (the solve on A^T, and the
// boundary columns of B^T; launch it with launch_field_opT)
template <typename real>
void make_plan(PlanDev<real> &pl, int kind, int periodic, int n, double dh, double alpha_f,
               int seg = 0, bool transpose = false);

// out = op(in) on every line of lm; map is MAP_ROW for contiguous lines.
template <typename real>
hipError_t launch_field_op(const PlanDev<real> &pl, int map, const LineMap &lm,
                           const real *in, real *out, hipStream_t st);

// out = coef * op^T(in), or out += coef * op^T(in) with accumulate, on every
// line of lm (pl: a transposed plan).  in and out must not overlap.
template <typename real>
hipError_t launch_field_opT(const PlanDev<real> &pl, int map, const LineMap &lm,
                            const real *in, real *out, real coef, int accumulate,
                            hipStream_t st);

// Batched operators along one direction (ops_batch.hip): the first- and
// second-derivative plans d1, d2 of one direction (same segmentation)
// applied to several fields in one solve batch.  Outputs must be distinct
// from each other and from every input.
constexpr int OPB_MAX = 12;
template <typename real>
struct OpBatch {
    const real *in[OPB_MAX];
    real *out[OPB_MAX];
    real coef[OPB_MAX];  // transposed launches: weight of each term
    int acc[OPB_MAX];    // transposed launches: out[f] += (1) or = (0)
};
// out[i] = D1(in[i]) for i < n1, out[n1 + i] = D2(in[i]) for i < n2
// (n1 = 4, 6 or 8; n2 = 4)
template <typename real>
hipError_t launch_ops_fwd(const PlanDev<real> &d1, const PlanDev<real> &d2, int map,
                          const LineMap &lm, int n1, int n2, const OpBatch<real> &b,
                          hipStream_t st);
// transposed plans.  n2 = 0 (n1 = 5, 7 or 9): out[i] (+)= coef[i] D1^T(in[i]);
// n1 = n2 = 4: out[f] (+)= coef[f] D1^T(in[f]) + coef[4 + f] D2^T(in[4 + f])
template <typename real>
hipError_t launch_ops_T(const PlanDev<real> &d1, const PlanDev<real> &d2, int map,
                        const LineMap &lm, int n1, int n2, const OpBatch<real> &b,
                        hipStream_t st);

// line maps of an Nx*Ny*Nz field in XYZ order (Macros.hpp:8), and of the
// reference's per-direction layouts YZX / ZXY (Macros.hpp:9-10)
LineMap xyz_lines(int dir, int Nx, int Ny, int Nz);
LineMap ref_lines(int dir, int Nx, int Ny, int Nz);

std::string hip_msg(hipError_t e, const char *what);

}  // namespace cs3d
