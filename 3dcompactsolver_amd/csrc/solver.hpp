// solver.hpp -- precision-independent interface of the GPU solver used by
// the C ABI (capi_solver.cpp).  Methods throw std::invalid_argument,
// std::logic_error (call out of order) or std::runtime_error (HIP failure).
#pragma once

#include "../../include/cs3d.h"

namespace cs3d {

// launch classes of the per-launch timer
constexpr int PROF_CLASSES = 6;  // X sweep, Y sweep, Z sweep, filter, reductions, other

struct SolverBase {
    virtual ~SolverBase() {}
    virtual void set_ic(const double *rho0, const double *U0, const double *V0,
                        const double *W0, const double *p0) = 0;
    virtual void pre_step() = 0;
    virtual void sub_step(int rkStep, int rkLast) = 0;
    virtual void post_step() = 0;
    virtual void step(int nsteps) = 0;
    virtual void rhs(int rkStep) = 0;
    virtual int check_finite() = 0;
    virtual void get_conserved(int set, double *const *out) = 0;
    virtual void set_conserved(int set, const double *const *in) = 0;
    virtual void get_primitives(double *const *out) = 0;
    virtual void get_time(double *time, double *dt, int *ts) = 0;
    virtual void get_run_state(cs3d_run_state *rs) = 0;
    virtual void shear_profiles(int fix_q12, double *out) = 0;
    // the same without allocating (device memory or host buffers)
    virtual void shear_profiles_noalloc(int fix_q12, double *out) = 0;
    // optical-path integrals of the density of conserved set `set` (1 or 4)
    virtual void aero_optics(int set, int num_angles, double max_angle, double *opl) = 0;
    virtual void set_run_state(const cs3d_run_state &rs) = 0;
    virtual void set_diag(int on) = 0;
    virtual int diag_tgv(double out[7]) = 0;
    virtual int diag_turb(double out[11]) = 0;
    virtual int kinetic_energy(double *ke) = 0;
    virtual void synchronize() = 0;
    virtual void *stream() = 0;
    virtual void profile_enable(int on) = 0;
    virtual void profile_report(double *ms, int *counts) = 0;
    // discrete adjoint:
    // lam^T d(dt*RHS)/dQ for the conserved set `set`, and the gradient of
    // the final volume-averaged kinetic energy after nsteps with respect to
    // the initial velocities
    virtual void rhs_vjp(int set, const double *const *lam, double *const *out) = 0;
    virtual void ke_gradient(int nsteps, int checkpoint, double *J, double *gU, double *gV,
                             double *gW) = 0;
    // frees the adjoint's work fields and state pool
    virtual void adjoint_release() = 0;
};

SolverBase *make_solver(const cs3d_config &c);

}  // namespace cs3d
