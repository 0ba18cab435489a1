// capi_solver.cpp -- C ABI of the GPU solver (include/cs3d.h).
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "../../include/cs3d.h"
#include "capi_common.hpp"
#include "solver.hpp"

using namespace cs3d;

struct cs3d_solver {
    std::unique_ptr<SolverBase> impl;
};

namespace {

// run f, translating exceptions into error codes
template <typename Fn>
int guard(const char *name, Fn f)
{
    try {
        return f();
    } catch (const std::invalid_argument &e) {
        return set_error(CS3D_ERR_ARG, std::string(name) + ": " + e.what());
    } catch (const std::logic_error &e) {
        return set_error(CS3D_ERR_STATE, std::string(name) + ": " + e.what());
    } catch (const std::exception &e) {
        return set_error(CS3D_ERR_HIP, std::string(name) + ": " + e.what());
    }
}

int check_config(const cs3d_config *c)
{
    if (c->Nx < 5 || c->Ny < 5 || c->Nz < 5)
        return set_error(CS3D_ERR_ARG, "cs3d_create: every dimension needs at least 5 points");
    for (int d = 0; d < 3; d++)
        if (c->bc_type[d] != CS3D_PERIODIC_SOLVE && c->bc_type[d] != CS3D_DIRICHLET_SOLVE)
            return set_error(CS3D_ERR_ARG, "cs3d_create: bad bc_type");
    for (int f = 0; f < 6; f++) {
        int k = c->bc_kind[f];
        if (k < CS3D_BC_PERIODIC || k > CS3D_BC_INLET)
            return set_error(CS3D_ERR_ARG, "cs3d_create: bad bc_kind");
        if (k == CS3D_BC_INLET)
            return set_error(CS3D_ERR_UNSUPPORTED,
                             "cs3d_create: the reference has no inlet boundary handling");
        const int n = f < 2 ? c->Nx : f < 4 ? c->Ny : c->Nz;
        if ((k == CS3D_BC_ADIABATIC_WALL || k == CS3D_BC_MOVING_ADIABATIC_WALL) && n < 7)
            return set_error(CS3D_ERR_ARG,
                             "cs3d_create: a wall needs at least 7 points along its direction");
    }
    if (c->rk != CS3D_RK4 && c->rk != CS3D_TVDRK3)
        return set_error(CS3D_ERR_ARG, "cs3d_create: rk must be 3 or 4");
    if (c->precision != CS3D_FP64 && c->precision != CS3D_FP32)
        return set_error(CS3D_ERR_ARG, "cs3d_create: bad precision");
    if (c->compensated && c->precision != CS3D_FP32)
        return set_error(CS3D_ERR_ARG, "cs3d_create: the compensated mode is an FP32 mode");
    if (c->filter_step < 1) return set_error(CS3D_ERR_ARG, "cs3d_create: filter_step < 1");
    if (!(c->gamma > 1.0) || !(c->Pr > 0.0) || !(c->mu_ref >= 0.0))
        return set_error(CS3D_ERR_ARG, "cs3d_create: bad gas constants");
    if (!(c->Lx > 0.0 && c->Ly > 0.0 && c->Lz > 0.0))
        return set_error(CS3D_ERR_ARG, "cs3d_create: domain lengths must be positive");
    if (!(c->cfl > 0.0) && !(c->dt_fixed > 0.0))
        return set_error(CS3D_ERR_ARG, "cs3d_create: need cfl > 0 or dt_fixed > 0");
    return CS3D_OK;
}

}  // namespace

extern "C" {

int cs3d_config_default(cs3d_config *c, int Nx, int Ny, int Nz)
{
    if (!c) return set_error(CS3D_ERR_ARG, "cs3d_config_default: NULL");
    std::memset(c, 0, sizeof(*c));
    c->Nx = Nx;
    c->Ny = Ny;
    c->Nz = Nz;
    c->Lx = c->Ly = c->Lz = 1.0;
    c->cfl = 0.5;
    c->filter_step = 1;
    c->max_steps = 1 << 30;
    c->max_time = 1e300;
    c->alpha_f = 0.495;          // 3DCompact.cpp:90
    c->mu_ref = 1.0;
    c->gamma = 1.4;              // IdealGas.hpp:39
    c->Pr = 0.7;                 // IdealGas.hpp:40
    c->rk = CS3D_RK4;
    c->precision = CS3D_FP64;
    c->sponge_frac[0] = 0.25;    // SpongeBC.hpp:57-59
    c->sponge_frac[1] = 0.125;
    c->sponge_frac[2] = 0.25;
    c->sponge_strength = 6.0;    // SpongeBC.hpp:56
    c->sponge_avgT = 15.0;       // SpongeBC.hpp:53
    c->sponge_epsP = 0.005;      // SpongeBC.hpp:54
    return CS3D_OK;
}

int cs3d_create(const cs3d_config *cfg, cs3d_solver **out)
{
    if (!cfg || !out) return set_error(CS3D_ERR_ARG, "cs3d_create: NULL argument");
    *out = nullptr;
    int rc = check_config(cfg);
    if (rc) return rc;
    return guard("cs3d_create", [&]() -> int {
        std::unique_ptr<cs3d_solver> s(new cs3d_solver());
        s->impl.reset(make_solver(*cfg));
        *out = s.release();
        return CS3D_OK;
    });
}

void cs3d_destroy(cs3d_solver *s) { delete s; }

#define CS3D_NEED(s, name) \
    if (!(s)) return set_error(CS3D_ERR_ARG, name ": NULL solver")

int cs3d_set_initial_conditions(cs3d_solver *s, const double *rho0, const double *U0,
                                const double *V0, const double *W0, const double *p0)
{
    CS3D_NEED(s, "cs3d_set_initial_conditions");
    if (!rho0 || !U0 || !V0 || !W0 || !p0)
        return set_error(CS3D_ERR_ARG, "cs3d_set_initial_conditions: NULL array");
    return guard("cs3d_set_initial_conditions", [&]() -> int {
        s->impl->set_ic(rho0, U0, V0, W0, p0);
        return CS3D_OK;
    });
}

int cs3d_pre_step(cs3d_solver *s)
{
    CS3D_NEED(s, "cs3d_pre_step");
    return guard("cs3d_pre_step", [&]() -> int {
        s->impl->pre_step();
        return CS3D_OK;
    });
}

int cs3d_sub_step(cs3d_solver *s, int rkStep, int rkLast)
{
    CS3D_NEED(s, "cs3d_sub_step");
    if (rkStep < 1 || rkStep > 4) return set_error(CS3D_ERR_ARG, "cs3d_sub_step: bad rkStep");
    return guard("cs3d_sub_step", [&]() -> int {
        s->impl->sub_step(rkStep, rkLast);
        return CS3D_OK;
    });
}

int cs3d_post_step(cs3d_solver *s)
{
    CS3D_NEED(s, "cs3d_post_step");
    return guard("cs3d_post_step", [&]() -> int {
        s->impl->post_step();
        return CS3D_OK;
    });
}

int cs3d_step(cs3d_solver *s, int nsteps)
{
    CS3D_NEED(s, "cs3d_step");
    if (nsteps < 0) return set_error(CS3D_ERR_ARG, "cs3d_step: nsteps < 0");
    return guard("cs3d_step", [&]() -> int {
        s->impl->step(nsteps);
        int rc = s->impl->check_finite();
        if (rc) return set_error(rc, "cs3d_step: non-finite value in the CFL reduction");
        return CS3D_OK;
    });
}

int cs3d_rhs(cs3d_solver *s, int rkStep)
{
    CS3D_NEED(s, "cs3d_rhs");
    if (rkStep < 1 || rkStep > 4) return set_error(CS3D_ERR_ARG, "cs3d_rhs: bad rkStep");
    return guard("cs3d_rhs", [&]() -> int {
        s->impl->rhs(rkStep);
        return CS3D_OK;
    });
}

int cs3d_get_conserved(cs3d_solver *s, int set, double *rho, double *rhoU, double *rhoV,
                       double *rhoW, double *rhoE)
{
    CS3D_NEED(s, "cs3d_get_conserved");
    if (set < 1 || set > 6) return set_error(CS3D_ERR_ARG, "cs3d_get_conserved: bad set");
    double *o[5] = {rho, rhoU, rhoV, rhoW, rhoE};
    return guard("cs3d_get_conserved", [&]() -> int {
        s->impl->get_conserved(set, o);
        return CS3D_OK;
    });
}

int cs3d_set_conserved(cs3d_solver *s, int set, const double *rho, const double *rhoU,
                       const double *rhoV, const double *rhoW, const double *rhoE)
{
    CS3D_NEED(s, "cs3d_set_conserved");
    if (set < 1 || set > 6) return set_error(CS3D_ERR_ARG, "cs3d_set_conserved: bad set");
    const double *in[5] = {rho, rhoU, rhoV, rhoW, rhoE};
    return guard("cs3d_set_conserved", [&]() -> int {
        s->impl->set_conserved(set, in);
        return CS3D_OK;
    });
}

int cs3d_get_primitives(cs3d_solver *s, double *U, double *V, double *W, double *p, double *T,
                        double *mu, double *Amu, double *sos)
{
    CS3D_NEED(s, "cs3d_get_primitives");
    double *o[8] = {U, V, W, p, T, mu, Amu, sos};
    return guard("cs3d_get_primitives", [&]() -> int {
        s->impl->get_primitives(o);
        return CS3D_OK;
    });
}

int cs3d_get_time(cs3d_solver *s, double *time, double *dt, int *timeStep)
{
    CS3D_NEED(s, "cs3d_get_time");
    return guard("cs3d_get_time", [&]() -> int {
        s->impl->get_time(time, dt, timeStep);
        return CS3D_OK;
    });
}

int cs3d_diag_shear_layer(cs3d_solver *s, int fix_q12, double *out)
{
    CS3D_NEED(s, "cs3d_diag_shear_layer");
    if (!out) return set_error(CS3D_ERR_ARG, "cs3d_diag_shear_layer: NULL");
    return guard("cs3d_diag_shear_layer", [&]() -> int {
        s->impl->shear_profiles(fix_q12, out);
        return CS3D_OK;
    });
}

int cs3d_diag_shear_layer_noalloc(cs3d_solver *s, int fix_q12, double *out)
{
    CS3D_NEED(s, "cs3d_diag_shear_layer_noalloc");
    if (!out) return set_error(CS3D_ERR_ARG, "cs3d_diag_shear_layer_noalloc: NULL");
    return guard("cs3d_diag_shear_layer_noalloc", [&]() -> int {
        s->impl->shear_profiles_noalloc(fix_q12, out);
        return CS3D_OK;
    });
}

int cs3d_aero_optics(cs3d_solver *s, int set, int num_angles, double max_angle, double *opl)
{
    CS3D_NEED(s, "cs3d_aero_optics");
    if (!opl) return set_error(CS3D_ERR_ARG, "cs3d_aero_optics: NULL");
    return guard("cs3d_aero_optics", [&]() -> int {
        s->impl->aero_optics(set, num_angles, max_angle, opl);
        return CS3D_OK;
    });
}

int cs3d_get_run_state(cs3d_solver *s, cs3d_run_state *rs)
{
    CS3D_NEED(s, "cs3d_get_run_state");
    if (!rs) return set_error(CS3D_ERR_ARG, "cs3d_get_run_state: NULL");
    return guard("cs3d_get_run_state", [&]() -> int {
        s->impl->get_run_state(rs);
        return CS3D_OK;
    });
}

int cs3d_set_run_state(cs3d_solver *s, const cs3d_run_state *rs)
{
    CS3D_NEED(s, "cs3d_set_run_state");
    if (!rs) return set_error(CS3D_ERR_ARG, "cs3d_set_run_state: NULL");
    return guard("cs3d_set_run_state", [&]() -> int {
        s->impl->set_run_state(*rs);
        return CS3D_OK;
    });
}

int cs3d_set_diagnostics(cs3d_solver *s, int enabled)
{
    CS3D_NEED(s, "cs3d_set_diagnostics");
    s->impl->set_diag(enabled);
    return CS3D_OK;
}

int cs3d_diag_taylor_green(cs3d_solver *s, double out[7])
{
    CS3D_NEED(s, "cs3d_diag_taylor_green");
    if (!out) return set_error(CS3D_ERR_ARG, "cs3d_diag_taylor_green: NULL");
    return guard("cs3d_diag_taylor_green", [&]() -> int {
        int rc = s->impl->diag_tgv(out);
        if (rc == CS3D_ERR_STATE)
            return set_error(rc, "cs3d_diag_taylor_green: enable diagnostics before the step");
        if (rc) return set_error(rc, "cs3d_diag_taylor_green: non-finite state");
        return CS3D_OK;
    });
}

int cs3d_diag_turbulence(cs3d_solver *s, double out[11])
{
    CS3D_NEED(s, "cs3d_diag_turbulence");
    if (!out) return set_error(CS3D_ERR_ARG, "cs3d_diag_turbulence: NULL");
    return guard("cs3d_diag_turbulence", [&]() -> int {
        int rc = s->impl->diag_turb(out);
        if (rc == CS3D_ERR_STATE)
            return set_error(rc, "cs3d_diag_turbulence: enable diagnostics and complete a step first");
        if (rc == CS3D_ERR_NONFINITE)
            return set_error(rc, "cs3d_diag_turbulence: non-finite state");
        return rc;
    });
}

int cs3d_kinetic_energy(cs3d_solver *s, double *ke)
{
    CS3D_NEED(s, "cs3d_kinetic_energy");
    if (!ke) return set_error(CS3D_ERR_ARG, "cs3d_kinetic_energy: NULL");
    return guard("cs3d_kinetic_energy", [&]() -> int {
        int rc = s->impl->kinetic_energy(ke);
        if (rc) return set_error(rc, "cs3d_kinetic_energy: non-finite state");
        return CS3D_OK;
    });
}

int cs3d_rhs_vjp(cs3d_solver *s, int set, const double *lam_rho, const double *lam_rhoU,
                 const double *lam_rhoV, const double *lam_rhoW, const double *lam_rhoE,
                 double *out_rho, double *out_rhoU, double *out_rhoV, double *out_rhoW,
                 double *out_rhoE)
{
    CS3D_NEED(s, "cs3d_rhs_vjp");
    if (set != 1 && set != 2) return set_error(CS3D_ERR_ARG, "cs3d_rhs_vjp: set must be 1 or 2");
    const double *lam[5] = {lam_rho, lam_rhoU, lam_rhoV, lam_rhoW, lam_rhoE};
    double *out[5] = {out_rho, out_rhoU, out_rhoV, out_rhoW, out_rhoE};
    for (int v = 0; v < 5; v++)
        if (!lam[v]) return set_error(CS3D_ERR_ARG, "cs3d_rhs_vjp: NULL cotangent");
    return guard("cs3d_rhs_vjp", [&]() -> int {
        s->impl->rhs_vjp(set, lam, out);
        return CS3D_OK;
    });
}

int cs3d_adjoint_ke_gradient(cs3d_solver *s, int nsteps, int checkpoint, double *J, double *gU,
                             double *gV, double *gW)
{
    CS3D_NEED(s, "cs3d_adjoint_ke_gradient");
    if (nsteps < 1 || checkpoint < 1)
        return set_error(CS3D_ERR_ARG, "cs3d_adjoint_ke_gradient: nsteps and checkpoint must be >= 1");
    return guard("cs3d_adjoint_ke_gradient", [&]() -> int {
        s->impl->ke_gradient(nsteps, checkpoint, J, gU, gV, gW);
        int rc = s->impl->check_finite();
        if (rc) return set_error(rc, "cs3d_adjoint_ke_gradient: non-finite value in the forward pass");
        return CS3D_OK;
    });
}

int cs3d_adjoint_release(cs3d_solver *s)
{
    CS3D_NEED(s, "cs3d_adjoint_release");
    return guard("cs3d_adjoint_release", [&]() -> int {
        s->impl->adjoint_release();
        return CS3D_OK;
    });
}

int cs3d_synchronize(cs3d_solver *s)
{
    CS3D_NEED(s, "cs3d_synchronize");
    return guard("cs3d_synchronize", [&]() -> int {
        s->impl->synchronize();
        return CS3D_OK;
    });
}

void *cs3d_stream(cs3d_solver *s) { return s ? s->impl->stream() : nullptr; }

int cs3d_profile_enable(cs3d_solver *s, int enabled)
{
    CS3D_NEED(s, "cs3d_profile_enable");
    s->impl->profile_enable(enabled);
    return CS3D_OK;
}

int cs3d_profile_report(cs3d_solver *s, double ms[6], int counts[6])
{
    CS3D_NEED(s, "cs3d_profile_report");
    if (!ms || !counts) return set_error(CS3D_ERR_ARG, "cs3d_profile_report: NULL");
    return guard("cs3d_profile_report", [&]() -> int {
        s->impl->profile_report(ms, counts);
        return CS3D_OK;
    });
}

}  // extern "C"
