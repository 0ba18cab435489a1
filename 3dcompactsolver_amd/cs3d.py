"""ctypes binding of lib3dcompact_hip.so (include/cs3d.h).

Host-side mirror of the reference interfaces for tests, the benchmark and
Python users: `Derivatives` / `Filter` (Derivatives.hpp:54-184,
Filter.hpp:11-117) and `Solver` (the AbstractCSolver / AbstractRK loop of
CSolver_AWS.hpp:17-234 and RK4.hpp / TVDRK3.hpp).  Every call goes through the
C ABI into the HIP library; there is no CPU fallback: loading fails loudly when
the library is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib3dcompact_hip.so")

# enums of include/cs3d.h
OK, ERR_ARG, ERR_HIP, ERR_UNSUPPORTED, ERR_STATE, ERR_NONFINITE = range(6)
PERIODIC_SOLVE, DIRICHLET_SOLVE = 0, 1
BC_PERIODIC, BC_SPONGE, BC_ADIABATIC_WALL, BC_MOVING_ADIABATIC_WALL, BC_INLET = range(5)
DIRX, DIRY, DIRZ = 0, 1, 2
FILTER, DERIV1, DERIV2 = 0, 1, 2
FILTER_T, DERIV1_T, DERIV2_T = 3, 4, 5   # transposed operators (the adjoint's)
FP64, FP32 = 0, 1
TVDRK3, RK4 = 3, 4
HOST, DEVICE = 0, 1
LAYOUT_XYZ, LAYOUT_REFERENCE = 0, 1

_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_int = ctypes.c_int


class Config(ctypes.Structure):
    """cs3d_config"""
    _fields_ = [
        ("Nx", _int), ("Ny", _int), ("Nz", _int),
        ("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("Lz", ctypes.c_double),
        ("bc_type", _int * 3), ("bc_kind", _int * 6),
        ("cfl", ctypes.c_double), ("filter_step", _int), ("max_steps", _int),
        ("max_time", ctypes.c_double), ("alpha_f", ctypes.c_double),
        ("mu_ref", ctypes.c_double), ("gamma", ctypes.c_double), ("Pr", ctypes.c_double),
        ("rk", _int), ("precision", _int), ("compensated", _int),
        ("sponge_frac", ctypes.c_double * 3), ("sponge_strength", ctypes.c_double),
        ("sponge_avgT", ctypes.c_double), ("sponge_epsP", ctypes.c_double),
        ("dt_fixed", ctypes.c_double),
        ("wall_vel", (ctypes.c_double * 3) * 6),
    ]


class CS3DError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cs3d error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    L = ctypes.CDLL(LIB_PATH)
    L.cs3d_last_error.restype = ctypes.c_char_p
    L.cs3d_version.restype = ctypes.c_char_p
    L.cs3d_config_default.argtypes = [ctypes.POINTER(Config), _int, _int, _int]
    L.cs3d_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(_vp)]
    L.cs3d_destroy.argtypes = [_vp]
    L.cs3d_destroy.restype = None
    L.cs3d_set_initial_conditions.argtypes = [_vp] + [_dp] * 5
    for f in ("cs3d_pre_step", "cs3d_post_step", "cs3d_synchronize"):
        getattr(L, f).argtypes = [_vp]
    L.cs3d_sub_step.argtypes = [_vp, _int, _int]
    L.cs3d_step.argtypes = [_vp, _int]
    L.cs3d_rhs.argtypes = [_vp, _int]
    L.cs3d_get_conserved.argtypes = [_vp, _int] + [_dp] * 5
    L.cs3d_set_conserved.argtypes = [_vp, _int] + [_dp] * 5
    L.cs3d_get_primitives.argtypes = [_vp] + [_dp] * 8
    L.cs3d_get_time.argtypes = [_vp, _dp, _dp, ctypes.POINTER(_int)]
    L.cs3d_diag_taylor_green.argtypes = [_vp, _dp]
    L.cs3d_diag_turbulence.argtypes = [_vp, _dp]
    L.cs3d_kinetic_energy.argtypes = [_vp, _dp]
    L.cs3d_rhs_vjp.argtypes = [_vp, _int] + [_dp] * 10
    L.cs3d_adjoint_ke_gradient.argtypes = [_vp, _int, _int] + [_dp] * 4
    L.cs3d_adjoint_release.argtypes = [_vp]
    L.cs3d_diag_shear_layer.argtypes = [_vp, _int, _dp]
    L.cs3d_diag_shear_layer_noalloc.argtypes = [_vp, _int, _dp]
    L.cs3d_aero_optics.argtypes = [_vp, _int, _int, ctypes.c_double, _dp]
    L.cs3d_aero_optics_field.argtypes = [_int, _int, _int] + [ctypes.c_double] * 3 + [
        _dp, _int, ctypes.c_double, _dp]
    L.cs3d_get_run_state.argtypes = [_vp, ctypes.POINTER(RunState)]
    L.cs3d_set_run_state.argtypes = [_vp, ctypes.POINTER(RunState)]
    L.cs3d_stream_probe.argtypes = [_int, ctypes.c_size_t, _int, _dp, _dp]
    L.cs3d_stream.argtypes = [_vp]
    L.cs3d_stream.restype = _vp
    L.cs3d_ops_create.argtypes = [_int, _int, _int, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, _int, _int, ctypes.c_double, _int,
                                  ctypes.POINTER(_vp)]
    L.cs3d_ops_apply.argtypes = [_vp, _int, _int, _vp, _vp, _int]
    L.cs3d_ops_synchronize.argtypes = [_vp]
    L.cs3d_ops_destroy.argtypes = [_vp]
    L.cs3d_ops_destroy.restype = None
    L.cs3d_set_diagnostics.argtypes = [_vp, _int]
    L.cs3d_profile_enable.argtypes = [_vp, _int]
    L.cs3d_profile_report.argtypes = [_vp, _dp, ctypes.POINTER(_int)]
    L.cs3d_plan_info.argtypes = [_int, _int, _int, ctypes.c_double, ctypes.c_double, _int,
                                 ctypes.POINTER(_int), _dp, ctypes.c_size_t,
                                 ctypes.POINTER(_int), ctypes.c_size_t]
    _lib = L
    return L


def check(code):
    if code != OK:
        raise CS3DError(code, lib().cs3d_last_error().decode())


def _dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


# ------------------------------------------------------------ operators
class _Ops:
    def __init__(self, dims, lengths, bc_type, direction, alpha_f, precision):
        self.dims = tuple(dims)
        self.precision = precision
        self.dtype = np.float32 if precision == FP32 else np.float64
        h = _vp()
        Nx, Ny, Nz = dims
        Lx, Ly, Lz = lengths
        check(lib().cs3d_ops_create(Nx, Ny, Nz, Lx, Ly, Lz, bc_type, direction, alpha_f,
                                    precision, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().cs3d_ops_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def apply(self, kind, data, layout=LAYOUT_XYZ):
        a = np.ascontiguousarray(data, dtype=self.dtype)
        out = np.empty_like(a)
        check(lib().cs3d_ops_apply(self.h, kind, layout, a.ctypes.data_as(_vp),
                                   out.ctypes.data_as(_vp), HOST))
        return out

    def apply_device(self, kind, in_ptr, out_ptr, layout=LAYOUT_XYZ):
        check(lib().cs3d_ops_apply(self.h, kind, layout, _vp(in_ptr), _vp(out_ptr), DEVICE))


class Derivatives(_Ops):
    """Derivatives(dom, bcType, currentDir) (Derivatives.hpp:54): compact
    6th-order first/second derivatives of a whole field along one direction.
    `dims` = (Nx, Ny, Nz), `lengths` = (Lx, Ly, Lz) as in Domain.hpp."""

    def __init__(self, dims, lengths, bc_type, direction, precision=FP64):
        super().__init__(dims, lengths, bc_type, direction, 0.0, precision)

    def calc1stDerivField(self, data, layout=LAYOUT_XYZ):
        return self.apply(DERIV1, data, layout)

    def calc2ndDerivField(self, data, layout=LAYOUT_XYZ):
        return self.apply(DERIV2, data, layout)

    # transposed operators: <D f, g> = <f, D^T g> over the field
    def calc1stDerivFieldTransposed(self, data, layout=LAYOUT_XYZ):
        return self.apply(DERIV1_T, data, layout)

    def calc2ndDerivFieldTransposed(self, data, layout=LAYOUT_XYZ):
        return self.apply(DERIV2_T, data, layout)


class Filter(_Ops):
    """Filter(alphaF, dom, bcType, currentDir) (Filter.hpp:31): 8th-order
    compact low-pass filter of a whole field along one direction."""

    def __init__(self, alpha_f, dims, lengths, bc_type, direction, precision=FP64):
        super().__init__(dims, lengths, bc_type, direction, alpha_f, precision)

    def filterField(self, data, layout=LAYOUT_XYZ):
        return self.apply(FILTER, data, layout)

    def filterFieldTransposed(self, data, layout=LAYOUT_XYZ):
        return self.apply(FILTER_T, data, layout)


# ------------------------------------------------------------ solver
def config(Nx, Ny, Nz, **kw):
    c = Config()
    check(lib().cs3d_config_default(ctypes.byref(c), Nx, Ny, Nz))
    for k, v in kw.items():
        if k in ("bc_type", "bc_kind", "sponge_frac"):
            arr = getattr(c, k)
            for i, x in enumerate(v):
                arr[i] = x
        elif k == "wall_vel":
            # {face: (U, V, W)} or a 6 x 3 sequence, faces X0, X1, Y0, Y1, Z0, Z1
            items = v.items() if isinstance(v, dict) else enumerate(v)
            for f, uvw in items:
                for m, x in enumerate(uvw):
                    c.wall_vel[f][m] = x
        else:
            setattr(c, k, v)
    return c


class RunState(ctypes.Structure):
    """cs3d_run_state"""
    _fields_ = [("timeStep", _int), ("filterTimeStep", _int), ("time", ctypes.c_double),
                ("dt", ctypes.c_double), ("cfl_max", ctypes.c_double),
                ("rho_ref", ctypes.c_double), ("rhoE_ref", ctypes.c_double), ("stale", _int)]


DUMP_NAMES = ("rho", "rhoU", "rhoV", "rhoW", "rhoE")


class Solver:
    """The GPU solver behind the AbstractCSolver / AbstractRK interfaces.
    Fields are numpy arrays shaped (Nz, Ny, Nx) (XYZ order)."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self.shape = (cfg.Nz, cfg.Ny, cfg.Nx)
        h = _vp()
        check(lib().cs3d_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().cs3d_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_initial_conditions(self, rho, U, V, W, p):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, U, V, W, p)]
        check(lib().cs3d_set_initial_conditions(self.h, *[_dptr(a) for a in arrs]))

    def pre_step(self):
        check(lib().cs3d_pre_step(self.h))

    def sub_step(self, rk_step, rk_last):
        check(lib().cs3d_sub_step(self.h, rk_step, 1 if rk_last else 0))

    def post_step(self):
        check(lib().cs3d_post_step(self.h))

    def step(self, n=1):
        check(lib().cs3d_step(self.h, n))

    def rhs(self, rk_step=1):
        check(lib().cs3d_rhs(self.h, rk_step))

    def synchronize(self):
        check(lib().cs3d_synchronize(self.h))

    def conserved(self, set_=1):
        out = [np.empty(self.shape) for _ in range(5)]
        check(lib().cs3d_get_conserved(self.h, set_, *[_dptr(a) for a in out]))
        return out

    def set_conserved(self, arrays, set_=1):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in arrays]
        check(lib().cs3d_set_conserved(self.h, set_, *[_dptr(a) for a in arrs]))

    def primitives(self):
        out = [np.empty(self.shape) for _ in range(8)]
        check(lib().cs3d_get_primitives(self.h, *[_dptr(a) for a in out]))
        return out

    def time(self):
        t, dt, n = ctypes.c_double(), ctypes.c_double(), _int()
        check(lib().cs3d_get_time(self.h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)))
        return t.value, dt.value, n.value

    def set_diagnostics(self, enabled=True):
        check(lib().cs3d_set_diagnostics(self.h, 1 if enabled else 0))

    def diag_taylor_green(self):
        out = np.zeros(7)
        check(lib().cs3d_diag_taylor_green(self.h, _dptr(out)))
        return out

    def diag_turbulence(self):
        out = np.zeros(11)
        check(lib().cs3d_diag_turbulence(self.h, _dptr(out)))
        return out

    PROFILE_CLASSES = ("x_sweep", "y_sweep", "z_sweep", "filter", "reduce", "other")

    def profile(self, enabled=True):
        """start (or stop) per-launch HIP-event timing on the solver stream"""
        check(lib().cs3d_profile_enable(self.h, 1 if enabled else 0))

    def profile_report(self):
        """{class: (total_ms, launches)} since profile() / the last report"""
        ms = np.zeros(6)
        cnt = (_int * 6)()
        check(lib().cs3d_profile_report(self.h, _dptr(ms), cnt))
        return {n: (float(ms[i]), int(cnt[i])) for i, n in enumerate(self.PROFILE_CLASSES)}

    SHEAR_FILES = ("rhoprime2", "uprime2", "vprime2", "wprime2", "pprime2", "tprime2",
                   "uvprime", "uwprime", "vwprime", "uAvg", "rhoAvg")

    def diag_shear_layer(self, fix_q12=False):
        """{file name: Ny values} of shearLayerInfoCalc for the current state"""
        out = np.zeros(11 * self.cfg.Ny)
        check(lib().cs3d_diag_shear_layer(self.h, 1 if fix_q12 else 0, _dptr(out)))
        return {n: out[i * self.cfg.Ny:(i + 1) * self.cfg.Ny].copy()
                for i, n in enumerate(self.SHEAR_FILES)}

    def aero_optics(self, num_angles=5, max_angle=70.0, set_=1):
        """optical-path integrals of the density (AeroOptics::computeAO),
        shaped (num_angles, Nz, Nx)"""
        out = np.zeros((num_angles, self.cfg.Nz, self.cfg.Nx))
        check(lib().cs3d_aero_optics(self.h, set_, num_angles, max_angle, _dptr(out)))
        return out

    def run_state(self):
        rs = RunState()
        check(lib().cs3d_get_run_state(self.h, ctypes.byref(rs)))
        return rs

    def set_run_state(self, rs):
        check(lib().cs3d_set_run_state(self.h, ctypes.byref(rs)))

    def dump(self, directory):
        """dumpSolution's files (CSolver_AWS.cpp:2493-2535: raw FP64 XYZ
        rho.out.<step> ... rhoE.out.<step>) plus restart.<step> (the run
        state as text) and, with sponges, the running averages and the
        stale-plane copy (spongeAvg*.out.<step>, spongeStale*.out.<step>).
        Returns the step."""
        rs = self.run_state()
        n = rs.timeStep
        sets = [(1, "")]
        if any(self.cfg.bc_kind[f] == BC_SPONGE for f in range(6)):
            sets += [(5, "spongeAvg"), (6, "spongeStale")]
        for set_, pre in sets:
            for name, a in zip(DUMP_NAMES, self.conserved(set_)):
                fn = f"{pre}{name}.out.{n}" if pre else f"{name}.out.{n}"
                a.astype(np.float64).tofile(os.path.join(directory, fn))
        with open(os.path.join(directory, f"restart.{n}"), "w") as f:
            for k, _ in RunState._fields_:
                v = getattr(rs, k)
                f.write(f"{k} {v!r}\n" if isinstance(v, float) else f"{k} {v}\n")
        return n

    def restart(self, directory, step):
        """continue from the files dump(directory) wrote at `step`"""
        rd = lambda fn: np.fromfile(os.path.join(directory, fn), dtype=np.float64).reshape(self.shape)
        q = [rd(f"{name}.out.{step}") for name in DUMP_NAMES]
        rs = RunState()
        with open(os.path.join(directory, f"restart.{step}")) as f:
            for line in f:
                k, v = line.split()
                setattr(rs, k, float(v) if k in ("time", "dt", "cfl_max", "rho_ref", "rhoE_ref")
                        else int(v))
        g = self.cfg.gamma
        rho = q[0]
        U, V, W = q[1] / rho, q[2] / rho, q[3] / rho
        p = (g - 1.0) * (q[4] - 0.5 * rho * (U * U + V * V + W * W))
        self.set_initial_conditions(rho, U, V, W, p)
        self.set_run_state(rs)
        self.set_conserved(q, 1)
        if any(self.cfg.bc_kind[f] == BC_SPONGE for f in range(6)):
            for set_, pre in ((5, "spongeAvg"), (6, "spongeStale")):
                self.set_conserved([rd(f"{pre}{name}.out.{step}") for name in DUMP_NAMES], set_)
        self.set_run_state(rs)  # setting set 1 recomputed the CFL maximum

    def rhs_vjp(self, lam, set_=1):
        """(d(dt*RHS)/dQ)^T lam at conserved set `set_`: five fields in, five out"""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in lam]
        out = [np.empty(self.shape) for _ in range(5)]
        check(lib().cs3d_rhs_vjp(self.h, set_, *[_dptr(a) for a in arrs],
                                 *[_dptr(a) for a in out]))
        return out

    def ke_gradient(self, nsteps, checkpoint=8, outputs=True):
        """advance nsteps; return (J, dJ/dU0, dJ/dV0, dJ/dW0) for the final
        volume-averaged kinetic energy J (only J when outputs is False)"""
        J = ctypes.c_double()
        g = [np.empty(self.shape) if outputs else None for _ in range(3)]
        check(lib().cs3d_adjoint_ke_gradient(self.h, nsteps, checkpoint, ctypes.byref(J),
                                             *[_dptr(a) for a in g]))
        return (J.value, g[0], g[1], g[2]) if outputs else J.value

    def adjoint_release(self):
        """free the adjoint's device memory"""
        check(lib().cs3d_adjoint_release(self.h))

    def kinetic_energy(self):
        v = ctypes.c_double()
        check(lib().cs3d_kinetic_energy(self.h, ctypes.byref(v)))
        return v.value


def stream_probe(kind="copy", nbytes=1 << 30, reps=5):
    """(best, mean) HBM rate in GB/s (bytes read + written) of the library's
    16-byte-vector copy, triad or column-sweep mix ("mix": 5 arrays read and
    9 written) kernel over arrays of nbytes each"""
    best, mean = ctypes.c_double(), ctypes.c_double()
    check(lib().cs3d_stream_probe({"copy": 0, "triad": 1, "mix": 2}[kind], nbytes, reps,
                                  ctypes.byref(best), ctypes.byref(mean)))
    return best.value, mean.value


def aero_optics_field(rho, lengths, num_angles=5, max_angle=70.0):
    """optical-path integrals of an FP64 density field shaped (Nz, Ny, Nx)"""
    Nz, Ny, Nx = rho.shape
    r = np.ascontiguousarray(rho, dtype=np.float64)
    out = np.zeros((num_angles, Nz, Nx))
    check(lib().cs3d_aero_optics_field(Nx, Ny, Nz, *lengths, _dptr(r), num_angles, max_angle,
                                       _dptr(out)))
    return out
