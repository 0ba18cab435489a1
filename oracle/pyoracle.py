"""ctypes bindings for the FP64 CPU oracle (liboracle.so) and, when it was
built, for the reference-compiled checker library (_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product path never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_dp = ctypes.POINTER(ctypes.c_double)

PERIODIC_SOLVE, DIRICHLET_SOLVE = 0, 1
PERIODIC, SPONGE, ADIABATIC_WALL, MOVING_ADIABATIC_WALL = 0, 1, 2, 3


def _ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class OrConfig(ctypes.Structure):
    _fields_ = [
        ("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("Nz", ctypes.c_int),
        ("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("Lz", ctypes.c_double),
        ("bc_type", ctypes.c_int * 3), ("bc_kind", ctypes.c_int * 6),
        ("alpha_f", ctypes.c_double), ("mu_ref", ctypes.c_double),
        ("gamma", ctypes.c_double), ("Pr", ctypes.c_double),
        ("cfl", ctypes.c_double), ("filter_step", ctypes.c_int),
        ("rk", ctypes.c_int), ("sponge_frac", ctypes.c_double * 3),
        ("sponge_strength", ctypes.c_double), ("sponge_avgT", ctypes.c_double),
        ("sponge_epsP", ctypes.c_double),
        ("wall_vel", (ctypes.c_double * 3) * 6),
    ]


_lib = None
_fast = None


def load_path(path):
    """an oracle library built elsewhere (the mutation tests' copies)"""
    return _load(path)


def lib(fast=False):
    """Load liboracle.so (strict) or liboracle_fast.so (reference flags)."""
    global _lib, _fast
    if fast:
        if _fast is None:
            _fast = _load(os.path.join(_HERE, "liboracle_fast.so"))
        return _fast
    if _lib is None:
        _lib = _load(os.path.join(_HERE, "liboracle.so"))
    return _lib


def _load(path):
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library missing: {path} (run __graft_entry__.build())")
    L = ctypes.CDLL(path)
    L.or_create.restype = ctypes.c_void_p
    L.or_create.argtypes = [ctypes.POINTER(OrConfig)]
    L.or_destroy.argtypes = [ctypes.c_void_p]
    L.or_set_ic.argtypes = [ctypes.c_void_p] + [_dp] * 5
    L.or_step.restype = ctypes.c_double
    L.or_step.argtypes = [ctypes.c_void_p]
    for f in ("or_pre_step", "or_post_step"):
        getattr(L, f).argtypes = [ctypes.c_void_p]
    L.or_sub_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.or_rhs.argtypes = [ctypes.c_void_p, ctypes.c_int]
    for f in ("or_dt", "or_time"):
        getattr(L, f).restype = ctypes.c_double
        getattr(L, f).argtypes = [ctypes.c_void_p]
    L.or_timestep.restype = ctypes.c_int
    L.or_timestep.argtypes = [ctypes.c_void_p]
    L.or_set_dt_override.argtypes = [ctypes.c_void_p, ctypes.c_double]
    L.or_field.restype = _dp
    L.or_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.or_prim.restype = _dp
    L.or_prim.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.or_deriv.restype = _dp
    L.or_deriv.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.or_diag_tgv.argtypes = [ctypes.c_void_p, _dp]
    L.or_diag_turb.argtypes = [ctypes.c_void_p, _dp]
    L.or_sponge_sigma.restype = _dp
    L.or_sponge_sigma.argtypes = [ctypes.c_void_p]
    L.or_sponge_avg.restype = _dp
    L.or_sponge_avg.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.or_solve_tri.argtypes = [_dp] * 5 + [_dp, ctypes.c_int]
    L.or_cyclic.argtypes = [_dp, _dp, _dp, ctypes.c_double, ctypes.c_double,
                            _dp, ctypes.c_int, _dp]
    L.or_line_op.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                             ctypes.c_double, ctypes.c_double, _dp, _dp]
    L.or_line_rhs.argtypes = L.or_line_op.argtypes
    L.or_line_lhs.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_double, _dp, _dp, _dp]
    L.or_rk_update_arrays.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long,
                                      ctypes.POINTER(_dp)]
    L.or_field_op.argtypes = [ctypes.c_int] * 6 + [ctypes.c_double, ctypes.c_double,
                                                   ctypes.c_int, _dp, _dp]
    L.or_calc_neumann.restype = ctypes.c_double
    L.or_calc_neumann.argtypes = [ctypes.c_double] * 6
    L.or_aero_optics.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [
        _dp, ctypes.c_int, ctypes.c_double, _dp]
    return L


# --------------------------------------------------------------- line level
def solve_tri(a, b, c, d):
    n = len(d)
    x = np.zeros(n)
    w = np.zeros(n)
    lib().or_solve_tri(_ptr(a), _ptr(b), _ptr(c), _ptr(d), _ptr(x), _ptr(w), n)
    return x


def cyclic(a, b, c, alpha, beta, r):
    n = len(r)
    x = np.zeros(n)
    lib().or_cyclic(_ptr(a), _ptr(b), _ptr(c), alpha, beta, _ptr(r), n, _ptr(x))
    return x


def line_op(kind, bctype, phi, dh=1.0, alpha_f=0.49):
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    out = np.zeros_like(phi)
    lib().or_line_op(kind, bctype, len(phi), dh, alpha_f, _ptr(phi), _ptr(out))
    return out


def line_rhs(kind, bctype, phi, dh=1.0, alpha_f=0.49):
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    out = np.zeros_like(phi)
    lib().or_line_rhs(kind, bctype, len(phi), dh, alpha_f, _ptr(phi), _ptr(out))
    return out


def line_lhs(kind, bctype, n, alpha_f=0.49):
    a, b, c = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().or_line_lhs(kind, bctype, n, alpha_f, _ptr(a), _ptr(b), _ptr(c))
    return a, b, c


def field_op(kind, direction, bctype, field, dh=1.0, alpha_f=0.49, ref_layout=False):
    """field: numpy array shaped (Nz, Ny, Nx) in XYZ order (or the
    direction's reference layout when ref_layout, shape given as dims)."""
    f = np.ascontiguousarray(field, dtype=np.float64)
    Nz, Ny, Nx = f.shape
    out = np.zeros_like(f)
    lib().or_field_op(kind, direction, bctype, Nx, Ny, Nz, dh, alpha_f,
                      1 if ref_layout else 0, _ptr(f), _ptr(out))
    return out


def calc_neumann(f):
    """calcNeumann of the 6 values next to a boundary"""
    return lib().or_calc_neumann(*[float(x) for x in f])


def rk_update(rk, rk_step, arrays):
    """arrays: list of 20 float64 arrays (*1, *k, *k2, *2 x 5 variables),
    updated in place."""
    ptrs = (_dp * 20)(*[_ptr(a) for a in arrays])
    lib().or_rk_update_arrays(rk, rk_step, arrays[0].size, ptrs)


def aero_optics(rho, lengths, num_angles=5, max_angle=70.0):
    """AeroOptics::computeAO (AeroOptics.cpp:45-93) of the density rho
    shaped (Nz, Ny, Nx): optical-path integrals shaped (num_angles, Nz, Nx)"""
    Nz, Ny, Nx = rho.shape
    r = np.ascontiguousarray(rho, dtype=np.float64)
    out = np.zeros((num_angles, Nz, Nx))
    lib().or_aero_optics(Nx, Ny, Nz, *lengths, _ptr(r), num_angles, max_angle, _ptr(out))
    return out


# --------------------------------------------------------------- solver
class Solver:
    """The oracle's solver (FP64, CPU).  Field arrays are (Nz, Ny, Nx)."""

    def __init__(self, cfg: OrConfig, fast=False, library=None):
        self.L = library if library is not None else lib(fast)
        self.cfg = cfg
        self.h = self.L.or_create(ctypes.byref(cfg))
        if not self.h:
            raise RuntimeError("oracle: unsupported configuration")
        self.shape = (cfg.Nz, cfg.Ny, cfg.Nx)
        self.N = cfg.Nx * cfg.Ny * cfg.Nz

    def close(self):
        if self.h:
            self.L.or_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_ic(self, rho, U, V, W, p):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, U, V, W, p)]
        self.L.or_set_ic(self.h, *[_ptr(a) for a in arrs])

    def step(self):
        return self.L.or_step(self.h)

    def pre_step(self):
        self.L.or_pre_step(self.h)

    def sub_step(self, rk_step, rk_last):
        self.L.or_sub_step(self.h, rk_step, 1 if rk_last else 0)

    def post_step(self):
        self.L.or_post_step(self.h)

    def rhs(self, rk_step=1):
        self.L.or_rhs(self.h, rk_step)

    def set_dt(self, dt):
        self.L.or_set_dt_override(self.h, dt)

    @property
    def dt(self):
        return self.L.or_dt(self.h)

    @property
    def time(self):
        return self.L.or_time(self.h)

    @property
    def timestep(self):
        return self.L.or_timestep(self.h)

    def _view(self, p):
        return np.ctypeslib.as_array(p, shape=(self.N,)).reshape(self.shape)

    def field(self, which, set_=1):
        """set_: 1 = state (*1), 2 = stage (*k), 3 = dt*RHS (*k2), 4 = *2"""
        return self._view(self.L.or_field(self.h, set_, which)).copy()

    def conserved(self, set_=1):
        return [self.field(v, set_) for v in range(5)]

    def set_field(self, which, arr, set_=1):
        self._view(self.L.or_field(self.h, set_, which))[...] = arr

    def prim(self, which):
        return self._view(self.L.or_prim(self.h, which)).copy()

    def deriv(self, which):
        return self._view(self.L.or_deriv(self.h, which)).copy()

    def diag_tgv(self):
        out = np.zeros(7)
        self.L.or_diag_tgv(self.h, _ptr(out))
        return out

    def diag_turb(self):
        out = np.zeros(11)
        self.L.or_diag_turb(self.h, _ptr(out))
        return out

    def sponge_avg(self):
        """the five sponge running averages (rho .. rhoE), or None"""
        if not self.L.or_sponge_avg(self.h, 0):
            return None
        return [self._view(self.L.or_sponge_avg(self.h, v)).copy() for v in range(5)]

    def set_sponge_avg(self, arrays):
        for v in range(5):
            self._view(self.L.or_sponge_avg(self.h, v))[...] = arrays[v]

    def sponge_sigma(self):
        p = self.L.or_sponge_sigma(self.h)
        if not p:
            return None
        return self._view(p).copy()


def make_config(Nx, Ny, Nz, **kw):
    c = OrConfig()
    c.Nx, c.Ny, c.Nz = Nx, Ny, Nz
    c.Lx = c.Ly = c.Lz = 1.0
    c.alpha_f = 0.49
    c.mu_ref = 1.0
    c.gamma = 1.4
    c.Pr = 0.7
    c.cfl = 0.5
    c.filter_step = 1
    c.rk = 4
    c.sponge_frac[0], c.sponge_frac[1], c.sponge_frac[2] = 0.25, 0.125, 0.25
    c.sponge_strength = 6.0
    c.sponge_avgT = 15.0
    c.sponge_epsP = 0.005
    for k, v in kw.items():
        if k in ("bc_type", "bc_kind", "sponge_frac"):
            arr = getattr(c, k)
            for i, x in enumerate(v):
                arr[i] = x
        elif k == "wall_vel":
            # {face: (U, V, W)} or a 6 x 3 sequence
            items = v.items() if isinstance(v, dict) else enumerate(v)
            for f, uvw in items:
                for m, x in enumerate(uvw):
                    c.wall_vel[f][m] = x
        else:
            setattr(c, k, v)
    return c


# --------------------------------------------------------------- reference
_ref = None


def ref_built():
    """whether _ref/libref.so exists (without loading it)"""
    return os.path.exists(os.path.join(_HERE, "_ref", "libref.so"))


def ref_lib():
    """The reference's own Utils/Derivatives/Filter compiled into
    _ref/libref.so, or None when it was not built (e.g. on a machine
    without the reference checkout)."""
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        L.ref_solve_tri.argtypes = [_dp] * 5 + [ctypes.c_int]
        L.ref_cyclic.argtypes = [_dp, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                 _dp, ctypes.c_int, _dp]
        L.ref_field_op.argtypes = [ctypes.c_int] * 6 + [ctypes.c_double] * 4 + [_dp, _dp]
        L.ref_domain.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [_dp] * 4
        L.ref_idealgas.argtypes = [ctypes.c_int, ctypes.c_double] + [_dp] * 11
        L.ref_sponge_sigma.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [
            ctypes.POINTER(ctypes.c_int), _dp]
        L.ref_rk_update.argtypes = [ctypes.c_int] * 5 + [ctypes.POINTER(_dp)]
        L.ref_calc_neumann.restype = ctypes.c_double
        L.ref_calc_neumann.argtypes = [ctypes.c_double] * 6
        L.ref_aero_optics.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [
            _dp, ctypes.c_int, ctypes.c_double, _dp]
        _ref = L
    return _ref
