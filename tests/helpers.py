"""Initial conditions and error metrics used by the tests."""
import numpy as np


def relerr(a, b):
    """max-norm error of a against b, relative to max|b| (per field)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.abs(b).max()
    if scale == 0.0:
        return float(np.abs(a).max())
    return float(np.abs(a - b).max() / scale)


def state_errs(got, ref, momentum_scale=False):
    """max-norm errors of the five conserved fields.  Each field is taken
    relative to its own maximum, or, with momentum_scale, the three momenta
    relative to the largest momentum of the three: a component that stays
    near zero by symmetry (rho W of the Taylor-Green vortex) is then judged
    against the flow's momentum, not against its own rounding-level size."""
    errs = [relerr(got[v], ref[v]) for v in range(5)]
    if momentum_scale:
        m = max(np.abs(np.asarray(ref[v], dtype=np.float64)).max() for v in (1, 2, 3))
        for v in (1, 2, 3):
            d = np.asarray(got[v], dtype=np.float64) - np.asarray(ref[v], dtype=np.float64)
            errs[v] = float(np.abs(d).max() / m)
    return errs


def grid(N, L):
    """x_i = i*L/(N-1) (Domain.hpp:30-45)"""
    return np.arange(N) * (L / (N - 1.0))


def tgv_ic(N, L, p0=100.0 / 1.4, rho0=1.0, U0=1.0, isothermal=True):
    """Taylor-Green vortex: u = U0 sin x cos y cos z, v = -U0 cos x sin y cos z,
    w = 0, p = p0 + rho0 U0^2/16 (cos 2x + cos 2y)(cos 2z + 2), rho from an
    isothermal state (rho = rho0 p/p0) or uniform.  Arrays are (Nz, Ny, Nx)."""
    x = grid(N, L)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    U = U0 * np.sin(X) * np.cos(Y) * np.cos(Z)
    V = -U0 * np.cos(X) * np.sin(Y) * np.cos(Z)
    W = np.zeros_like(U)
    p = p0 + rho0 * U0 ** 2 / 16.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
    rho = rho0 * p / p0 if isothermal else np.full_like(U, rho0)
    return rho, U, V, W, p


def smooth(a, passes=2):
    for _ in range(passes):
        for ax in range(a.ndim):
            a = (np.roll(a, 1, ax) + 2 * a + np.roll(a, -1, ax)) / 4
    return a


def random_ic(shape, seed, amp=0.3):
    """smooth random velocity, near-uniform density and pressure"""
    rng = np.random.default_rng(seed)
    U = amp * smooth(rng.standard_normal(shape))
    V = amp * smooth(rng.standard_normal(shape))
    W = amp * smooth(rng.standard_normal(shape))
    rho = 1.0 + 0.01 * smooth(rng.standard_normal(shape))
    p = 1.0 / 1.4 + 0.01 * smooth(rng.standard_normal(shape))
    return rho, U, V, W, p


def shear_ic(Nx, Ny, Nz, Ly, seed, dU=1.0, amp=0.05):
    """tanh shear layer centred in y with a windowed random perturbation"""
    rng = np.random.default_rng(seed)
    y = grid(Ny, Ly)
    Y = np.broadcast_to(y[None, :, None], (Nz, Ny, Nx))
    win = np.exp(-(Y - Ly / 2) ** 2)
    U = 0.5 * dU * np.tanh(-(Y - Ly / 2)) + amp * smooth(rng.standard_normal((Nz, Ny, Nx))) * win
    V = amp * smooth(rng.standard_normal((Nz, Ny, Nx))) * win
    W = amp * smooth(rng.standard_normal((Nz, Ny, Nx))) * win
    rho = np.ones_like(U)
    p = np.full_like(U, 1.0 / 1.4)
    return rho, U, V, W, p
