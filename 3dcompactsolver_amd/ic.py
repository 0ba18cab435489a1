"""Initial-condition generators for the decaying isotropic turbulence and
temporal shear-layer configurations (host side, numpy).

The reference builds these fields outside the solver:

- isotropic turbulence: `homogenous_turbulence/generate_init_hit.f90`
  (random-phase solenoidal field with E(k) ~ (k/k0)^4 exp(-2 (k/k0)^2),
  written with its periodic copies as `init_field_hit.dat`) and
  `MiscTools/loadinithit.py:43-92` (rescaled to u' = sqrt(1/3), written as
  `U/V/W_Mt0p3_N128_k8.dat`, which the driver
  `homogenous_turbulence/Samtaney_D4/3DCompact.cpp_D4:97-130` reads);
- shear layer: `3DCompact.cpp:114-228` (perturbation fields read from files
  that are not in the reference tree, scaled, placed in a band around the
  mid-plane, windowed by exp(-y^2), over a tanh mean profile).

`hit_velocity` restates the generator step by step, including its random
number generator (randset/randvec, a three-LCG generator with a 97-entry
shuffle table); FFTW's unnormalised complex-to-real transform becomes
numpy's irfftn times N^3.  The reference's recorded run of its 128^3 field
(Samtaney_D4/turbdata_N128_D4_k8.out) is reproduced statistically; no
reference output fixes the individual random phases.
`shear_layer_ic` keeps the driver's construction and draws the missing
perturbation fields from seeded white noise.  Fields are numpy arrays
shaped (Nz, Ny, Nx) (XYZ order, x fastest).
"""
import numpy as np

# randset / randvec constants (generate_init_hit.f90:20-27)
_M1, _IA1, _IC1 = 259200, 7141, 54773
_M2, _IA2, _IC2 = 134456, 8121, 28441
_M3, _IA3, _IC3 = 243000, 4561, 51349


def _fmod(a, b):
    """Fortran MOD (sign of the dividend)"""
    r = abs(a) % b
    return r if a >= 0 else -r


class ReferenceRandom:
    """randset(idum) then successive randvec calls (generate_init_hit.f90:
    16-87): uniform numbers in [0, 1] from three linear congruential
    generators and a 97-entry shuffle table."""

    def __init__(self, idum=0):
        rm1, rm2 = 1.0 / _M1, 1.0 / _M2
        ix1 = _fmod(_IC1 - idum, _M1)
        ix1 = _fmod(_IA1 * ix1 + _IC1, _M1)
        ix2 = _fmod(ix1, _M2)
        ix1 = _fmod(_IA1 * ix1 + _IC1, _M1)
        ix3 = _fmod(ix1, _M3)
        r = []
        for _ in range(97):
            ix1 = _fmod(_IA1 * ix1 + _IC1, _M1)
            ix2 = _fmod(_IA2 * ix2 + _IC2, _M2)
            r.append(min(max((ix1 + ix2 * rm2) * rm1, 0.0), 1.0))
        self.r, self.ix = r, [ix1, ix2, ix3]

    def vec(self, n):
        """randvec: the next n numbers"""
        rm1, rm2 = 1.0 / _M1, 1.0 / _M2
        r = self.r
        ix1, ix2, ix3 = self.ix
        out = np.empty(n)
        for m in range(n):
            # every operand is non-negative here, so % is Fortran's MOD
            ix1 = (_IA1 * ix1 + _IC1) % _M1
            ix2 = (_IA2 * ix2 + _IC2) % _M2
            ix3 = (_IA3 * ix3 + _IC3) % _M3
            j = ((97 * ix3) // _M3) % 97
            out[m] = r[j]
            r[j] = min(max((ix1 + ix2 * rm2) * rm1, 0.0), 1.0)
        self.ix = [ix1, ix2, ix3]
        return out


def hit_velocity(N=128, k0=8.0, seed=0, spectral_divergence_free=True):
    """The velocity field of generate_init_hit.f90:150-365 on an N^3 periodic
    box of side 2 pi (grid spacing 2 pi / N), N even.  `seed` is randset's
    idum (the reference calls randset(0)).  Returns U, V, W shaped
    (N, N, N).  spectral_divergence_free=False takes the reference's
    divergence_flag = 1 (modified wavenumbers of the second-order central
    difference in the solenoidal projection)."""
    if N % 2:
        raise ValueError("the generator needs an even number of points")
    kxm = N // 2 + 1
    twopi = 2.0 * np.pi
    tol = 1e-4
    rng = ReferenceRandom(seed)
    # phi(i, j, k) in Fortran order (i fastest): C arrays indexed [k, j, i]
    phis = [twopi * (rng.vec(kxm * N * N).reshape(N, N, kxm) - 0.5) for _ in range(3)]
    i = np.arange(kxm, dtype=np.float64)
    jy = np.arange(N)
    kk = np.where(jy < kxm, jy, jy - N).astype(np.float64)
    KZ, KY, KX = np.meshgrid(kk, kk, i, indexing="ij")
    K = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    live = K > tol
    Ks = np.where(live, K, 1.0)
    Eq = (Ks / k0) ** 4 * np.exp(-2.0 * (Ks / k0) ** 2)
    amp = np.where(live, np.sqrt(Eq / (twopi * Ks ** 2)), 0.0)
    a = amp * np.exp(1j * phis[0]) * np.cos(phis[2])
    b = amp * np.exp(1j * phis[1]) * np.sin(phis[2])
    if not spectral_divergence_free:
        h = twopi / N
        KX, KY, KZ = (np.sin(x * h) / h for x in (KX, KY, KZ))
    K = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    Kxy = np.sqrt(KX ** 2 + KY ** 2)
    live = K > tol
    flat = Kxy <= tol
    Ks = np.where(live, K, 1.0)
    Kxys = np.where(flat, 1.0, Kxy)
    s2 = np.sqrt(2.0)
    comps = [
        np.where(flat, (a + b) / s2, (a * Ks * KY + b * KX * KZ) / (Ks * Kxys)),
        np.where(flat, (b - a) / s2, (b * KY * KZ - a * Ks * KX) / (Ks * Kxys)),
        -b * Kxy / Ks,
    ]
    out = []
    for uk in comps:
        uk = np.where(live, uk, 0.0).astype(np.complex128)
        # the kx = 0 plane made Hermitian (generate_init_hit.f90:331-341)
        jj = np.arange(kxm, N)
        uk[1:, kxm:, 0] = np.conj(uk[N - np.arange(1, N)][:, N - jj, 0])
        uk[kxm:, 0, 0] = np.conj(uk[N - jj, 0, 0])
        uk[0, kxm:, 0] = np.conj(uk[0, N - jj, 0])
        # odd-ball wavenumbers and the mean removed (:343-349)
        uk[:, :, kxm - 1] = 0.0
        uk[:, kxm - 1, :] = 0.0
        uk[kxm - 1, :, :] = 0.0
        uk[0, 0, 0] = 0.0
        # FFTW's unnormalised c2r
        out.append(np.fft.irfftn(uk, s=(N, N, N), axes=(0, 1, 2)) * float(N) ** 3)
    return out[0], out[1], out[2]


def hit_rescale(U, V, W, uprime=np.sqrt(1.0 / 3.0)):
    """MiscTools/loadinithit.py:28-59: scale the field so that
    u' = sqrt(<|u|^2>/3) takes the given value"""
    cur = np.sqrt(np.mean(U * U + V * V + W * W) / 3.0)
    f = uprime / cur
    return U * f, V * f, W * f


def hit_initial_conditions(N=128, k0=8.0, seed=0, Mt=0.3, gamma=1.4):
    """rho, U, V, W, p of the decaying-turbulence runs
    (3DCompact.cpp_D4:118-132): the generated field rescaled to
    <|u|^2> = 1, uniform density 1 and pressure p = 1/(gamma Mt^2) so that
    the turbulent Mach number sqrt(<|u|^2>)/c is Mt (the reference writes
    11.1111111/gamma for Mt = 0.3)."""
    U, V, W = hit_rescale(*hit_velocity(N, k0, seed))
    rho = np.ones_like(U)
    p = np.full_like(U, 1.0 / (Mt * Mt * gamma))
    return rho, U, V, W, p


def write_reference_files(prefix, U, V, W):
    """the driver's input files: one value per line, x fastest
    (loadinithit.py:94-104), named <prefix>U.dat etc."""
    for name, f in (("U", U), ("V", V), ("W", W)):
        np.savetxt(f"{prefix}{name}.dat", f.reshape(-1), fmt="%.17g")


def shear_layer_ic(Nx=512, Ny=128, Nz=42, Ly=128.0, delta_u=1.2, delta0=1.0, intensity=None,
                   band=None, seed=0, gamma=1.4):
    """rho, U, V, W, p of the temporal shear layer (3DCompact.cpp:114-228):
    U = (delta_u / 2) tanh(-(y - Ly/2) / delta0) plus velocity perturbations
    of rms `intensity` (default 0.1 delta_u) windowed by exp(-(y - Ly/2)^2)
    in a band of `band` points (default Nz, as the driver's perturbation
    files) around the mid-plane; density 1, pressure 1/gamma.  The driver
    reads its perturbations from files absent from the reference tree; here
    they are seeded white noise (numpy PCG64).  y_j = j Ly / (Ny - 1)
    (Domain.hpp:30-45)."""
    if intensity is None:
        intensity = 0.1 * delta_u
    if band is None:
        band = min(Nz, Ny)
    rng = np.random.default_rng(seed)
    y = np.arange(Ny) * (Ly / (Ny - 1.0))
    Y = np.broadcast_to(y[None, :, None], (Nz, Ny, Nx))
    start = Ny // 2 - band // 2
    fl = []
    for _ in range(3):
        f = np.zeros((Nz, Ny, Nx))
        f[:, start:start + band, :] = rng.standard_normal((Nz, band, Nx))
        fl.append(f)
    rms = np.sqrt(np.mean(sum(f[:, start:start + band, :] ** 2 for f in fl)) / 3.0)
    win = np.exp(-(Y - Ly / 2.0) ** 2)
    u1, v1, w1 = (intensity / rms * f * win for f in fl)
    U = 0.5 * delta_u * np.tanh(-(Y - Ly / 2.0) / delta0) + u1
    rho = np.ones((Nz, Ny, Nx))
    p = np.full((Nz, Ny, Nx), 1.0 / gamma)
    return rho, U, v1, w1, p
