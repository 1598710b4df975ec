"""Seeded initial conditions in the reference's spectral layout
([ky][kz][kx = 0..N/2][3], true Fourier coefficients).

``solenoidal_vorticity``: vorticity = i k x u_hat of a Gaussian random
solenoidal velocity whose shell spectrum is E(k) ~ k^4 exp(-2 (k/k0)^2),
normalised to total energy 1/2 sum |u_hat|^2 (with the kx > 0 modes counted
twice, as kspace::cospectrum weights them) equal to `energy`; modes with
k^2 > kM^2 are left at zero.  This is the initial condition the benchmark
configurations specify; the reference itself generates a different family
(bfps/tools.py:96-144), so this generator has no reference counterpart.
"""
import numpy as np


def wavenumbers(n):
    k1 = np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n).astype(np.float64)
    kx = np.arange(n // 2 + 1, dtype=np.float64)
    return kx, k1, k1


def solenoidal_vorticity(n, seed, k0=3.0, energy=0.5, kM=None, dtype=np.complex64, kcut=None):
    """The curl i k x u of solenoidal_velocity(n, seed, ...), computed here on
    the host.  Modes are drawn for |kx|, |ky|, |kz| <= kc = min(kM,
    ceil(8 k0)) only: beyond that E(k) ~ exp(-2 (k/k0)^2) < exp(-128), i.e.
    zero in double precision, so large grids cost no more than small ones."""
    return _solenoidal(n, seed, k0, energy, kM, dtype, kcut, curl=True)


def solenoidal_velocity(n, seed, k0=3.0, energy=0.5, kM=None, dtype=np.complex64, kcut=None):
    """The Gaussian random solenoidal velocity u_hat itself (same draw as
    solenoidal_vorticity); a solver's compute_vorticity turns it into the
    initial vorticity on the device."""
    return _solenoidal(n, seed, k0, energy, kM, dtype, kcut, curl=False)


def _solenoidal(n, seed, k0, energy, kM, dtype, kcut, curl):
    kM = (n // 2 - 1) if kM is None else kM
    kc = int(min(kM, np.ceil(8 * k0))) if kcut is None else int(kcut)
    rng = np.random.default_rng(seed)
    k1 = np.arange(-kc, kc + 1, dtype=np.float64)
    kx1 = np.arange(0, kc + 1, dtype=np.float64)
    KY, KZ, KX = np.meshgrid(k1, k1, kx1, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    k = np.sqrt(k2)
    amp = np.where((k2 > 0) & (k2 <= kM * kM), k ** 4 * np.exp(-2.0 * (k / k0) ** 2), 0.0)
    amp = np.sqrt(amp / np.maximum(4 * np.pi * k2, 1e-300))  # |u_hat|^2 ~ E(k) / k^2
    u = (rng.standard_normal((3,) + k2.shape) + 1j * rng.standard_normal((3,) + k2.shape)) * amp[None]
    kv = np.stack([KX, KY, KZ])
    kdotu = (kv * u).sum(axis=0)
    u -= kv * np.where(k2 > 0, kdotu / np.maximum(k2, 1e-300), 0.0)[None]  # solenoidal
    nxm = np.where(KX == 0, 1.0, 2.0)
    E = 0.5 * float((nxm * (np.abs(u) ** 2).sum(axis=0)).sum())
    if E > 0:
        u *= np.sqrt(energy / E)
    w = 1j * np.cross(kv, u, axis=0) if curl else u  # vorticity = i k x u_hat
    out = np.zeros((n, n, n // 2 + 1, 3), dtype=dtype)
    iy = np.where(k1 >= 0, k1, k1 + n).astype(np.int64)
    sub = np.moveaxis(w, 0, -1).astype(dtype)
    out[np.ix_(iy, iy, np.arange(kc + 1))] = sub
    return out


def uniform_positions(np_, seed, box=2 * np.pi):
    rng = np.random.default_rng(seed)
    return rng.random((int(np_), 3)) * box
