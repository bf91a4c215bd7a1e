"""Seeded synthetic inputs for the benchmark configurations (host, NumPy).

Spectra: "k^-s spectrum" means a power spectrum proportional to |k|^-s, so
Fourier amplitudes scale as |k|^(-s/2) (the tracer amplitude |k|^-1.5 of the
advection configuration is the k^-3 case).  Wavenumbers are integers in
units of 2 pi / L; phases are uniform random from ``numpy.random.default_rng``.
"""

from __future__ import annotations

import numpy as np


def spectral_field(ny: int, nx: int, power_slope: float, seed: int, kmax: float | None = None,
                   periodic: bool = True) -> np.ndarray:
    """Zero-mean, unit-std random-phase field with power spectrum ~ |k|^-power_slope."""
    rng = np.random.default_rng(seed)
    ky = np.fft.fftfreq(ny) * ny
    kx = np.fft.rfftfreq(nx) * nx
    KY, KX = np.meshgrid(ky, kx, indexing="ij")
    k = np.sqrt(KX**2 + KY**2)
    amp = np.zeros_like(k)
    nz = k > 0
    amp[nz] = k[nz] ** (-0.5 * power_slope)
    if kmax is not None:
        amp[k > kmax] = 0.0
    phase = rng.uniform(0.0, 2.0 * np.pi, size=k.shape)
    spec = amp * np.exp(1j * phase)
    f = np.fft.irfft2(spec, s=(ny, nx))
    f -= f.mean()
    s = f.std()
    return f / s if s > 0 else f


def qg_initial_pv(ny: int, nx: int, rms: float = 1e-5, seed: int = 2, mask=None) -> np.ndarray:
    """Interior PV: k^-3 random-phase field with the given rms (configs 2 and the headline)."""
    q = spectral_field(ny, nx, 3.0, seed) * rms
    if mask is not None:
        q = q * (np.asarray(mask) > 0)
    return q


def basin_mask(ny: int, nx: int, seed: int = 3, percentile: float = 30.0) -> np.ndarray:
    """Irregular coastline mask: smooth k^-4 field > its percentile, largest
    4-connected wet component (config 3's generator)."""
    from scipy import ndimage

    f = spectral_field(ny, nx, 4.0, seed)
    wet = f > np.percentile(f, percentile)
    lab, n = ndimage.label(wet)
    if n > 1:
        sizes = ndimage.sum(wet, lab, index=np.arange(1, n + 1))
        wet = lab == (1 + int(np.argmax(sizes)))
    return wet


def helmholtz_config3(n: int = 2048, seeds=(3, 4, 5)):
    """Masked variable-coefficient Helmholtz problem (config 3)."""
    m = basin_mask(n, n, seeds[0])
    g = spectral_field(n, n, 4.0, seeds[1])
    coeff = np.exp(0.5 * g)
    rhs = spectral_field(n, n, 3.0, seeds[2]) * m
    return m.astype(np.float64), coeff, rhs


def advection_config1(n: int = 256, L: float = 5.12e6, ng: int = 3, seeds=(0, 1), kmax=16.0,
                      cfl: float = 0.4):
    """Periodic WENO5 advection inputs (config 1): h, u, v on (n+2ng)^2, dx, dt."""
    dx = L / n
    q = spectral_field(n, n, 3.0, seeds[0], kmax=kmax)
    psi = spectral_field(n, n, 3.0, seeds[1], kmax=kmax)
    N = n + 2 * ng

    def wrap(a):
        return np.pad(a, ng, mode="wrap")

    h = wrap(q)
    P = wrap(psi)
    # psi at X points (NE corners) by 4-point averaging, then u = -dpsi/dy at U,
    # v = dpsi/dx at V: discretely non-divergent.
    X = np.zeros((N, N))
    X[:-1, :-1] = 0.25 * (P[:-1, :-1] + P[:-1, 1:] + P[1:, :-1] + P[1:, 1:])
    u = np.zeros((N, N))
    v = np.zeros((N, N))
    u[1:, :] = -(X[1:, :] - X[:-1, :]) / dx
    v[:, 1:] = (X[:, 1:] - X[:, :-1]) / dx
    s = max(np.abs(u).max(), np.abs(v).max())
    u /= s
    v /= s
    umax, vmax = np.abs(u).max(), np.abs(v).max()
    dt = cfl / (umax / dx + vmax / dx)
    return h, u, v, dx, dt
