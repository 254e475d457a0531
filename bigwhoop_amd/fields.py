"""Deterministic synthetic fields for the benchmark and parity workloads.

All generators are numpy-only and seeded (``numpy.random.Generator(PCG64)``),
so the same call gives the same bytes on every machine with this numpy.

* :func:`modes_field` -- sum of sinusoidal modes with integer wavevectors
  (|k| <= kmax), amplitudes U(0.5, 1), phases U(0, 2*pi), plus N(0, sigma)
  noise: the 64^3 CPU-runnable case.
* :func:`grf` -- Gaussian random field with an isotropic energy spectrum
  E(k) ~ k^slope, synthesised by filtering white noise in Fourier space,
  rescaled to the requested mean and standard deviation.
"""
from __future__ import annotations

import numpy as np


def _shape4(shape):
    shape = tuple(int(s) for s in shape)
    if not 1 <= len(shape) <= 4:
        raise ValueError("fields have 1 to 4 axes")
    return shape


def modes_field(shape, seed=0, n_modes=4, kmax=8, noise=1e-3, dtype=np.float64):
    """Sum of ``n_modes`` sinusoids on a periodic grid, axes ordered (x, y, z, t).

    The returned array is C-ordered with x fastest, i.e. numpy shape is the
    reverse of ``shape``.
    """
    shape = _shape4(shape)
    rng = np.random.Generator(np.random.PCG64(seed))
    d = len(shape)
    field = np.zeros(shape[::-1], dtype=np.float64)
    axes = [np.arange(n, dtype=np.float64) / n for n in shape]
    for _ in range(n_modes):
        while True:
            k = rng.integers(-kmax, kmax + 1, size=d)
            if 0 < np.sqrt((k * k).sum()) <= kmax:
                break
        amp = rng.uniform(0.5, 1.0)
        phase = rng.uniform(0.0, 2.0 * np.pi)
        arg = np.zeros(shape[::-1], dtype=np.float64)
        for a in range(d):
            sl = [None] * d
            sl[d - 1 - a] = slice(None)
            arg = arg + 2.0 * np.pi * k[a] * axes[a][tuple(sl)]
        field += amp * np.sin(arg + phase)
    field += rng.normal(0.0, noise, size=field.shape)
    return np.ascontiguousarray(field.astype(dtype))


def grf(shape, slope=-5.0 / 3.0, seed=0, mean=10.0, std=1.0, dtype=np.float64):
    """Isotropic Gaussian random field with energy spectrum E(k) ~ k**slope.

    In d dimensions the spectral density is P(k) = E(k) / k**(d-1), so the
    Fourier amplitude filter is k**((slope - (d - 1)) / 2); the k = 0 mode is
    removed and the result is rescaled to exactly ``mean`` and ``std``.
    """
    shape = _shape4(shape)
    d = len(shape)
    rng = np.random.Generator(np.random.PCG64(seed))
    nshape = shape[::-1]
    white = rng.standard_normal(nshape)
    spec = np.fft.rfftn(white)
    k2 = np.zeros(spec.shape, dtype=np.float64)
    for ax, n in enumerate(nshape):
        if ax == len(nshape) - 1:
            f = np.fft.rfftfreq(n) * n
        else:
            f = np.fft.fftfreq(n) * n
        sl = [None] * len(nshape)
        sl[ax] = slice(None)
        k2 = k2 + (f ** 2)[tuple(sl)]
    kmag = np.sqrt(k2)
    with np.errstate(divide="ignore"):
        filt = np.where(kmag > 0, kmag ** ((slope - (d - 1)) / 2.0), 0.0)
    spec *= filt
    del k2, kmag, filt
    out = np.fft.irfftn(spec, s=nshape, axes=tuple(range(len(nshape))))
    out -= out.mean()
    sd = out.std()
    if sd > 0:
        out *= std / sd
    out += mean
    return np.ascontiguousarray(out.astype(dtype))


def grf_device(n: int, slope=-5.0 / 3.0, seed=1, mean=10.0, std=1.0):
    """3-D n^3 float64 Gaussian random field with E(k) ~ k**slope, made in HBM
    with torch's FFT (seeded white noise filtered by k**((slope - 2) / 2)),
    rescaled to ``mean`` and ``std``. Returns a contiguous CUDA/HIP tensor
    (numpy order z, y, x)."""
    import torch

    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    white = torch.randn((n, n, n), generator=g, device="cuda", dtype=torch.float64)
    spec = torch.fft.rfftn(white)
    del white
    f = torch.fft.fftfreq(n, d=1.0 / n, device="cuda", dtype=torch.float64)
    fr = torch.fft.rfftfreq(n, d=1.0 / n, device="cuda", dtype=torch.float64)
    k2 = f[:, None, None] ** 2 + f[None, :, None] ** 2 + fr[None, None, :] ** 2
    kmag = torch.sqrt(k2)
    del k2
    filt = torch.where(kmag > 0, kmag.clamp_min(1e-30) ** ((slope - 2.0) / 2.0), torch.zeros_like(kmag))
    del kmag
    spec *= filt
    del filt
    out = torch.fft.irfftn(spec, s=(n, n, n))
    del spec
    out -= out.mean()
    out *= std / out.std()
    out += mean
    torch.cuda.synchronize()
    return out.contiguous()


def turbulence_4d(shape, nvar=5, seed=7, rho=0.95):
    """(x, y, z, t) float64 fields of ``nvar`` variables, numpy layout
    [variable][t][z][y][x]: variables 0-2 (u, v, w) from independent
    E(k) ~ k^-5/3 Gaussian random fields, the others (p, rho) from k^-7/3
    fields plus a 5% Gaussian pulse background (centred, width nx/6). In
    time the complex spectral coefficients follow a first-order AR(rho)
    process, so the phases evolve smoothly; every time step is rescaled to
    unit variance and mean 10. Seeded and numpy-only, like :func:`grf`."""
    nx, ny, nz, nt = (list(int(s) for s in shape) + [1, 1, 1])[:4]
    rng = np.random.Generator(np.random.PCG64(seed))
    sshape = (nz, ny, nx)
    kz = np.fft.fftfreq(nz) * nz
    ky = np.fft.fftfreq(ny) * ny
    kx = np.fft.rfftfreq(nx) * nx
    kmag = np.sqrt(kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2)
    cshape = kmag.shape
    az = np.arange(nz, dtype=np.float64) - nz / 2
    ay = np.arange(ny, dtype=np.float64) - ny / 2
    ax = np.arange(nx, dtype=np.float64) - nx / 2
    r2 = az[:, None, None] ** 2 + ay[None, :, None] ** 2 + ax[None, None, :] ** 2
    pulse = np.exp(-r2 / (2 * (nx / 6.0) ** 2))
    out = np.empty((nvar, nt) + sshape, np.float64)
    s = (1.0 - rho * rho) ** 0.5
    for v in range(nvar):
        slope = -5.0 / 3.0 if v < 3 else -7.0 / 3.0
        with np.errstate(divide="ignore"):
            filt = np.where(kmag > 0, kmag ** ((slope - 2.0) / 2.0), 0.0)
        w = rng.standard_normal(cshape) + 1j * rng.standard_normal(cshape)
        for t in range(nt):
            if t:
                w = rho * w + s * (rng.standard_normal(cshape) + 1j * rng.standard_normal(cshape))
            x = np.fft.irfftn(w * filt, s=sshape, axes=(0, 1, 2))
            x -= x.mean()
            sd = x.std()
            if sd > 0:
                x /= sd
            if v >= 3:
                x += 0.05 * pulse * np.abs(x).max()
            out[v, t] = x + 10.0
    return out
