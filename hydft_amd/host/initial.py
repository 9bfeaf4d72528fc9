"""Initial conditions (initial.f90) and the seeded spectral random fields of
the benchmark configurations.

Reference kinds: uniform | mode | random | gaussian (initial.f90:23-75).
'random' draws uniform deviates from a seeded numpy generator (the reference
uses the compiler's intrinsic generator, which no other compiler reproduces),
then dealiases, removes the mean and normalises to unit rms on the GPU.

Benchmark kinds, generated directly on the kept half spectrum (Hermitian,
masked, zero mean) from a counter-free seeded generator and transformed on
the GPU:
  gauss_spectrum: |dn_k| = A n0 exp(-q^2/2), phases uniform in [0, 2 pi)
  shell_rms:      uniform-amplitude random phases in q_lo < |q| < q_hi,
                  scaled so that the real-space rms equals `rms`.
"""
from __future__ import annotations

import math

import numpy as np

from ..spectral import Spec

TWOPI = 2.0 * math.pi


class InitialError(RuntimeError):
    pass


def _coords(p):
    x = np.arange(p.nx) * (p.lx / p.nx)
    y = np.arange(p.ny) * (p.ly / p.ny)
    z = np.arange(p.nz) * (p.lz / p.nz)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def set_initial_condition(p, n0, kind=None, amp=None, mode=None, width=None, seed=None, spec=None):
    """Density n (nz, ny, nx); mom = 0 and Pi = 0 as in the reference."""
    kind = (kind or p.ic).lower()
    amp = p.ic_amplitude if amp is None else amp
    mode = p.ic_mode if mode is None else mode
    width = p.ic_width if width is None else width
    seed = p.ic_seed if seed is None else seed
    shape = (p.nz, p.ny, p.nx)
    if kind == "uniform":
        d = np.zeros(shape)
    elif kind == "mode":
        kv = [TWOPI * mode[0] / p.lx, TWOPI * mode[1] / p.ly, TWOPI * mode[2] / p.lz]
        if p.ny == 1:
            kv[1] = 0.0
        if p.nz == 1:
            kv[2] = 0.0
        X, Y, Z = _coords(p)
        d = np.cos(kv[0] * X + kv[1] * Y + kv[2] * Z)
    elif kind == "random":
        spec = spec or Spec(p.nx, p.ny, p.nz, p.lx, p.ly, p.lz, p.dealias)
        rng = np.random.default_rng(seed)
        d = spec.dealias(rng.random(shape) - 0.5)
        d = d - spec.mean(d)
        rms = math.sqrt(spec.mean(d * d))
        if rms > 0.0:
            d = d / rms
    elif kind == "gaussian":
        spec = spec or Spec(p.nx, p.ny, p.nz, p.lx, p.ly, p.lz, p.dealias)
        X, Y, Z = _coords(p)
        x = X - 0.5 * p.lx
        y = (Y - 0.5 * p.ly) if p.ny > 1 else 0.0 * Y
        z = (Z - 0.5 * p.lz) if p.nz > 1 else 0.0 * Z
        d = spec.dealias(np.exp(-0.5 * (x * x + y * y + z * z) / width ** 2))
        d = d - spec.mean(d)
    else:
        raise InitialError("unknown initial condition: " + kind)
    n = n0 * (1.0 + amp * d)
    if n.min() <= 0.0:
        raise InitialError("initial density is not positive; reduce the amplitude")
    return n


def _kept_half_spectrum(p):
    """Wavenumber arrays and the 2/3 mask on the (nz, ny, nkx) half spectrum."""
    spec = Spec(p.nx, p.ny, p.nz, p.lx, p.ly, p.lz, p.dealias)
    mask = spec.mask()
    nkx = p.nx // 2 + 1
    kx = TWOPI * np.arange(nkx) / p.lx
    wy = np.where(np.arange(p.ny) <= p.ny // 2, np.arange(p.ny), np.arange(p.ny) - p.ny)
    wz = np.where(np.arange(p.nz) <= p.nz // 2, np.arange(p.nz), np.arange(p.nz) - p.nz)
    ky = TWOPI * wy / p.ly if p.ny > 1 else np.zeros(1)
    kz = TWOPI * wz / p.lz if p.nz > 1 else np.zeros(1)
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    return spec, mask, np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)


def _hermitian_field(spec, amp_k, rng, mask):
    """Real field whose half spectrum has modulus amp_k (masked, zero mean)
    and uniform random phases.  On the kx = 0 plane a mode and its partner
    (0, -ky, -kz) are one Fourier pair of a real field: the phase is drawn for
    the first of the pair (kz > 0, or kz = 0 and ky >= 0, with signed wrapped
    indices) and the partner takes its negative, so the modulus is amp_k
    there too.  Self-partnered modes other than k = 0 lie on Nyquist planes,
    which the mask removes; kx Nyquist is masked likewise."""
    ph = rng.random(amp_k.shape) * TWOPI
    nz, ny = amp_k.shape[0], amp_k.shape[1]
    wy = np.where(np.arange(ny) <= ny // 2, np.arange(ny), np.arange(ny) - ny)
    wz = np.where(np.arange(nz) <= nz // 2, np.arange(nz), np.arange(nz) - nz)
    first = (wz[:, None] > 0) | ((wz[:, None] == 0) & (wy[None, :] >= 0))
    iy = (-np.arange(ny)) % ny
    iz = (-np.arange(nz)) % nz
    p0 = ph[:, :, 0]
    ph[:, :, 0] = np.where(first, p0, -p0[np.ix_(iz, iy)])
    c = amp_k * np.exp(1j * ph) * mask
    c[0, 0, 0] = 0.0
    return spec.bwd(c)


def gauss_spectrum_field(p, n0, amp, seed):
    """dn with per-mode modulus amp n0 exp(-q^2/2), zero mean (config: 128^3 Maxwell)."""
    spec, mask, kabs = _kept_half_spectrum(p)
    rng = np.random.default_rng(seed)
    d = _hermitian_field(spec, amp * n0 * np.exp(-0.5 * kabs ** 2), rng, mask)
    n = n0 + d
    if n.min() <= 0.0:
        raise InitialError("initial density is not positive; reduce the amplitude or enlarge the box")
    return n


def shell_rms_field(p, rms, qlo, qhi, seed):
    """Zero-mean field with random phases on qlo < |q| < qhi and real-space rms `rms`."""
    spec, mask, kabs = _kept_half_spectrum(p)
    rng = np.random.default_rng(seed)
    sel = ((kabs > qlo) & (kabs < qhi)).astype(float)
    d = _hermitian_field(spec, sel, rng, mask)
    s = math.sqrt(float(np.mean(d * d)))
    if s == 0.0:
        raise InitialError("no modes in the requested shell")
    return d * (rms / s)
