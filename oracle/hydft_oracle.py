"""CPU FP64 oracle: a numpy restatement of HyDFT's pseudo-spectral RHS and RK3 loop.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``hydft_amd``) may import,
call or execute this module; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

It follows the reference dataflow statement by statement: real-space state,
every forward/backward transform, every dealias call, the same k = 0 rules.
Arrays are numpy C-order with shape (nz, ny, nx), i.e. the same memory as the
reference's Fortran (nx, ny, nz) column-major arrays; half spectra are
(nz, ny, nx//2+1), which is FFTW's r2c_3d(nz, ny, nx) layout (fft.f90:46).

Pinning (see tests/test_oracle_*.py): the FFT, grid and operators against the
known answers of test_fft.f90 and test_ops.f90; the functional terms against
the linearisation check of test_kernels.f90 and against closed forms; the
Fermi-Dirac functions against test_fermi_dirac.f90 and scipy quadrature; the
stepping against test_conservation.f90, the single-mode, driven-mode and
ground-state tests, and the exact RK3 polynomial of a linear mode.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.fft as _sfft

# Threads of the oracle's transforms (scipy.fft workers; 1: numpy's FFT).  The
# tests run with 1; the CPU baseline of the benchmark sets the host's cores.
FFT_WORKERS = 1

PI = 3.14159265358979323846264338327950288
TWOPI = 2.0 * PI
SQRT2 = 1.41421356237309504880168872420969808
SQRTPI = 1.77245385090551602729816748334114518

# Wray low-storage RK3 coefficients (model.f90:51-52)
RK_A = (32.0 / 60.0, 25.0 / 60.0, 45.0 / 60.0)
RK_B = (0.0, -17.0 / 60.0, -25.0 / 60.0)
NCOMP = 6


class OracleError(RuntimeError):
    """Raised where the reference calls fatal() (kinds.f90:26-30)."""


# --------------------------------------------------------------------------- grid
class Grid:
    """grid_t%init (grid.f90:34-113)."""

    def __init__(self, nx, ny, nz, lx, ly, lz, dealias=True):
        if nx < 2:
            raise OracleError("grid: nx must be >= 2")
        if nx % 2:
            raise OracleError("grid: nx must be even")
        if ny > 1 and ny % 2:
            raise OracleError("grid: ny must be even (or 1)")
        if nz > 1 and nz % 2:
            raise OracleError("grid: nz must be even (or 1)")
        self.nx, self.ny, self.nz = nx, ny, nz
        self.nkx = nx // 2 + 1
        self.ndim = 1
        if ny > 1:
            self.ndim = 2
        if nz > 1:
            self.ndim = 3
        self.lx, self.ly, self.lz = float(lx), float(ly), float(lz)
        self.dx, self.dy, self.dz = lx / nx, ly / ny, lz / nz
        self.volume = float(lx)
        if ny > 1:
            self.volume *= ly
        if nz > 1:
            self.volume *= lz
        self.x = np.arange(nx) * self.dx
        self.y = np.arange(ny) * self.dy
        self.z = np.arange(nz) * self.dz
        self.kx = np.array([TWOPI * i / lx for i in range(self.nkx)])
        self.ky = np.zeros(ny)
        self.kz = np.zeros(nz)
        if ny > 1:
            self.ky = np.array([TWOPI * _wrap(j, ny) / ly for j in range(ny)])
        if nz > 1:
            self.kz = np.array([TWOPI * _wrap(k, nz) / lz for k in range(nz)])
        KZ, KY, KX = np.meshgrid(self.kz, self.ky, self.kx, indexing="ij")
        self.k2 = KX ** 2 + KY ** 2 + KZ ** 2
        self.kabs = np.sqrt(self.k2)
        kxc = TWOPI * (nx // 2) / lx * (2.0 / 3.0)
        kyc = kzc = np.finfo(np.float64).max
        if ny > 1:
            kyc = TWOPI * (ny // 2) / ly * (2.0 / 3.0)
        if nz > 1:
            kzc = TWOPI * (nz // 2) / lz * (2.0 / 3.0)
        mask = np.ones_like(self.k2)
        if dealias:
            mask[(np.abs(KX) >= kxc) | (np.abs(KY) >= kyc) | (np.abs(KZ) >= kzc)] = 0.0
        else:
            mask[:, :, self.nkx - 1] = 0.0
            if ny > 1:
                mask[:, ny // 2, :] = 0.0
            if nz > 1:
                mask[nz // 2, :, :] = 0.0
        self.mask = mask
        self.kmax_dealiased = float(self.kabs[mask > 0].max()) if (mask > 0).any() else 0.0
        self.KX, self.KY, self.KZ = KX, KY, KZ

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    def kdir(self, d):
        return (self.KX, self.KY, self.KZ)[d]

    def coords(self):
        """Node coordinates broadcast to (nz, ny, nx) (grid.f90:63-71)."""
        Z, Y, X = np.meshgrid(self.z, self.y, self.x, indexing="ij")
        return X, Y, Z


def _wrap(m, n):
    """grid.f90:116-123."""
    return m if m <= n // 2 else m - n


# ----------------------------------------------------------------------- spectral
class Spec:
    """spec_t (spectral.f90:19-188) on top of the FFT of fft.f90:52-69."""

    def __init__(self, nx, ny, nz, lx, ly, lz, dealias=True):
        self.g = Grid(nx, ny, nz, lx, ly, lz, dealias)
        self.nfwd = 0
        self.nbwd = 0

    # fft.f90:52-59: forward normalised by 1/(nx ny nz)
    def fwd(self, u):
        self.nfwd += 1
        if FFT_WORKERS > 1:
            return _sfft.rfftn(u, norm="forward", workers=FFT_WORKERS)
        return np.fft.rfftn(u, norm="forward")

    # fft.f90:62-69: backward unnormalised, input preserved
    def bwd(self, uhat):
        self.nbwd += 1
        if FFT_WORKERS > 1:
            return _sfft.irfftn(uhat, s=self.g.shape, axes=(0, 1, 2), norm="forward", workers=FFT_WORKERS)
        return np.fft.irfftn(uhat, s=self.g.shape, axes=(0, 1, 2), norm="forward")

    def grad_hat(self, uhat, idir):
        """spectral.f90:67-90 (idir 0-based)."""
        g = self.g
        w1 = (1j * g.kdir(idir)) * uhat * g.mask
        return self.bwd(w1)

    def grad(self, u):
        """spectral.f90:93-106; returns a list of 3 real fields."""
        w2 = self.fwd(u)
        out = []
        for d in range(3):
            if d >= self.g.ndim:
                out.append(np.zeros(self.g.shape))
            else:
                out.append(self.grad_hat(w2, d))
        return out

    def lap(self, u):
        """spectral.f90:109-116."""
        w2 = self.fwd(u)
        return self.bwd(-self.g.k2 * w2 * self.g.mask)

    def div(self, v):
        """spectral.f90:119-144; v is a list of 3 real fields."""
        g = self.g
        w1 = np.zeros(g.k2.shape, dtype=complex)
        for d in range(g.ndim):
            w2 = self.fwd(v[d])
            w1 = w1 + (1j * g.kdir(d)) * w2
        w1 = w1 * g.mask
        return self.bwd(w1)

    def convolve(self, kernel, uhat):
        """spectral.f90:147-154."""
        return self.bwd(kernel * uhat * self.g.mask)

    def dealias(self, u):
        """spectral.f90:157-163 (returns the dealiased field)."""
        w2 = self.fwd(u)
        return self.bwd(w2 * self.g.mask)

    def mean(self, u):
        """spectral.f90:172-176."""
        return float(np.sum(u)) / (self.g.nx * self.g.ny * self.g.nz)

    def mode(self, u, mx, my, mz):
        """spectral.f90:179-188."""
        w2 = self.fwd(u)
        return complex(w2[mz % self.g.nz, my % self.g.ny, mx])


# -------------------------------------------------------------------- Fermi-Dirac
class FermiDirac:
    """fermi_dirac.f90: 6-column table on alpha in [-30, 300], h = 0.05, cubic
    Hermite interpolation, Sommerfeld / classical tails, safeguarded Newton inverse."""

    AMIN, AMAX, H = -30.0, 300.0, 0.05
    NTAB = int(round((300.0 - (-30.0)) / 0.05)) + 1
    NGL = 20

    def __init__(self):
        self.tab = self._build()

    @staticmethod
    def gauss_legendre(n):
        """quadrature.f90:17-38 (Newton on Legendre polynomials)."""
        x = np.zeros(n)
        w = np.zeros(n)
        for i in range(1, (n + 1) // 2 + 1):
            z = math.cos(PI * (i - 0.25) / (n + 0.5))
            for _ in range(100):
                p1, p2 = 1.0, 0.0
                for j in range(1, n + 1):
                    p3 = p2
                    p2 = p1
                    p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j
                pp = n * (z * p1 - p2) / (z * z - 1.0)
                z1 = z
                z = z1 - p1 / pp
                if abs(z - z1) < 1.0e-15:
                    break
            x[i - 1] = -z
            x[n - i] = z
            w[i - 1] = 2.0 / ((1.0 - z * z) * pp * pp)
            w[n - i] = w[i - 1]
        return x, w

    @staticmethod
    def _fermi_derivs(y):
        """fermi_dirac.f90:78-90, vectorised."""
        g = np.exp(-np.abs(y))
        f = np.where(y > 0.0, g / (1.0 + g), 1.0 / (1.0 + g))
        p1 = f * (1.0 - f)
        p2 = p1 * (1.0 - 2.0 * f)
        p3 = p1 * (1.0 - 6.0 * f + 6.0 * f * f)
        return f, p1, p2, p3

    def _build(self):
        """fd_init (fermi_dirac.f90:30-74), vectorised over alpha."""
        xg, wg = self.gauss_legendre(self.NGL)
        alpha = self.AMIN + np.arange(self.NTAB) * self.H
        xmax = np.maximum(alpha, 0.0) + 45.0
        s = np.zeros((6, self.NTAB))
        for q in range(self.NGL):
            t = 0.5 * (xg[q] + 1.0)
            x = t * t
            f, p1, p2, p3 = self._fermi_derivs(x - alpha)
            e = wg[q] * 0.5 * 2.0
            s[0] += e * f
            s[1] += e * t * t * f
            s[2] += e * t ** 4 * f
            s[3] += e * p1
            s[4] += e * p2
            s[5] += e * p3
        npan = np.ceil(xmax - 1.0).astype(int)
        for ip in range(1, int(npan.max()) + 1):
            act = ip <= npan
            a = 1.0 + (ip - 1)
            b = np.minimum(a + 1.0, xmax)
            xm = 0.5 * (a + b)
            xr = 0.5 * (b - a)
            for q in range(self.NGL):
                x = xm + xr * xg[q]
                f, p1, p2, p3 = self._fermi_derivs(x - alpha)
                e = np.where(act, wg[q] * xr, 0.0)
                t = 1.0 / np.sqrt(x)
                sx = np.sqrt(x)
                s[0] += e * t * f
                s[1] += e * sx * f
                s[2] += e * x * sx * f
                s[3] += e * t * p1
                s[4] += e * t * p2
                s[5] += e * t * p3
        return s

    def _node(self, a):
        i = np.floor((a - self.AMIN) / self.H).astype(np.int64)  # int() of a value >= 0 in the table
        return np.clip(i, 0, self.NTAB - 2)

    def _hermite(self, col, dcol_vals, a):
        """fermi_dirac.f90:94-107 with the derivative columns given by the caller."""
        i = self._node(a)
        t = (a - (self.AMIN + i * self.H)) / self.H
        t2 = t * t
        t3 = t2 * t
        h00 = 2 * t3 - 3 * t2 + 1
        h10 = t3 - 2 * t2 + t
        h01 = -2 * t3 + 3 * t2
        h11 = t3 - t2
        y0 = self.tab[col, i]
        y1 = self.tab[col, i + 1]
        d0 = dcol_vals[0][i]
        d1 = dcol_vals[1][i + 1]
        return h00 * y0 + h10 * self.H * d0 + h01 * y1 + h11 * self.H * d1

    @staticmethod
    def sommerfeld(p, a, m):
        """fermi_dirac.f90:116-133."""
        c = [1.0, (PI ** 2 / 6.0) * p * (p + 1.0),
             (7.0 * PI ** 4 / 360.0) * (p + 1.0) * p * (p - 1.0) * (p - 2.0)]
        v = 0.0
        for j in range(3):
            ex = p + 1.0 - 2 * j
            coef = c[j] / (p + 1.0)
            for l in range(m):
                coef = coef * (ex - l)
            v = v + coef * a ** (ex - m)
        return v

    @staticmethod
    def classical(p, a, m):
        """fermi_dirac.f90:136-147."""
        v = 0.0
        sgn = 1.0
        for j in range(1, 13):
            v = v + sgn * np.exp(j * a) * float(j) ** (m - p - 1.0)
            sgn = -sgn
        return v * math.gamma(p + 1.0)

    def _eval(self, a, p, m, col, dsrc, dscale):
        a = np.asarray(a, dtype=float)
        out = np.empty_like(a)
        hi = a > self.AMAX
        lo = a < self.AMIN
        mid = ~(hi | lo)
        if hi.any():
            out[hi] = self.sommerfeld(p, a[hi], m)
        if lo.any():
            out[lo] = self.classical(p, a[lo], m)
        if mid.any():
            d = dscale * self.tab[dsrc]
            out[mid] = self._hermite(col, (d, d), a[mid])
        return out

    def im12(self, a):
        return self._eval(a, -0.5, 0, 0, 3, 1.0)

    def ip12(self, a):
        return self._eval(a, 0.5, 0, 1, 0, 0.5)

    def ip32(self, a):
        return self._eval(a, 1.5, 0, 2, 1, 1.5)

    def dim12(self, a):
        return self._eval(a, -0.5, 1, 3, 4, 1.0)

    def d2im12(self, a):
        return self._eval(a, -0.5, 2, 4, 5, 1.0)

    def xi(self, a):
        """fermi_dirac.f90:225-230."""
        im = self.im12(a)
        return self.dim12(a) / (im * im)

    def dxi(self, a):
        """fermi_dirac.f90:233-238."""
        im = self.im12(a)
        d1 = self.dim12(a)
        d2 = self.d2im12(a)
        return (d2 * im - 2.0 * d1 * d1) / im ** 3

    def alpha_from_ip12(self, y):
        """fermi_dirac.f90:241-267, vectorised with per-point exit."""
        y = np.asarray(y, dtype=float)
        if (y <= 0.0).any():
            raise OracleError("fermi_dirac: I_{1/2} inverse needs y > 0")
        huge = np.finfo(np.float64).max
        alpha = np.maximum(np.log(2.0 * y / SQRTPI), (1.5 * y) ** (2.0 / 3.0) - 1.0)
        alpha = np.where(y < 1.0, np.log(2.0 * y / SQRTPI), alpha)
        lo = np.full_like(y, -huge)
        hi = np.full_like(y, huge)
        act = np.ones(y.shape, dtype=bool)
        for _ in range(100):
            if not act.any():
                break
            f = self.ip12(alpha) - y
            hi = np.where(act & (f > 0.0), np.minimum(hi, alpha), hi)
            lo = np.where(act & ~(f > 0.0), np.maximum(lo, alpha), lo)
            df = 0.5 * self.im12(alpha)
            step = f / df
            done = np.abs(step) < 1.0e-14 * np.maximum(1.0, np.abs(alpha))
            act = act & ~done
            alpha = np.where(act, alpha - step, alpha)
            bis = act & ((alpha <= lo) | (alpha >= hi)) & (lo > -huge) & (hi < huge)
            alpha = np.where(bis, 0.5 * (lo + hi), alpha)
        return alpha


_FD = None


def fermi_dirac():
    global _FD
    if _FD is None:
        _FD = FermiDirac()
    return _FD


# -------------------------------------------------------------------------- units
@dataclass
class Units:
    """units_t%init (units.f90:41-70)."""
    kind: str = "ion"
    gamma: float = 1.0
    kappa: float = 0.0
    rs: float = 1.0
    mass: float = 1.0
    n0: float = 3.0 / (4.0 * PI)
    e2: float = 1.0 / 3.0
    beta: float = 3.0
    hbar: float = 0.0
    cn: float = 0.0
    alpha0: float = 0.0
    theta: float = float(np.finfo(np.float64).max)
    lambda_qsp: float = 0.0
    lambda_tf: float = 0.0
    quantum: bool = False

    def __post_init__(self):
        self.beta = 3.0 * self.gamma
        if self.kind == "electron":
            fd = fermi_dirac()
            self.quantum = True
            self.kappa = 0.0
            self.hbar = 1.0 / math.sqrt(3.0 * self.rs)
            self.cn = SQRT2 / (PI * PI * self.hbar ** 3 * self.beta ** 1.5) * self.mass ** 1.5
            self.alpha0 = float(fd.alpha_from_ip12(np.array([self.n0 / self.cn]))[0])
            tf = self.hbar ** 2 * (3.0 * PI * PI * self.n0) ** (2.0 / 3.0) / (2.0 * self.mass)
            self.theta = (1.0 / self.beta) / tf
            self.lambda_qsp = 2.0 * math.sqrt(PI * self.gamma / self.rs)
            self.lambda_tf = 1.0 / math.sqrt(4.0 * PI * self.e2 * self.beta * self.cn * 0.5
                                             * float(fd.im12(np.array([self.alpha0]))[0]))
        else:
            self.lambda_tf = 1.0 / math.sqrt(3.0 * self.gamma)


# ---------------------------------------------------------------- free-energy terms
def interp_linear(x, y, xq):
    """utils.f90:30-55, vectorised (clamped binary search, linear extrapolation)."""
    xq = np.asarray(xq, dtype=float)
    n = len(x)
    if n == 1:
        return np.full_like(xq, y[0])
    lo = np.searchsorted(x, xq, side="right") - 1  # largest lo with x[lo] <= xq
    lo = np.where(xq <= x[0], 0, lo)
    lo = np.where(xq >= x[n - 1], n - 2, lo)
    lo = np.clip(lo, 0, n - 2)
    w = (xq - x[lo]) / (x[lo + 1] - x[lo])
    return (1.0 - w) * y[lo] + w * y[lo + 1]


class Term:
    name = ""

    def __init__(self, un: Units):
        self.n0 = un.n0
        self.beta = un.beta

    def setup(self, spc):
        pass


class Ideal(Term):
    """fe_ideal_classical.f90:33-52."""
    name = "ideal"

    def add_mu(self, spc, n, nhat, mu):
        if n.min() <= 0.0:
            raise OracleError("ideal term: density is not positive")
        return mu + np.log(n) / self.beta

    def kernel(self, k):
        return np.full_like(np.asarray(k, float), 1.0 / (self.beta * self.n0))

    def energy(self, spc, n, nhat):
        return spc.mean(n * (np.log(n) - 1.0)) * spc.g.volume / self.beta


class Hartree(Term):
    """fe_hartree.f90:34-74."""
    name = "hartree"

    def __init__(self, un: Units):
        super().__init__(un)
        self.e2 = un.e2
        self.kappa = un.kappa

    def kernel(self, k):
        k = np.asarray(k, float)
        den = k * k + self.kappa ** 2
        return np.where(den > 0.0, 4.0 * PI * self.e2 / np.where(den > 0.0, den, 1.0), 0.0)

    def setup(self, spc):
        self.vk = self.kernel(spc.g.kabs)

    def add_mu(self, spc, n, nhat, mu):
        return mu + spc.convolve(self.vk, nhat)

    def energy(self, spc, n, nhat):
        w = spc.convolve(self.vk, nhat)
        return 0.5 * spc.mean(n * w) * spc.g.volume


class RY(Term):
    """fe_correlation_ry.f90:39-98; (k_tab, c_tab) is the structure table c(k)."""
    name = "ry"

    def __init__(self, un: Units, k_tab, c_tab, subtract_mean_field):
        super().__init__(un)
        self.e2 = un.e2
        self.kappa = un.kappa
        self.k_tab = np.asarray(k_tab, float)
        self.c_tab = np.asarray(c_tab, float)
        self.subtract_mean_field = subtract_mean_field

    def c_of_k(self, k):
        """structure_source.f90:37-41."""
        return interp_linear(self.k_tab, self.c_tab, np.minimum(k, self.k_tab[-1]))

    def c_cor(self, k):
        k1 = np.maximum(np.asarray(k, float), self.k_tab[0])
        c = self.c_of_k(k1)
        if self.subtract_mean_field:
            c = c + self.beta * 4.0 * PI * self.e2 / (k1 * k1 + self.kappa ** 2)
        return c

    def setup(self, spc):
        self.ck = self.c_cor(spc.g.kabs)

    def _no_mean(self, nhat):
        b = nhat.copy()
        b[0, 0, 0] = nhat[0, 0, 0] - self.n0
        return b

    def add_mu(self, spc, n, nhat, mu):
        w = spc.convolve(self.ck, self._no_mean(nhat))
        return mu - w / self.beta

    def kernel(self, k):
        return -self.c_cor(k) / self.beta

    def energy(self, spc, n, nhat):
        w = spc.convolve(self.ck, self._no_mean(nhat))
        return -0.5 * spc.mean((n - self.n0) * w) * spc.g.volume / self.beta


class TF(Term):
    """fe_kinetic_tfk.f90:63-111 (Thomas-Fermi)."""
    name = "tf"

    def __init__(self, un: Units):
        super().__init__(un)
        self.cn = un.cn
        self.alpha0 = un.alpha0

    def alpha_of_n(self, n):
        if n.min() <= 0.0:
            raise OracleError("tf term: density is not positive")
        return fermi_dirac().alpha_from_ip12(n / self.cn)

    def add_mu(self, spc, n, nhat, mu):
        return mu + self.alpha_of_n(n) / self.beta

    def kernel(self, k):
        fd = fermi_dirac()
        v = 2.0 / (self.beta * self.cn * float(fd.im12(np.array([self.alpha0]))[0]))
        return np.full_like(np.asarray(k, float), v)

    def energy(self, spc, n, nhat):
        fd = fermi_dirac()
        a = self.alpha_of_n(n)
        s = np.sum(n * a - (2.0 / 3.0) * self.cn * fd.ip32(a))
        return s / n.size * spc.g.volume / self.beta


class Gradient(Term):
    """fe_kinetic_tfk.f90:114-160 (Kirzhnits gradient correction)."""
    name = "gradient"

    def __init__(self, un: Units, tfk_gamma=1.0 / 9.0):
        super().__init__(un)
        self.cn = un.cn
        self.alpha0 = un.alpha0
        self.gamma_g = tfk_gamma
        self.cg = tfk_gamma * (3.0 * SQRT2 * PI * PI / 8.0) * un.hbar ** 5 * un.beta ** 1.5 / un.mass ** 2.5

    def add_mu(self, spc, n, nhat, mu):
        if n.min() <= 0.0:
            raise OracleError("gradient term: density is not positive")
        fd = fermi_dirac()
        gn = [spc.grad_hat(nhat, d) if d < spc.g.ndim else np.zeros(spc.g.shape) for d in range(3)]
        lapn = spc.convolve(-spc.g.k2, nhat)
        a = fd.alpha_from_ip12(n / self.cn)
        xi = fd.xi(a)
        dxidn = fd.dxi(a) / (0.5 * self.cn * fd.im12(a))
        g2 = gn[0] ** 2 + gn[1] ** 2 + gn[2] ** 2
        return mu - self.cg * (dxidn * g2 + 2.0 * xi * lapn)

    def kernel(self, k):
        fd = fermi_dirac()
        k = np.asarray(k, float)
        return 2.0 * self.cg * float(fd.xi(np.array([self.alpha0]))[0]) * k * k

    def energy(self, spc, n, nhat):
        fd = fermi_dirac()
        gn = [spc.grad_hat(nhat, d) if d < spc.g.ndim else np.zeros(spc.g.shape) for d in range(3)]
        a = fd.alpha_from_ip12(n / self.cn)
        s = np.sum(fd.xi(a) * (gn[0] ** 2 + gn[1] ** 2 + gn[2] ** 2))
        return self.cg * s / n.size * spc.g.volume


class Functional:
    """functional_t (functional.f90:56-90)."""

    def __init__(self, terms):
        self.terms = list(terms)

    def setup(self, spc):
        for t in self.terms:
            t.setup(spc)

    def mu(self, spc, n, nhat):
        mu = np.zeros(spc.g.shape)
        for t in self.terms:
            mu = t.add_mu(spc, n, nhat, mu)
        return mu

    def kernel(self, k):
        kk = np.zeros_like(np.asarray(k, float))
        for t in self.terms:
            kk = kk + t.kernel(k)
        return kk

    def energy(self, spc, n, nhat):
        return sum(t.energy(spc, n, nhat) for t in self.terms)


def build_functional(un: Units, terms: str, k_tab=None, c_tab=None, tfk_gamma=1.0 / 9.0):
    """The term factory of species.f90:97-146."""
    names = [s.strip() for s in terms.split(",") if s.strip()]
    has_hartree = "hartree" in terms
    out = []
    for nm in names:
        if nm == "ideal":
            out.append(Ideal(un))
        elif nm == "tf":
            if not un.quantum:
                raise OracleError("tf term needs kind = electron")
            out.append(TF(un))
        elif nm == "gradient":
            if not un.quantum:
                raise OracleError("gradient term needs kind = electron")
            out.append(Gradient(un, tfk_gamma))
        elif nm == "hartree":
            out.append(Hartree(un))
        elif nm == "ry":
            if k_tab is None:
                raise OracleError("ry term needs a structure source (hnc or file)")
            out.append(RY(un, k_tab, c_tab, has_hartree))
        else:
            raise OracleError("unknown free-energy term: " + nm)
    return Functional(out)


# ------------------------------------------------------------------------- stress
class Stress:
    """stress_t (stress.f90:37-105)."""

    IDX = ((0, 3, 4), (3, 1, 5), (4, 5, 2))  # stress.f90:92, 0-based

    def __init__(self, name, eta_bar, xi_bar, tau, n0):
        name = name.lower()
        if name == "newtonian":
            self.evolve, self.tau = False, 0.0
        elif name == "maxwell":
            self.evolve, self.tau = tau > 0.0, tau
        else:
            raise OracleError("unknown stress closure: " + name)
        self.eta = eta_bar * n0
        self.xi = xi_bar * n0

    def pi0(self, spc, u):
        g = spc.g
        gu = [[None] * 3 for _ in range(3)]
        for i in range(3):
            if i < g.ndim:
                gu[i] = spc.grad(u[i])
            else:
                gu[i] = [np.zeros(g.shape) for _ in range(3)]
        divu = gu[0][0] + gu[1][1] + gu[2][2]
        pi = [None] * 6
        for i in range(3):
            pi[i] = self.eta * (2.0 * gu[i][i] - (2.0 / 3.0) * divu) + self.xi * divu
        pi[3] = self.eta * (gu[0][1] + gu[1][0])
        pi[4] = self.eta * (gu[0][2] + gu[2][0])
        pi[5] = self.eta * (gu[1][2] + gu[2][1])
        return pi

    def divergence(self, spc, pi):
        f = []
        for i in range(3):
            if i >= spc.g.ndim:
                f.append(np.zeros(spc.g.shape))
                continue
            f.append(spc.div([pi[self.IDX[i][j]] for j in range(3)]))
        return f


# ------------------------------------------------------------- external potential
class Vext:
    """vext_t (external_potential.f90:39-93)."""

    def __init__(self, spc, kind="none", amp=0.0, mode=(1, 0, 0), omega=0.0, width=1.0, ramp=0.0):
        g = spc.g
        self.kind = kind.lower()
        self.amp, self.mode, self.omega, self.width, self.ramp = amp, tuple(mode), omega, width, ramp
        self.kvec = [TWOPI * mode[0] / g.lx, TWOPI * mode[1] / g.ly, TWOPI * mode[2] / g.lz]
        if g.ny == 1:
            self.kvec[1] = 0.0
        if g.nz == 1:
            self.kvec[2] = 0.0
        if self.kind == "none":
            self.active, self.time_dependent = False, False
        elif self.kind in ("mode", "gaussian"):
            self.active, self.time_dependent = True, False
        elif self.kind == "driven":
            self.active, self.time_dependent = True, True
        else:
            raise OracleError("unknown external potential: " + kind)
        self.v = np.zeros(g.shape)
        self.f = [np.zeros(g.shape) for _ in range(3)]
        if self.active:
            self.eval(spc, 0.0)

    def eval(self, spc, t):
        if not self.active:
            return
        g = spc.g
        a = self.amp
        if self.time_dependent and self.ramp > 0.0:
            a = a * min(1.0, t / self.ramp)
        X, Y, Z = g.coords()
        if self.kind in ("mode", "driven"):
            ph = self.kvec[0] * X + self.kvec[1] * Y + self.kvec[2] * Z - self.omega * t
            self.v = a * np.cos(ph)
            s = a * np.sin(ph)
            self.f = [s * self.kvec[0], s * self.kvec[1], s * self.kvec[2]]
        elif self.kind == "gaussian":
            x = X - 0.5 * g.lx
            y = (Y - 0.5 * g.ly) if g.ny > 1 else 0.0 * Y
            z = (Z - 0.5 * g.lz) if g.nz > 1 else 0.0 * Z
            r2 = x * x + y * y + z * z
            self.v = -a * np.exp(-0.5 * r2 / self.width ** 2)
            gv = spc.grad(self.v)
            self.f = [-gv[0], -gv[1], -gv[2]]


# -------------------------------------------------------------------------- model
@dataclass
class Model:
    """model_t (model.f90:29-227) with the species state of species.f90:39-41."""
    spc: Spec
    fun: Functional
    stress: Stress
    ext: Vext
    n0: float
    mass: float = 1.0
    dt: float = 0.0
    cfl: float = 0.4
    eta_l: float = 0.0
    t: float = 0.0
    step: int = 0
    n: np.ndarray = None
    mom: list = field(default_factory=list)
    pi: list = field(default_factory=list)

    def __post_init__(self):
        g = self.spc.g
        if self.n is None:
            self.n = np.full(g.shape, self.n0)
        if not self.mom:
            self.mom = [np.zeros(g.shape) for _ in range(3)]
        if self.stress.evolve and not self.pi:
            self.pi = [np.zeros(g.shape) for _ in range(NCOMP)]
        self.dn_o = np.zeros(g.shape)
        self.dmom_o = [np.zeros(g.shape) for _ in range(3)]
        self.dpi_o = [np.zeros(g.shape) for _ in range(NCOMP)]
        self.u = [np.zeros(g.shape) for _ in range(3)]

    def velocity(self):
        """model.f90:76-82."""
        self.u = [self.mom[d] / (self.mass * self.n) for d in range(3)]

    def rhs(self, t):
        """model.f90:85-146."""
        spc, g = self.spc, self.spc.g
        if self.n.min() <= 0.0:
            raise OracleError("density became non-positive; reduce the amplitude or the time step")
        self.velocity()
        nhat = spc.fwd(self.n)
        dn = -spc.div(self.mom) / self.mass
        mu = self.fun.mu(spc, self.n, nhat)
        gmu = spc.grad(mu)
        if self.ext.active and self.ext.time_dependent:
            self.ext.eval(spc, t)
        if self.stress.evolve:
            pi = [p.copy() for p in self.pi]
        else:
            pi = self.stress.pi0(spc, self.u)
        fpi = self.stress.divergence(spc, pi)
        dmom = [None] * 3
        for i in range(3):
            if i >= g.ndim:
                dmom[i] = np.zeros(g.shape)
                continue
            flux = [self.mom[i] * self.u[j] for j in range(3)]
            divf = spc.div(flux)
            divf = divf + self.n * gmu[i]
            dmom[i] = -(divf - spc.mean(divf)) + fpi[i]
            if self.ext.active:
                dmom[i] = dmom[i] + self.n * self.ext.f[i]
        dpi = []
        if self.stress.evolve:
            for c in range(NCOMP):
                gpi = spc.grad(self.pi[c])
                d = np.zeros(g.shape)
                for dd in range(g.ndim):
                    d = d - self.u[dd] * gpi[dd]
                dpi.append(d)
        dn = spc.dealias(dn)
        for i in range(g.ndim):
            dmom[i] = spc.dealias(dmom[i])
        if self.stress.evolve:
            dpi = [spc.dealias(d) for d in dpi]
        return dn, dmom, dpi

    def compute_dt(self):
        """model.f90:150-183."""
        g = self.spc.g
        if self.dt_fixed > 0.0:
            self.dt = self.dt_fixed
            return self.dt
        sel = (g.mask != 0.0) & (g.kabs != 0.0)
        k = g.kabs[sel]
        wmax = 0.0
        if k.size:
            c2 = self.n0 * self.fun.kernel(k) / self.mass
            if self.stress.evolve:
                c2 = c2 + self.eta_l / (self.mass * self.stress.tau)
            kk = k * k * np.maximum(c2, 0.0)
            wmax = float(np.sqrt(kk).max())
        numax = 0.0
        if not self.stress.evolve:
            numax = self.eta_l / self.mass * g.kmax_dealiased ** 2
        self.velocity()
        umax = max(float(np.abs(u).max()) for u in self.u)
        dxmin = g.dx
        if g.ny > 1:
            dxmin = min(dxmin, g.dy)
        if g.nz > 1:
            dxmin = min(dxmin, g.dz)
        huge = np.finfo(np.float64).max
        dtm = huge
        if wmax > 0.0:
            dtm = min(dtm, 1.7 / wmax)
        if numax > 0.0:
            dtm = min(dtm, 2.5 / numax)
        if umax > 0.0:
            dtm = min(dtm, dxmin / umax)
        if dtm == huge:
            raise OracleError("cannot determine a time step; set dt in &time")
        self.dt = self.cfl * dtm
        return self.dt

    dt_fixed: float = 0.0

    def advance(self):
        """model.f90:186-227."""
        self._ts = self.t
        for s in range(3):
            self.advance_stage(s)
        self.t = self.t + self.dt
        self.step += 1

    def advance_stage(self, s):
        """Stage s of model_advance (model.f90:192-223); the stage time is self._ts."""
        spc, g = self.spc, self.spc.g
        dt = self.dt
        if s == 0:
            self._ts = self.t
        ts = self._ts
        dn, dmom, dpi = self.rhs(ts)
        if self.stress.evolve:
            pi0_old = self.stress.pi0(spc, self.u)
        if s == 0:
            self.n = self.n + dt * RK_A[0] * dn
            self.mom = [self.mom[i] + dt * RK_A[0] * dmom[i] for i in range(3)]
            if self.stress.evolve:
                self.pi = [self.pi[c] + dt * RK_A[0] * dpi[c] for c in range(NCOMP)]
        else:
            self.n = self.n + dt * (RK_A[s] * dn + RK_B[s] * self.dn_o)
            self.mom = [self.mom[i] + dt * (RK_A[s] * dmom[i] + RK_B[s] * self.dmom_o[i]) for i in range(3)]
            if self.stress.evolve:
                self.pi = [self.pi[c] + dt * (RK_A[s] * dpi[c] + RK_B[s] * self.dpi_o[c])
                           for c in range(NCOMP)]
        self.dn_o = dn
        self.dmom_o = dmom
        if self.stress.evolve:
            self.dpi_o = dpi
        dts = (RK_A[s] + RK_B[s]) * dt
        self._ts = ts + dts
        if self.stress.evolve:
            self.velocity()
            pi0 = self.stress.pi0(spc, self.u)
            fac = math.exp(-dts / self.stress.tau)
            phi = (self.stress.tau / dts) * (1.0 - fac)
            self.pi = [pi0[c] - (pi0[c] - pi0_old[c]) * phi + (self.pi[c] - pi0_old[c]) * fac
                       for c in range(NCOMP)]
            self.pi = [spc.dealias(p) for p in self.pi]
        self.n = spc.dealias(self.n)
        for i in range(g.ndim):
            self.mom[i] = spc.dealias(self.mom[i])

    # diagnostics.f90:41-77
    def budgets(self):
        spc, g = self.spc, self.spc.g
        v = g.volume
        mass = spc.mean(self.n) * v
        mom = [spc.mean(self.mom[i]) * v for i in range(3)]
        ekin = 0.0
        for i in range(g.ndim):
            ekin += spc.mean(self.mom[i] ** 2 / self.n)
        ekin = 0.5 * ekin * v / self.mass
        nhat = spc.fwd(self.n)
        fe = self.fun.energy(spc, self.n, nhat)
        if self.ext.active:
            fe += spc.mean(self.n * self.ext.v) * v
        return mass, mom, ekin, fe

    def probe(self, mode=(1, 0, 0)):
        a = self.spc.mode(self.n, *mode) / self.n0
        b = self.spc.mode(self.mom[0], *mode)
        return a, b


def make_model(nx, ny, nz, lx, ly, lz, *, kind="ion", gamma=2.0, kappa=0.1, rs=1.86, mass=1.0,
               terms=None, k_tab=None, c_tab=None, tfk_gamma=1.0 / 9.0, closure="newtonian",
               eta=0.0, xi=0.0, tau=0.0, eta_l=None, dt=0.0, cfl=0.4, dealias=True,
               vext="none", vext_amplitude=0.0, vext_mode=(1, 0, 0), vext_omega=0.0,
               vext_width=1.0, vext_ramp=0.0):
    """model_init (model.f90:56-74) with the species factory (species.f90:49-153);
    transport coefficients are passed in already resolved (eta_bar, xi_bar, tau)."""
    spc = Spec(nx, ny, nz, lx, ly, lz, dealias)
    un = Units(kind=kind, gamma=gamma, kappa=kappa, rs=rs, mass=mass)
    if terms is None:
        terms = "tf,gradient,hartree,ry" if un.quantum else "ideal,hartree,ry"
    fun = build_functional(un, terms, k_tab, c_tab, tfk_gamma)
    fun.setup(spc)
    stress = Stress(closure, eta, xi, tau, un.n0)
    ext = Vext(spc, vext, vext_amplitude, vext_mode, vext_omega, vext_width, vext_ramp)
    if eta_l is None:
        eta_l = 4.0 * eta / 3.0 + xi
    m = Model(spc=spc, fun=fun, stress=stress, ext=ext, n0=un.n0, mass=mass, cfl=cfl, eta_l=eta_l)
    m.dt_fixed = dt
    m.dt = dt
    m.units = un
    return m


def ic_mode(spc, n0, amp, mode):
    """set_initial_condition 'mode' (initial.f90:40-47, 74)."""
    g = spc.g
    kv = [TWOPI * mode[0] / g.lx, TWOPI * mode[1] / g.ly, TWOPI * mode[2] / g.lz]
    if g.ny == 1:
        kv[1] = 0.0
    if g.nz == 1:
        kv[2] = 0.0
    X, Y, Z = g.coords()
    d = np.cos(kv[0] * X + kv[1] * Y + kv[2] * Z)
    n = n0 * (1.0 + amp * d)
    if n.min() <= 0.0:
        raise OracleError("initial density is not positive; reduce the amplitude")
    return n


def ic_random_field(spc, n0, amp, d_raw):
    """set_initial_condition 'random' (initial.f90:48-58, 74) given the raw
    uniform deviates d_raw in [0, 1) (the reference draws them from the
    compiler's intrinsic generator, which is not reproducible across compilers)."""
    d = d_raw - 0.5
    d = spc.dealias(d)
    d = d - spc.mean(d)
    rms = math.sqrt(spc.mean(d * d))
    if rms > 0.0:
        d = d / rms
    n = n0 * (1.0 + amp * d)
    if n.min() <= 0.0:
        raise OracleError("initial density is not positive; reduce the amplitude")
    return n
