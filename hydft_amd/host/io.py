"""The reference's output files: raw field dumps and the diag / probe tables.

Field files (src/model/io.f90:21-36): `<prefix>_fld_NNNNNN.bin`, little-endian
stream with no record markers:
    int32   nx, ny, nz, ndim, npi
    float64 t, lx, ly, lz
    float64 n(nx, ny, nz)            x fastest
    float64 mom(nx, ny, nz, 3)       component-major
    float64 pi(nx, ny, nz, npi)      only when npi > 0 (Maxwell closure)
Arrays here are numpy C-order (nz, ny, nx) and (c, nz, ny, nx), the same
memory.

Tables (src/model/diagnostics.f90:35-77), Fortran edit descriptors:
    <prefix>_diag.dat   '(i9,11es18.9e3)': step t dt mass momx momy momz
                        E_kin F E_tot n_min n_max
    <prefix>_probe.dat  '(6es18.9e3)': t Re(dn_k/n0) Im(dn_k/n0) |dn_k/n0|
                        Re(mom_k) Im(mom_k)
each after a one-line '#' header.
"""
from __future__ import annotations

import math
import struct

import numpy as np

DIAG_HEADER = "# step   t   dt   mass   momx   momy   momz   E_kin   F   E_tot   n_min   n_max"


def field_name(prefix: str, step: int) -> str:
    """'(a,a,i6.6,a)': prefix, '_fld_', step with 6 digits, '.bin' (io.f90:24)."""
    return f"{prefix}_fld_{step:06d}.bin"


def write_fields(path: str, t: float, lx: float, ly: float, lz: float, ndim: int, n, mom, pi=None) -> None:
    n = np.ascontiguousarray(n, dtype="<f8")
    nz, ny, nx = n.shape
    mom = np.ascontiguousarray(mom, dtype="<f8").reshape((3, nz, ny, nx))
    npi = 0 if pi is None else int(np.asarray(pi).shape[0])
    with open(path, "wb") as f:
        f.write(struct.pack("<5i", nx, ny, nz, ndim, npi))
        f.write(struct.pack("<4d", t, lx, ly, lz))
        f.write(n.tobytes())
        f.write(mom.tobytes())
        if npi:
            f.write(np.ascontiguousarray(pi, dtype="<f8").reshape((npi, nz, ny, nx)).tobytes())


def read_fields(path: str) -> dict:
    with open(path, "rb") as f:
        nx, ny, nz, ndim, npi = struct.unpack("<5i", f.read(20))
        t, lx, ly, lz = struct.unpack("<4d", f.read(32))
        npts = nx * ny * nz
        n = np.frombuffer(f.read(8 * npts), dtype="<f8").reshape((nz, ny, nx)).copy()
        mom = np.frombuffer(f.read(24 * npts), dtype="<f8").reshape((3, nz, ny, nx)).copy()
        pi = None
        if npi > 0:
            pi = np.frombuffer(f.read(8 * npi * npts), dtype="<f8").reshape((npi, nz, ny, nx)).copy()
    return {"nx": nx, "ny": ny, "nz": nz, "ndim": ndim, "t": t, "lx": lx, "ly": ly, "lz": lz,
            "n": n, "mom": mom, "pi": pi}


def fmt_es(v: float, w: int = 18, d: int = 9, e: int = 3) -> str:
    """Fortran ESw.dEe: one nonzero digit before the point, d after, an
    exponent of e digits with its sign, right-justified in w columns."""
    v = float(v)
    if math.isnan(v):
        return "NaN".rjust(w)
    if math.isinf(v):
        return ("-Infinity" if v < 0 else "Infinity").rjust(w)
    mant, ex = f"{v:.{d}E}".split("E")
    ev = int(ex)
    if v == 0.0:
        ev = 0
    s = f"{mant}E{'-' if ev < 0 else '+'}{abs(ev):0{e}d}"
    if len(s) > w:
        return "*" * w
    return s.rjust(w)


def diag_line(step: int, t: float, dt: float, mass: float, mom, ekin: float, fe: float, nmin: float,
              nmax: float) -> str:
    """'(i9,11es18.9e3)' (diagnostics.f90:66)."""
    vals = [t, dt, mass, mom[0], mom[1], mom[2], ekin, fe, ekin + fe, nmin, nmax]
    return f"{step:9d}" + "".join(fmt_es(v) for v in vals)


def probe_header(mode) -> str:
    """'(a,3i4,a)' (diagnostics.f90:38)."""
    return ("# probe mode" + "".join(f"{int(k):4d}" for k in mode) +
            ":   t   Re(dn_k/n0)   Im(dn_k/n0)   |dn_k/n0|   Re(mom_k)   Im(mom_k)")


def probe_line(t: float, a: complex, b: complex) -> str:
    """'(6es18.9e3)' (diagnostics.f90:75)."""
    return "".join(fmt_es(v) for v in (t, a.real, a.imag, abs(a), b.real, b.imag))


def read_table(path: str) -> np.ndarray:
    """Rows of a diag or probe table (header lines skipped)."""
    return np.atleast_2d(np.loadtxt(path, comments="#"))


class DiagFiles:
    """diag_t%open/write/probe/close (diagnostics.f90:29-84) on files."""

    def __init__(self, prefix: str, probe_mode):
        self.probe_mode = tuple(int(k) for k in probe_mode)
        self.fd = open(f"{prefix}_diag.dat", "w")
        self.fd.write(DIAG_HEADER + "\n")
        self.fp = open(f"{prefix}_probe.dat", "w")
        self.fp.write(probe_header(self.probe_mode) + "\n")

    def write(self, step, t, dt, mass, mom, ekin, fe, nmin, nmax):
        self.fd.write(diag_line(step, t, dt, mass, mom, ekin, fe, nmin, nmax) + "\n")

    def probe(self, t, a, b):
        self.fp.write(probe_line(t, a, b) + "\n")

    def close(self):
        for f in (self.fd, self.fp):
            if not f.closed:
                f.close()
