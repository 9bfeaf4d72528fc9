"""Output formats of the reference (src/model/io.f90, diagnostics.f90)."""
import struct

import numpy as np

from hydft_amd.host import io


def test_field_file_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    nx, ny, nz = 6, 4, 3
    n = rng.random((nz, ny, nx))
    mom = rng.random((3, nz, ny, nx))
    pi = rng.random((6, nz, ny, nx))
    path = tmp_path / io.field_name("run", 1200)
    assert path.name == "run_fld_001200.bin"  # '(a,a,i6.6,a)'
    io.write_fields(str(path), 1.25, 10.0, 5.0, 2.0, 3, n, mom, pi)
    raw = path.read_bytes()
    # header: 5 int32 + 4 float64, then n, mom(3), pi(6), all little-endian, x fastest
    assert struct.unpack("<5i", raw[:20]) == (nx, ny, nz, 3, 6)
    assert struct.unpack("<4d", raw[20:52]) == (1.25, 10.0, 5.0, 2.0)
    assert len(raw) == 52 + 8 * nx * ny * nz * (1 + 3 + 6)
    # element (i, j, k) of the Fortran array n(nx, ny, nz) sits at i + nx (j + ny k)
    i, j, k = 4, 2, 1
    off = 52 + 8 * (i + nx * (j + ny * k))
    assert struct.unpack("<d", raw[off:off + 8])[0] == n[k, j, i]
    c = 2  # mom(i, j, k, c+1)
    off = 52 + 8 * nx * ny * nz * (1 + c) + 8 * (i + nx * (j + ny * k))
    assert struct.unpack("<d", raw[off:off + 8])[0] == mom[c, k, j, i]
    d = io.read_fields(str(path))
    assert (d["nx"], d["ny"], d["nz"], d["ndim"]) == (nx, ny, nz, 3)
    assert d["t"] == 1.25 and d["lx"] == 10.0
    np.testing.assert_array_equal(d["n"], n)
    np.testing.assert_array_equal(d["mom"], mom)
    np.testing.assert_array_equal(d["pi"], pi)
    io.write_fields(str(path), 0.0, 1.0, 1.0, 1.0, 1, n, mom)
    assert io.read_fields(str(path))["pi"] is None


def test_fortran_es_edit_descriptor():
    # ES18.9E3 as the reference writes its tables
    assert io.fmt_es(1.0) == "  1.000000000E+000"
    assert io.fmt_es(-2.5e-7) == " -2.500000000E-007"
    assert io.fmt_es(0.0) == "  0.000000000E+000"
    assert io.fmt_es(123456.789) == "  1.234567890E+005"
    assert io.fmt_es(9.9999999999) == "  1.000000000E+001"
    assert io.fmt_es(1e-300) == "  1.000000000E-300"
    assert len(io.fmt_es(-1.0e123)) == 18


def test_diag_and_probe_tables(tmp_path):
    prefix = str(tmp_path / "r")
    d = io.DiagFiles(prefix, (1, 0, 2))
    d.write(20, 0.2, 0.01, 3.0, (1e-3, 0.0, -2e-4), 0.5, -1.5, 0.9, 1.1)
    d.probe(0.2, complex(1e-4, -2e-4), complex(3e-5, 0.0))
    d.close()
    lines = open(prefix + "_diag.dat").read().splitlines()
    assert lines[0] == io.DIAG_HEADER
    assert len(lines[1]) == 9 + 11 * 18  # '(i9,11es18.9e3)'
    assert lines[1][:9] == "       20"
    row = io.read_table(prefix + "_diag.dat")[0]
    assert row[0] == 20 and row[9] == 0.5 - 1.5  # E_tot = E_kin + F
    plines = open(prefix + "_probe.dat").read().splitlines()
    assert plines[0].startswith("# probe mode   1   0   2:")  # '(a,3i4,a)'
    pr = io.read_table(prefix + "_probe.dat")[0]
    assert len(plines[1]) == 6 * 18
    np.testing.assert_allclose(pr[3], abs(complex(1e-4, -2e-4)), rtol=1e-9)
