"""The C-ABI library loads and exports every symbol include/hydft.h declares
(CPU: no device call is made), and its host-only entry points work."""
import ctypes as C

import numpy as np

from hydft_amd import abi


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(abi.LIBPATH))
    declared = abi.declared_symbols()
    assert len(declared) >= 35
    missing = [s for s in declared if not hasattr(lib, s)]
    assert missing == []
    assert abi.lib().missing == []


def test_host_entry_points_without_gpu():
    assert abi.lib().lib.hy_version() >= 100
    x = np.array([0.0, 2.0])
    y = np.zeros(2)
    abi.call("hy_fd_eval", 1, 2, x.ctypes.data, y.ctypes.data)
    assert abs(y[0] - 0.678093895) < 1e-8
    try:
        abi.call("hy_fd_eval", 7, 1, np.array([-1.0]).ctypes.data, y.ctypes.data)
        raise AssertionError("expected an error")
    except abi.HydftError as e:
        assert e.code == abi.HY_ERR_INVALID and "y > 0" in str(e)
    assert abi.lib().lib.hy_profile_class_name(0) == b"pencil_z_inverse"


def test_fortran_descriptors_match_c_layout(tmp_path):
    """The bind(C) mirrors in hydft_hip_model.f90 have the size and field
    offsets of the C structs (checked against the ctypes mirror)."""
    import os
    import shutil
    import subprocess
    import pytest
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    mod = os.path.join(os.path.dirname(abi.__file__), "_obj", "fmod")
    if not os.path.exists(fc) or not os.path.exists(os.path.join(mod, "hydft_hip_model.mod")):
        pytest.skip("amdflang or the built Fortran module is missing")
    src = tmp_path / "lay.f90"
    src.write_text("""program lay
  use, intrinsic :: iso_c_binding
  use hydft_hip_model
  type(hy_term_desc), target :: t
  type(hy_model_desc), target :: d
  integer(c_intptr_t) :: b
  b = transfer(c_loc(d), b)
  print '(6(i0,1x))', c_sizeof(t), c_sizeof(d), transfer(c_loc(d%terms), b) - b, &
    transfer(c_loc(d%vext_mode), b) - b, transfer(c_loc(d%cfl), b) - b, &
    transfer(c_loc(t%c_tab), b) - transfer(c_loc(t), b)
end program
""")
    exe = tmp_path / "lay"
    subprocess.run([fc, "-I", mod, str(src), "-o", str(exe)], check=True, capture_output=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = [C.sizeof(abi.TermDesc), C.sizeof(abi.ModelDesc), abi.ModelDesc.terms.offset,
            abi.ModelDesc.vext_mode.offset, abi.ModelDesc.cfl.offset, abi.TermDesc.c_tab.offset]
    assert got == want


def test_grid_validation_matches_reference():
    """hy_grid_create rejects the grids grid_init stops on (grid.f90:45-48), with
    its messages, before any device call; lengths outside the supported radix
    set are HY_ERR_UNSUPPORTED."""
    h = C.c_void_p()

    def err(nx, ny, nz):
        try:
            abi.call("hy_grid_create", nx, ny, nz, 1.0, 1.0, 1.0, 1, abi.HY_PREC_FP64, C.byref(h))
        except abi.HydftError as e:
            return e.code, str(e)
        raise AssertionError(f"grid {nx}x{ny}x{nz} accepted")

    code, msg = err(1, 1, 1)
    assert code == abi.HY_ERR_INVALID and "nx must be >= 2" in msg
    assert err(0, 8, 8)[0] == abi.HY_ERR_INVALID
    code, msg = err(7, 1, 1)
    assert code == abi.HY_ERR_INVALID and "nx must be even" in msg
    code, msg = err(8, 3, 1)
    assert code == abi.HY_ERR_INVALID and "ny must be even (or 1)" in msg
    code, msg = err(8, 8, 5)
    assert code == abi.HY_ERR_INVALID and "nz must be even (or 1)" in msg
    for shape in [(6, 1, 1), (10, 10, 10), (512, 1, 1), (16, 20, 1), (16, 16, 40)]:
        assert err(*shape)[0] == abi.HY_ERR_UNSUPPORTED, shape
    try:
        abi.call("hy_grid_create", 16, 1, 1, 1.0, 1.0, 1.0, 1, 7, C.byref(h))
        raise AssertionError("unknown precision accepted")
    except abi.HydftError as e:
        assert e.code == abi.HY_ERR_INVALID
