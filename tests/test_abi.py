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
