"""The Fortran binding (hydft_amd/fortran/hydft_hip.f90): the reference's
test_fft.f90 and test_ops.f90 checks, written in Fortran against hip_spec_t
(the drop-in for spec_t) and run on the GPU as a separate process."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hydft_amd", "f_spec")


def test_fortran_spec_drop_in(gpu):
    if not os.path.exists(EXE):
        pytest.fail("Fortran program not built (run __graft_entry__.build())")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "drv_spec: all passed" in r.stdout
    assert r.stdout.count("  pass  ") == 17
