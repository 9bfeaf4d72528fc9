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


EXE_MODEL = os.path.join(os.path.dirname(EXE), "f_model")


def test_fortran_model_drop_in(gpu):
    """The model entry points through hydft_hip_model: a single-mode ion run
    against linear theory (1 %), mass conservation (1e-12), the kept 2/3-rule
    box, batched and captured steps equal to hy_model_advance bitwise, and a
    group launch of a batch whose runs stop at 20 and 40 steps."""
    if not os.path.exists(EXE_MODEL):
        pytest.fail("Fortran program not built (run __graft_entry__.build())")
    r = subprocess.run([EXE_MODEL], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "drv_model: all passed" in r.stdout
    assert r.stdout.count("  pass  ") == 9
