"""bench.py's measurement contract on the CPU: one GPU only, and the traffic
bookkeeping that feeds roofline.traffic (no device is used)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr and r.stdout.strip() == ""


def test_kernel_traffic_picks_the_class_kernel_of_the_run_type():
    """kernel_traffic sums FETCH_SIZE and WRITE_SIZE of the class's kernel in
    the run's transform type, and is None when either counter is missing."""
    tr = {"per_kernel": {
        "void hy::k_phaseA<double, 256, hy::XWave<double, 256>, hy::XCfg<0> >(...)":
            {"FETCH_SIZE": 5.0e8, "WRITE_SIZE": 3.0e8},
        "void hy::k_phaseA<float, 256, hy::XWave<float, 256>, hy::XCfg<2> >(...)":
            {"FETCH_SIZE": 4.0e8, "WRITE_SIZE": 1.0e8},
        "void hy::k_pencil_inv<float, 256, 32, hy::AxisZ>(hy::PencilK<float>)":
            {"FETCH_SIZE": 2.0e8},
    }, "factors": {"FETCH_SIZE": 2.0, "WRITE_SIZE": 1.0}}
    assert bench.kernel_traffic(tr, "x_phaseA", "mixed") == 5.0e8
    assert bench.kernel_traffic(tr, "x_phaseA", "fp64") == 8.0e8
    assert bench.kernel_traffic(tr, "pencil_z_inverse", "mixed") is None
    assert bench.kernel_traffic(tr, "not_a_class", "mixed") is None
    assert bench.kernel_traffic(None, "x_phaseA", "mixed") is None
