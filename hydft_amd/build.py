"""Build libhydft.so (HIP, gfx950) in-tree, and the Fortran binding objects.

    python -m hydft_amd.build            # compile everything
Each HIP translation unit is compiled in parallel to an object, then linked.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_obj"
LIB = PKG / "libhydft.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HYDFT_ARCH", "gfx950")
HIPFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _compile(src: Path) -> Path:
    obj = OBJ / (src.stem + ".o")
    _run([HIPCC, *HIPFLAGS, "-c", str(src), "-o", str(obj)])
    return obj


def build_lib(verbose=True) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp),
          "-L/opt/rocm/lib", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_fortran(verbose=True):
    """Compile the iso_c_binding module and its drop-in test programs (amdflang)."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    fdir = PKG / "fortran"
    if not Path(fc).exists() or not fdir.exists():
        return []
    out = []
    mod_dir = OBJ / "fmod"
    mod_dir.mkdir(parents=True, exist_ok=True)
    srcs = sorted(fdir.glob("hydft_hip*.f90"))
    objs = []
    for s in srcs:
        o = mod_dir / (s.stem + ".o")
        _run([fc, "-O2", "-c", str(s), "-J", str(mod_dir), "-o", str(o)])
        objs.append(str(o))
    for prog in sorted(fdir.glob("drv_*.f90")):
        exe = PKG / ("f_" + prog.stem[4:])
        _run([fc, "-O2", str(prog), *objs, "-I", str(mod_dir), "-J", str(mod_dir),
              f"-L{PKG}", "-lhydft", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN", "-o", str(exe)])
        out.append(exe)
    if verbose and out:
        print("built", " ".join(str(p.name) for p in out))
    return out


def main():
    build_lib()
    build_fortran()


if __name__ == "__main__":
    sys.exit(main())
