"""hydft_run (src/hydft.f90:21-55) on the GPU path.

Reads the reference's namelist input, builds the species (units, HNC c(k),
transport) and the device model, sets the initial condition, and runs the
time loop with the reference's cadence:
- compute_dt every 20 steps;
- the last step clamped to tend;
- a probe row every iprobe steps;
- a diag row every idiag steps;
- fields every iout steps;
- a final diag row.

The RK3 steps between output events run on the device in one
hy_model_advance call (the dt cadence and the tend clamp are applied there).
The probe values of every step are recorded on the device and written here.
The files are the reference's (`hydft_amd.host.io`).

    python -m hydft_amd.run input.nml
"""
from __future__ import annotations

import sys
import time

from .host import io
from .host.initial import set_initial_condition
from .host.params import Params, read_params

DT_EVERY = 20  # src/hydft.f90:42-43


def _write_diag(d: io.DiagFiles, m) -> None:
    t, dt, step = m.time
    mass, mom, ekin, fe = m.budgets()
    n, _, _ = m.get_state()
    d.write(step, t, dt, mass, mom, ekin, fe, float(n.min()), float(n.max()))


def _write_fields(p: Params, m) -> str:
    t, _, step = m.time
    n, mom, pi = m.get_state()
    path = io.field_name(p.prefix, step)
    io.write_fields(path, t, p.lx, p.ly, p.lz, _ndim(p), n, mom, pi)
    return path


def _ndim(p: Params) -> int:
    return 3 if p.nz > 1 else (2 if p.ny > 1 else 1)


def run(p: Params, verbose: bool = True, model=None) -> dict:
    """Run p to tend / nsteps_max.  Returns a summary (steps, t, wall seconds,
    file names)."""
    from .model import Model, build_species

    m = model
    if m is None:
        sp = build_species(p)
        m = Model(p, species=sp)
    n = set_initial_condition(p, m.n0)
    m.set_state(n)
    d = io.DiagFiles(p.prefix, p.probe_mode)
    fields = []
    try:
        dt0 = m.compute_dt()
        if verbose:
            print(f"  initial dt = {dt0:12.4E}")
        _write_diag(d, m)
        a, b = m.probe(p.probe_mode)
        d.probe(m.time[0], a, b)
        if p.iout > 0:
            fields.append(_write_fields(p, m))
        w0 = time.perf_counter()
        while True:
            t, _, step = m.time
            if not (t < p.tend and step < p.nsteps_max):
                break
            events = [p.iprobe, p.idiag] + ([p.iout] if p.iout > 0 else [])
            nxt = min(e - step % e for e in events)
            nxt = max(1, min(nxt, p.nsteps_max - step))
            taken, rows = m.advance(nxt, dt_every=DT_EVERY, tend=p.tend, probe=p.probe_mode)
            for i in range(taken):
                s = step + i + 1
                if s % p.iprobe == 0:
                    r = rows[i]
                    d.probe(r[0], complex(r[1], r[2]), complex(r[3], r[4]))
            step += taken
            if taken and step % p.idiag == 0:
                _write_diag(d, m)
            if taken and p.iout > 0 and step % p.iout == 0:
                fields.append(_write_fields(p, m))
            if taken < nxt:
                break
        wall = time.perf_counter() - w0
        t, _, step = m.time
        if step % p.idiag != 0:
            _write_diag(d, m)
    finally:
        d.close()
    if verbose:
        print(f"hydft: finished {step} steps, t = {t:12.4f}  ({wall:10.2f} s)")
    return {"steps": step, "t": t, "wall_s": wall, "diag": f"{p.prefix}_diag.dat",
            "probe": f"{p.prefix}_probe.dat", "fields": fields}


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 1:
        print("usage: python -m hydft_amd.run input.nml", file=sys.stderr)
        return 2
    from . import abi
    abi.call("hy_init", 0)
    p = read_params(argv[0])
    run(p)
    return 0


if __name__ == "__main__":
    sys.exit(main())
