#!/usr/bin/env python3
"""RK3 steps/s of the 3-D pseudo-spectral DDFT-QHD model on one MI355X.

Workload (default): quantum electrons, Thomas-Fermi + Kirzhnits gradient +
Coulomb Hartree + RY correlation from the Hansen-McDonald HNC c_ee, Newtonian
electron_fit viscosity, r_s = 1, theta = 1, 256^3 periodic grid in a box
L = 2 pi/0.05; FP32 spectral kernels with an FP64 RK3 state.  Initial dn:
seeded random phases on the shell 0.1 < q < 1.5 with rms 5e-3 n0; u: three
independent seeded fields on the same shell with 1e-3 rms each, mom = n u.
dt is the CFL step (cfl = 0.4) at t = 0, frozen.  Synthetic, seeded inputs.

One JSON line on stdout.  `roofline` is for the kernel class with the largest
summed time in the timed region (HIP events on the model's stream; algorithmic
bytes = compulsory reads + writes of each launch).  Its `traffic` is the
measured HBM bytes per launch of that kernel: two rocprofv3 counter runs
(FETCH_SIZE, WRITE_SIZE) of a one-step child process, started before this
process uses the GPU, each scaled by what the counter reads for a plain 1 GiB
copy in the same child.  `cpu_baseline` times one RK3 stage of the CPU FP64
oracle (the reference dataflow, single core) on the same workload and scales
it to steps/s.
"""
from __future__ import annotations

import argparse
import csv
import ctypes as C
import glob
import json
import math
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def config(n=256, precision="mixed"):
    from hydft_amd.host.params import Params
    p = Params()
    p.kind, p.rs, p.theta = "electron", 1.0, 1.0
    p.nx = p.ny = p.nz = n
    p.lx = p.ly = p.lz = 2.0 * math.pi / 0.05
    p.terms = "tf,gradient,hartree,ry"
    p.transport, p.closure = "electron_fit", "newtonian"
    p.cfl = 0.4
    p.precision = precision
    return p.finalize()


def initial_state(p, n0, seed=1234):
    from hydft_amd.host.initial import shell_rms_field
    dn = shell_rms_field(p, 5e-3 * n0, 0.1, 1.5, seed)
    n = n0 + dn
    u = np.stack([shell_rms_field(p, 1e-3, 0.1, 1.5, seed + 1 + d) for d in range(3)])
    return n, n[None] * u


def profile_table(m, ncls=10):
    from hydft_amd import abi
    launches = (C.c_int * ncls)()
    ms = (C.c_double * ncls)()
    by = (C.c_double * ncls)()
    n = abi.lib().lib.hy_model_profile_read(m.h, ncls, launches, ms, by)
    out = []
    for c in range(min(n, ncls)):
        if launches[c]:
            out.append({"kernel": abi.lib().lib.hy_profile_class_name(c).decode(), "cls": c,
                        "launches": launches[c], "total_ms": ms[c], "bytes": by[c]})
    return out


def cpu_baseline(p, sp, n, mom):
    """One RK3 stage of the oracle (reference dataflow, FP64) on the same model
    as the GPU run (species, c(k), electron-fit viscosity), its transforms
    threaded over the host cores this process may use."""
    from oracle import hydft_oracle as O
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cores = max(1, min(cores, 16))  # the GPU machine's host share per GPU
    st = sp.structure
    om = O.make_model(p.nx, p.ny, p.nz, p.lx, p.ly, p.lz, kind=p.kind, gamma=p.gamma, rs=p.rs,
                      terms=p.terms, k_tab=st.k, c_tab=st.c, tfk_gamma=p.tfk_gamma, closure=sp.closure,
                      eta=sp.transport.eta, xi=sp.transport.xi, tau=sp.tau, eta_l=sp.transport.eta_l, dt=1e-3)
    om.n = n.copy()
    om.mom = [mom[d].copy() for d in range(3)]
    O.FFT_WORKERS = cores
    try:
        t0 = time.perf_counter()
        om.advance_stage(0)
        el = time.perf_counter() - t0
    finally:
        O.FFT_WORKERS = 1
    return {"value": 1.0 / (3.0 * el), "unit": "steps/s", "cores": cores, "kind": "port",
            "sample": f"one RK3 stage (of 3) of the FP64 oracle at {p.nx}^3 (transforms on {cores} threads, "
                      f"the rest on one), {el:.1f} s, scaled to steps/s"}


def pencil_plans():
    """The autotuner's choices so far: {fp64, L, inverse, A, B, nfields, batch,
    columns per workgroup} per pass shape (hy_plan_list)."""
    from hydft_amd import abi
    lib = abi.lib().lib
    n = lib.hy_plan_list(0, None)
    buf = (C.c_int * (8 * max(n, 1)))()
    lib.hy_plan_list(n, buf)
    return [list(buf[8 * i:8 * i + 8]) for i in range(n)]


def copy_probe(mib, timed=False):
    """A plain copy kernel of known size: reads and writes mib MiB once per launch, 3 launches.
    timed: also time 10 more launches with events and report the copy rate, the
    practical HBM ceiling of a kernel that reads and writes once (read + write
    bytes over time)."""
    import torch
    x = torch.ones(mib * 262144, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    out = {"copy_probe_bytes_read": mib * 1048576, "copy_probe_bytes_written": mib * 1048576, "launches": 3}
    if timed:
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            y.copy_(x)
        e1.record()
        torch.cuda.synchronize()
        sec = e0.elapsed_time(e1) * 1e-3 / 10
        out["copy_GBps"] = 2.0 * mib * 1048576 / sec / 1e9
        out["launches"] += 10
        e0.record()
        for _ in range(10):
            y.fill_(1.0)
        e1.record()
        torch.cuda.synchronize()
        out["fill_GBps"] = mib * 1048576 / (e0.elapsed_time(e1) * 1e-3 / 10) / 1e9
    print(json.dumps(out))
    return 0


# kernel-class name -> device kernel of that class, and the tag in its name
# (the single-run pencil kernels carry their axis: AxisY or AxisZ)
CLASS_KERNEL = {"x_phaseA": ("k_phaseA", ""), "x_phaseB": ("k_phaseB", ""), "spec_update": ("k_spec2", ""),
                "spec_mu": ("k_spec1", ""), "pencil_z_inverse": ("k_pencil_inv", "AxisZ"),
                "pencil_y_inverse": ("k_pencil_inv", "AxisY"), "pencil_y_forward": ("k_pencil_fwd", "AxisY"),
                "pencil_z_forward": ("k_pencil_fwd", "AxisZ"), "x_umax": ("k_umax", "")}
COPY_MIB = 1024


def pmc_child(args):
    """Body of the counter runs: a plain copy of known size, then one RK3 step."""
    import torch
    from hydft_amd import abi
    from hydft_amd.model import Model, build_species
    copy_probe(COPY_MIB)
    abi.call("hy_init", 0)
    p = config(args.n, args.precision)
    sp = build_species(p)
    m = Model(p, species=sp, precision=args.precision)
    n, mom = initial_state(p, m.n0)
    m.set_state(n, mom)
    m.set_dt(m.compute_dt())
    m.advance(1)
    torch.cuda.synchronize()
    return 0


def pmc_traffic(args):
    """HBM bytes per launch of every kernel, from two rocprofv3 runs of the
    one-step child (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch).  Each
    counter is scaled by the ratio of known bytes to what it reads for the
    child's plain 1 GiB copy.  Runs before this process touches the GPU."""
    rocprof = shutil.which("rocprofv3")
    if not rocprof:
        return None, "rocprofv3 not found"
    per = {}
    factors = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="hydft_pmc_")
        cmd = [rocprof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "pmc", "--",
               sys.executable, os.path.abspath(__file__), "--pmc-child", "--n", str(args.n),
               "--precision", args.precision]
        try:
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
        except subprocess.TimeoutExpired:
            shutil.rmtree(d, ignore_errors=True)
            return None, f"{counter} run timed out"
        files = glob.glob(os.path.join(d, "**", "pmc_counter_collection.csv"), recursive=True)
        if r.returncode != 0 or not files:
            shutil.rmtree(d, ignore_errors=True)
            return None, f"{counter} run failed (rc {r.returncode})"
        vals = {}
        for row in csv.DictReader(open(files[0])):
            vals.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]) * 1024.0)
        shutil.rmtree(d, ignore_errors=True)
        copies = [v for k, vs in vals.items() if "copyBuffer" in k for v in vs]
        if not copies:
            return None, "no copy kernel in the counter run"
        known = COPY_MIB * 1048576.0
        factors[counter] = known / max(copies)
        for k, vs in vals.items():
            per.setdefault(k, {})[counter] = sum(vs) / len(vs) * factors[counter]
    return {"per_kernel": per, "factors": factors}, None


def kernel_traffic(tr, cls, precision):
    """(fetch + write) bytes per launch of the device kernel of class `cls` in
    the run's transform type (mean over the dispatches of that kernel)."""
    if not tr or cls not in CLASS_KERNEL:
        return None
    want = "<float" if precision == "mixed" else "<double"
    kern, tag = CLASS_KERNEL[cls]
    names = [k for k in tr["per_kernel"] if kern + "<" in k and tag in k]
    names = [k for k in names if want in k] or names
    for name in names:
        c = tr["per_kernel"][name]
        if "FETCH_SIZE" in c and "WRITE_SIZE" in c:
            return c["FETCH_SIZE"] + c["WRITE_SIZE"]
    return None


def _one_at_a_time_estimate(reqs, cache, per_class=2):
    """The reference's way (one model at a time, hy_model_advance) timed on a
    stratified sample: `per_class` requests of each (grid, species, closure)
    class, scaled by the class counts.  Returns (estimated seconds for all
    requests, sample size, the sample's results)."""
    from hydft_amd import serve
    classes = {}
    for r in reqs:
        classes.setdefault((r.p.nx, r.p.kind, r.p.closure), []).append(r)
    total, n, results = 0.0, 0, []
    for key, rs in classes.items():
        pick = rs[:per_class]
        t0 = time.perf_counter()
        res = serve.one_at_a_time(pick, precision="mixed", cache=cache)
        el = time.perf_counter() - t0
        steps_pick = sum(r.nsteps for r in pick)
        steps_all = sum(r.nsteps for r in rs)
        total += el * steps_all / steps_pick
        n += len(pick)
        results += res
    return total, n, results


def scan_sample(nruns, steps, grids=(32, 48, 64)):
    """The batched driven-mode scan: `nruns` Yukawa-ion chi(k, w) runs, grids
    cycling through 32^3, 48^3, 64^3, Gamma log-uniform in [1, 100], kappa in
    [0.5, 3], Newtonian / Maxwell alternating (tau = 1), drive 1e-4 cos(q x - w t)
    with q in [0.2, 2] and w in [0.1, 3] w_p, CFL dt every 20 steps, `steps`
    steps each, FP32-mixed, all submitted at once.  The c(k) of all state points
    are solved together on the GPU first (timed apart: `hnc_device_s`).  Warm
    cache: batched (one launch per kernel for each batch of same-shape runs)
    against one at a time, the latter timed on a stratified sample.  Cold cache,
    the reference's way on a small sample: host HNC then the run, one at a
    time."""
    from hydft_amd import serve
    from hydft_amd.host import hnc
    from hydft_amd.model import hnc_request
    reqs = serve.chi_requests(nruns, seed=11, grid=grids, steps=steps)
    cache = hnc.StructureCache()
    t0 = time.perf_counter()
    cache.solve_many([hnc_request(r.p) for r in reqs], device=True)
    t_hnc = time.perf_counter() - t0
    print(f"# scan: {nruns} c(k) on the GPU in {t_hnc:.1f} s", file=sys.stderr, flush=True)
    srv = serve.Server(precision="mixed", cache=cache, max_slots=64, batch=True, max_batch=512)
    t0 = time.perf_counter()
    batch = srv.serve(reqs, realtime=False)
    w_batch = time.perf_counter() - t0
    print(f"# scan: batched {nruns} runs in {w_batch:.1f} s", file=sys.stderr, flush=True)
    w_one, n_one, one = _one_at_a_time_estimate(reqs, cache)
    print(f"# scan: one at a time sample of {n_one} done", file=sys.stderr, flush=True)
    by = {r.rid: r for r in batch}
    same = all(np.array_equal(r.probe, by[r.rid].probe) for r in one)
    # cold cache, the reference's way: host HNC solve, then the run
    ncold = 4
    cache_h = hnc.StructureCache()
    t0 = time.perf_counter()
    serve.one_at_a_time(reqs[:ncold], precision="mixed", cache=cache_h)
    w_cold = (time.perf_counter() - t0) / ncold
    s_b = serve.summary(batch, w_batch)
    return {"runs": nruns, "grids": list(grids), "steps_per_run": steps,
            "hnc_device_s": t_hnc,
            "batched_runs_per_s": s_b["runs_per_s"],
            "batched_launch_groups": srv.batches_launched,
            "one_at_a_time_runs_per_s": nruns / w_one,
            "one_at_a_time_sample": n_one,
            "batched_over_one_at_a_time": s_b["runs_per_s"] * w_one / nruns,
            "cold_ck_one_at_a_time_runs_per_s": 1.0 / w_cold, "cold_ck_sample": ncold,
            "cold_ck_batched_runs_per_s": nruns / (w_batch + t_hnc),
            "cold_ck_batched_over_one_at_a_time": (nruns / (w_batch + t_hnc)) * w_cold,
            "sample_results_equal_batched": bool(same),
            "chi_vs_linear_median": s_b.get("chi_vs_linear_median")}


def stream_sample(nreq, rate=50.0):
    """The serving stream: `nreq` chi(k, w) / dispersion requests with Poisson
    arrivals at `rate`/s (hydft_amd.serve.stream_requests: grids 32/64/96/128^3
    at weights .4/.3/.2/.1, ions or electrons, Newtonian or Maxwell, 500-5000
    steps, FP32-mixed), served with continuous batching: requests that have
    arrived are admitted in batches as slots and HBM free up while earlier ones
    run.  The c(k) are solved on the GPU before the stream starts (timed
    apart).  Reports runs/s and p50/p99 latency (arrival to result); one at a
    time is timed on a stratified sample."""
    from hydft_amd import serve
    from hydft_amd.host import hnc
    from hydft_amd.model import hnc_request
    reqs = serve.stream_requests(nreq, seed=5, arrival_rate=rate)
    cache = hnc.StructureCache()
    t0 = time.perf_counter()
    cache.solve_many([hnc_request(r.p) for r in reqs], device=True)
    t_hnc = time.perf_counter() - t0
    print(f"# stream: {nreq} c(k) on the GPU in {t_hnc:.1f} s", file=sys.stderr, flush=True)
    srv = serve.Server(precision="mixed", cache=cache, max_slots=64, batch=True, max_batch=256)
    t0 = time.perf_counter()
    served = srv.serve(reqs, realtime=True)
    w = time.perf_counter() - t0
    print(f"# stream: {nreq} requests served in {w:.1f} s", file=sys.stderr, flush=True)
    w_one, n_one, one = _one_at_a_time_estimate(reqs, cache, per_class=1)
    by = {r.rid: r for r in served}
    same = all(np.array_equal(r.probe, by[r.rid].probe) for r in one)
    s = serve.summary(served, w)
    mix = {}
    for r in reqs:
        k = f"{r.p.nx}^3 {r.p.kind}"
        mix[k] = mix.get(k, 0) + 1
    return {"requests": nreq, "arrival_rate_per_s": rate, "mix": mix,
            "mean_steps": float(np.mean([r.nsteps for r in reqs])),
            "hnc_device_s": t_hnc,
            "served_runs_per_s": s["runs_per_s"], "latency_p50_s": s["latency_p50_s"],
            "latency_p99_s": s["latency_p99_s"], "launch_groups": srv.batches_launched,
            "one_at_a_time_runs_per_s": nreq / w_one, "one_at_a_time_sample": n_one,
            "served_over_one_at_a_time": s["runs_per_s"] * w_one / nreq,
            "sample_results_equal_served": bool(same)}


def maxwell_sample(n=128, steps=1000):
    """The 128^3 Maxwell configuration: Yukawa ions, Gamma = 50, kappa = 2,
    Maxwell closure with Pi as state (Stanton-Murillo eta, Ichimaru tau from
    the HNC excess energy), dt = 0.005, 1000 steps, FP64 and FP32-mixed from
    the same state.  Initial dn: a seeded Gaussian random field with per-mode
    modulus 1e-2 n0 exp(-q^2/2) and uniform random phases, zero mean, in a box
    of 4 pi (a box of 2 pi/0.2 would put an rms of 0.26 n0 on this spectrum
    and the density below zero); u = 0, Pi = 0.  Returns steps/s of both and
    the mixed density error against FP64 after the run."""
    from hydft_amd import abi
    from hydft_amd.host.initial import gauss_spectrum_field
    from hydft_amd.host.params import Params
    from hydft_amd.model import Model, build_species
    p = Params()
    p.kind, p.gamma, p.kappa = "ion", 50.0, 2.0
    p.nx = p.ny = p.nz = n
    p.lx = p.ly = p.lz = 4.0 * math.pi
    p.closure, p.transport, p.tau_model = "maxwell", "stanton_murillo", "ichimaru"
    p.dt = 0.005
    p.finalize()
    sp = build_species(p)
    n0 = sp.units.n0
    dens = gauss_spectrum_field(p, n0, 1e-2, 1234)
    out = {"grid": n, "steps": steps, "dt": p.dt, "box": p.lx,
           "dn_rms_over_n0": float(np.sqrt(np.mean((dens - n0) ** 2)) / n0)}
    fields = {}
    for prec in ("fp64", "mixed"):
        m = Model(p, species=sp, precision=prec)
        m.set_state(dens)
        m.advance(5)
        abi.call("hy_device_sync")
        t0 = time.perf_counter()
        m.advance(steps - 5)
        abi.call("hy_device_sync")
        out[f"{prec}_steps_per_s"] = (steps - 5) / (time.perf_counter() - t0)
        fields[prec] = m.get_state()[0]
        m.close()
    a, b = fields["fp64"], fields["mixed"]
    out["mixed_over_fp64"] = out["mixed_steps_per_s"] / out["fp64_steps_per_s"]
    out["mixed_density_rel_l2_vs_fp64"] = float(np.linalg.norm(b - a) / np.linalg.norm(a))
    out["mixed_dn_rel_l2_vs_fp64"] = float(np.linalg.norm(b - a) / np.linalg.norm(a - n0))
    return out


def mixed_error_run(p, species, n, mom, dt, steps=500):
    """Density error of the FP32-mixed run against the FP64 run after `steps`
    steps of the benchmark configuration (the configuration's 500)."""
    from hydft_amd.model import Model
    out = {}
    for prec in ("fp64", "mixed"):
        m = Model(p, species=species, precision=prec)
        m.set_state(n, mom)
        m.set_dt(dt)
        m.advance(steps)
        out[prec] = m.get_state()[0]
        m.close()
    a, b = out["fp64"], out["mixed"]
    n0 = species.units.n0
    return {"steps": steps, "density_rel_l2": float(np.linalg.norm(b - a) / np.linalg.norm(a)),
            "dn_rel_l2": float(np.linalg.norm(b - a) / np.linalg.norm(a - n0))}


def batch_only(spec, steps):
    """Step one batch of same-shape driven-mode runs (spec "B,N,closure") and
    print model-steps/s: the batched kernels alone, for profiling."""
    from hydft_amd import abi, serve
    from hydft_amd.host import hnc
    from hydft_amd.model import Batch, hnc_request
    b_s, n_s, clos = spec.split(",")
    nb, ng = int(b_s), int(n_s)
    abi.call("hy_init", 0)
    reqs = [r for r in serve.chi_requests(4 * nb, seed=11, grid=(ng,), steps=steps) if r.p.closure == clos][:nb]
    cache = hnc.StructureCache()
    cache.solve_many([hnc_request(r.p) for r in reqs])
    ms = [serve.make_model(r, "mixed", cache) for r in reqs]
    b = Batch(ms, 20, dt_every=20, probe=(1, 0, 0), ring=steps)
    b.launch(1)
    b.sync()
    t0 = time.perf_counter()
    b.launch(steps // 20 - 1)
    b.sync()
    w = time.perf_counter() - t0
    print(json.dumps({"batch": nb, "grid": ng, "closure": clos, "steps": steps - 20,
                      "batch_steps_per_s": (steps - 20) / w, "model_steps_per_s": nb * (steps - 20) / w}))
    b.close()
    for m in ms:
        m.close()
    return 0


def hnc_only(nstate):
    """Solve the c(k) of the first nstate state points of the driven-mode scan
    together on the GPU (hy_hnc_iterate), twice (the first call also builds
    the DST-I table), and print the wall time of each: the batched HNC alone,
    for profiling."""
    from hydft_amd import abi, serve
    from hydft_amd.host import hnc
    from hydft_amd.model import hnc_request
    abi.call("hy_init", 0)
    reqs = serve.chi_requests(nstate, seed=11, grid=(32, 48, 64), steps=400)
    pts = [hnc_request(r.p) for r in reqs]
    nr, rmax, tol, mix, maxiter = pts[0][2:7]
    out = {"states": nstate, "nr": nr}
    for rep in range(2):
        t0 = time.perf_counter()
        res = hnc.hnc_solve_many_device([p[0] for p in pts], [p[1] for p in pts], nr, rmax, tol, mix, maxiter)
        out[f"wall_s_{rep}"] = time.perf_counter() - t0
        its = sorted(r.niter for r in res)
        out["iterations_p50"], out["iterations_max"] = its[len(its) // 2], its[-1]
        out["converged"] = all(r.converged for r in res)
    print(json.dumps(out))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates HBM byte counters)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 HBM counter runs")
    ap.add_argument("--scan", type=int, default=512, metavar="NRUNS",
                    help="batched driven-mode scan: number of runs (0: skip)")
    ap.add_argument("--scan-steps", type=int, default=2000)
    ap.add_argument("--stream-rate", type=float, default=50.0, metavar="PER_S",
                    help="Poisson arrival rate of the serving stream (requests/s)")
    ap.add_argument("--stream-only", type=int, default=0, metavar="NREQ",
                    help="run only the serving stream of NREQ requests at --stream-rate")
    ap.add_argument("--stream", type=int, default=100, metavar="NREQ",
                    help="serving stream: number of requests (0: skip)")
    ap.add_argument("--serve-steps", type=int, default=400, help="steps per run of --batch-only")
    ap.add_argument("--no-maxwell", action="store_true", help="skip the 128^3 Maxwell FP64/mixed sample")
    ap.add_argument("--error-steps", type=int, default=500,
                    help="steps of the mixed-vs-FP64 density error run (0: skip)")
    ap.add_argument("--hnc-only", type=int, default=0, metavar="NSTATE",
                    help="only solve NSTATE state points of the scan with the batched GPU HNC, and time it")
    ap.add_argument("--batch-only", default="", metavar="B,N,CLOSURE",
                    help="only step one batch of B driven-mode runs on N^3 (profiling)")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pmc_child:
        return pmc_child(args)
    if args.hnc_only:
        return hnc_only(args.hnc_only)
    if args.batch_only:
        return batch_only(args.batch_only, args.serve_steps)
    if args.stream_only:
        from hydft_amd import abi
        abi.call("hy_init", 0)
        print(json.dumps(stream_sample(args.stream_only, rate=args.stream_rate)))
        return 0
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; runs on more than one GPU are not part of it "
              "(independent runs would shard with no exchange).", file=sys.stderr)
        return 1

    if args.copy_probe:
        return copy_probe(args.copy_probe, timed=True)
    traffic, traffic_note = (None, "skipped") if args.no_traffic else pmc_traffic(args)
    import torch
    from hydft_amd import abi
    from hydft_amd.model import Model, build_species

    abi.call("hy_init", 0)
    p = config(args.n, args.precision)
    t_setup = time.perf_counter()
    un_sp = build_species(p)  # HNC (Hansen-McDonald), electron_fit transport
    st = un_sp.structure
    m = Model(p, species=un_sp, precision=args.precision)
    n, mom = initial_state(p, m.n0)
    m.set_state(n, mom)
    dt = m.compute_dt()
    m.set_dt(dt)
    m.p.dt = dt
    print(f"# setup {time.perf_counter() - t_setup:.1f} s, dt = {dt:.4e}, n0 = {m.n0:.6f}", file=sys.stderr)
    m.advance(args.warmup)
    torch.cuda.synchronize()
    # per-class pass (untimed, K steps): a HIP event pair around every launch
    # gives each kernel class's time and names the dominant class
    abi.call("hy_model_profile", m.h, 1)
    m.advance(args.steps)
    prof = profile_table(m)
    dom_cls = max(prof, key=lambda r: r["total_ms"])["cls"]
    # timed region: K steps, events only around the dominant class's launches
    # (an event record costs a few us of stream time; one around all ~45
    # launches of a step would slow the step by several per cent)
    abi.call("hy_model_profile_classes", m.h, 1 << dom_cls)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m.advance(args.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    dom = profile_table(m)[0]
    abi.call("hy_model_profile", m.h, 0)
    sps = args.steps / el
    avg_ms = dom["total_ms"] / dom["launches"]
    bpl = dom["bytes"] / dom["launches"]
    achieved = bpl / (avg_ms * 1e-3) / 1e9
    tot_bytes = sum(r["bytes"] for r in prof)
    tot_ms = sum(r["total_ms"] for r in prof)
    extra = {"kernels": [{**r, "avg_us": 1e3 * r["total_ms"] / r["launches"],
                          "GBps": r["bytes"] / (r["total_ms"] * 1e-3) / 1e9} for r in prof],
             "step_hbm_GBps": tot_bytes / (tot_ms * 1e-3) / 1e9,
             "bytes_per_step": tot_bytes / args.steps}
    extra["pencil_plans"] = pencil_plans()
    if traffic:
        extra["traffic_per_launch"] = {r["kernel"]: kernel_traffic(traffic, r["kernel"], args.precision)
                                       for r in prof}
        extra["counter_scale"] = traffic["factors"]
    else:
        extra["traffic_note"] = traffic_note
    # FP64 comparison on the same input: speed and the density error of the mixed run
    if not args.no_fp64 and args.precision == "mixed":
        nm, _, _ = m.get_state()
        m64 = Model(p, species=un_sp, precision="fp64")
        m64.set_state(n, mom)
        m64.set_dt(dt)
        m64.advance(args.warmup + args.steps)  # the mixed run's warmup and per-class pass
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m64.advance(args.steps)
        torch.cuda.synchronize()
        el64 = time.perf_counter() - t0
        n64, _, _ = m64.get_state()
        extra["fp64_steps_per_s"] = args.steps / el64
        extra["mixed_over_fp64"] = sps / extra["fp64_steps_per_s"]
        extra["mixed_density_rel_l2_vs_fp64"] = float(np.linalg.norm(nm - n64) / np.linalg.norm(n64))
        extra["mixed_dn_rel_l2_vs_fp64"] = float(np.linalg.norm(nm - n64) / np.linalg.norm(n64 - m.n0))
        extra["steps_compared"] = args.warmup + 2 * args.steps
        m64.close()
        if args.error_steps > 0:
            extra["mixed_vs_fp64_long"] = mixed_error_run(p, un_sp, n, mom, dt, args.error_steps)
    if args.scan > 0:
        extra["scan"] = scan_sample(args.scan, args.scan_steps)
    if args.stream > 0:
        extra["stream"] = stream_sample(args.stream, rate=args.stream_rate)
    if not args.no_maxwell:
        extra["maxwell_128"] = maxwell_sample()
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(p, un_sp, n, mom)
    out = {
        "metric": "RK3 steps/s at 256^3 (DDFT-QHD electrons, FP32-mixed)" if args.n == 256 else
                  f"RK3 steps/s at {args.n}^3 (DDFT-QHD electrons)",
        "value": sps,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / sps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 transforms, f64 state" if args.precision == "mixed" else "f64",
        "data": "synthetic (seeded shell-random dn and u)",
        "config": {"workload": f"electrons_ddft_qhd_{args.n}^3", "grid": [args.n] * 3, "box": p.lx,
                   "rs": p.rs, "theta": p.theta, "terms": p.terms, "closure": p.closure,
                   "transport": p.transport, "precision": args.precision, "dt": dt},
        "roofline": {"bound": "hbm", "kernel": dom["kernel"], "achieved": achieved, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                     "traffic": kernel_traffic(traffic, dom["kernel"], args.precision),
                     "bytes_per_launch": bpl, "avg_launch_us": avg_ms * 1e3},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
