"""Batched and served chi(k, omega) and dispersion runs on one GPU.

A request is one run: its grid, species, closure, time step, step count and
either a drive (kind "chi": v_ext = A cos(k x - w t), the response
chi = 2 <dn_k/n0 e^{i w t}> / A over the settled tail, analysis.f90:78-98) or a
single-mode initial condition (kind "dispersion": frequency and damping of
dn_k/n0, analysis.f90:21-74).

`Server.serve` admits requests in arrival order.  A request takes a slot
when fewer than `max_slots` runs are active and the active runs' device bytes
plus its own stay within `hbm_budget`.  An admitted run has its own model on
its own HIP stream.  Its steps are captured once as a graph of `chunk` steps,
and all ceil(nsteps / chunk) replays are enqueued at once.  The device records
every step's probe row in a ring.  The host loop then only polls the streams.
A finished run is read and analysed and frees its slot; its device memory is
released once no run is in flight, or earlier when the budget needs it
(freeing device memory waits for all queued device work).  The next requests
are admitted as slots and memory free up.  Runs overlap on the device while each
run's own work keeps its stream order, so a run's result does not depend on
what else runs beside it.

With `batch=True` the server admits the arrived requests in groups of the
same shape: same grid, species kind, closure, drive kind, probed mode, time
step policy and step counts within a factor of two, split further by the
models' kept box (for N divisible by 3 the 2/3-rule boundary mode depends on
the box length's rounding).  A batch stops each run at its own step count
(hy_batch_set_stops) and runs as long as its longest member.  A group runs as one batch (hy_batch_*): each
kernel of a step is launched once for all its models (grid z = model), so a
step of B small runs costs the launches of one.  A batch is one slot; its
models' device bytes are charged to the budget as a whole.  A group the
library refuses as a batch runs as separate captured runs.  Batching changes
how the kernels are launched, not what they compute: each run's result is
bitwise the result it has alone.

With `hnc_device=True` the c(k) of arrived requests that are not yet in the
structure cache are solved together on the GPU (hnc.hnc_solve_many_device)
before they are admitted, instead of one host solve per model.

`one_at_a_time` is the reference's way of running the same requests: one
model at a time, steps taken by hy_model_advance calls.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import abi as _abi
from .host import hnc as _hnc
from .host.analysis import fit_damped_mode, fit_driven_response
from .host.initial import InitialError, set_initial_condition
from .host.params import Params

TWOPI = 2.0 * math.pi
DT_EVERY = 20  # src/hydft.f90:42-43


@dataclass
class Request:
    rid: int
    p: Params
    nsteps: int
    kind: str = "chi"            # "chi" | "dispersion"
    mode: tuple = (1, 0, 0)      # probed (and driven / excited) mode
    arrival: float = 0.0         # seconds after the start of the stream
    t_fit: float = -1.0          # analysis window start; < 0: the second half of the samples


@dataclass
class Result:
    rid: int
    kind: str
    arrival: float
    start: float
    done: float
    steps: int
    probe: np.ndarray = field(repr=False)
    value: dict = field(default_factory=dict)
    error: str = ""              # the run's error (a non-positive density, no time step), if any

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def estimate_bytes(p: Params, precision: str) -> int:
    """Upper bound of a model's device memory, the terms of
    hy_model_device_bytes from the grid alone (kept box of the 2/3 rule)."""
    kx = p.nx // 3 + 1
    ky = 2 * (p.ny // 3) + 1 if p.ny > 1 else 1
    kz = 2 * (p.nz // 3) + 1 if p.nz > 1 else 1
    nt = kx * ky * kz
    esz = 8 if precision == "mixed" else 16
    xs = 37 if p.closure == "maxwell" else 17
    b = 16 * nt * 33 + esz * (kx * p.ny * p.nz * xs + kx * ky * p.nz * 24 + nt * 20)
    b += 8 * (nt + max(p.ny * p.nz * 4, 4096)) + 8 * 6601 * 6 + (24 << 20)
    return int(b)


def make_model(req: Request, precision: str, cache=None):
    """Species (c(k) from `cache`), model and initial state of a request."""
    from .model import Model, build_species
    p = req.p
    sp = build_species(p, cache=cache)
    m = Model(p, species=sp, precision=precision)
    if req.kind == "chi":
        m.set_state(np.full((p.nz, p.ny, p.nx), m.n0))
    else:
        m.set_state(set_initial_condition(p, m.n0, kind="mode", mode=req.mode))
    return m


def linear_for_model(m):
    """Closed-form linear response of a model's species (linear_response.f90:
    59-147): the functional kernel K(k) of the model, eta_l = 4/3 eta + xi per
    particle, the Maxwell time."""
    from .host.linear import Linear
    sp, un = m.sp, m.sp.units
    eta_l = (4.0 * sp.eta_code / 3.0 + sp.xi_code) / un.n0
    return Linear(un.n0, lambda k: np.array([m.kernel(float(np.atleast_1d(k)[0]))]), m.p.mass, eta_l, sp.tau,
                  un.beta, un.hbar, un.quantum)


def _mode_k(req: Request) -> float:
    p = req.p
    kx, ky, kz = (TWOPI * req.mode[0] / p.lx, TWOPI * req.mode[1] / p.ly, TWOPI * req.mode[2] / p.lz)
    return math.sqrt(kx * kx + ky * ky + kz * kz)


def analyse(req: Request, rows: np.ndarray, lin=None) -> dict:
    """chi (kind "chi") or omega, gamma (kind "dispersion") from the probe
    series; with `lin` (linear_for_model), the closed form beside it as the
    per-request check."""
    t = rows[:, 0]
    z = rows[:, 1] + 1j * rows[:, 2]
    # default window: the second half of the samples, by index (a start time of
    # t_end / 2 would fall exactly on a sample at a constant dt, and rounding
    # would then decide whether that sample is in)
    n = len(t)
    t0 = req.t_fit if req.t_fit >= 0.0 else (0.5 * (t[n // 2 - 1] + t[n // 2]) if n >= 2 else t[0])
    if req.kind == "chi":
        a, scatter = fit_driven_response(t, z, req.p.vext_omega, t0)
        out = {"chi": 2.0 * a / req.p.vext_amplitude, "scatter": scatter}
        if lin is not None:
            cl = lin.chi(_mode_k(req), req.p.vext_omega)
            out["chi_linear"] = cl
            out["chi_rel_dev"] = float(abs(out["chi"] - cl) / abs(cl))
        return out
    om, ga, nz = fit_damped_mode(t, z.real, t0)
    out = {"omega": om, "gamma": ga, "nzeros": nz}
    if lin is not None:
        w, _ = lin.dispersion(_mode_k(req))
        out["omega_linear"], out["gamma_linear"] = w.real, w.imag
    return out


@dataclass
class _Slot:
    reqs: list
    models: list
    start: float
    bytes: int
    steps: int
    batch: object = None

    def idle(self):
        return self.batch.idle() if self.batch is not None else self.models[0].idle()

    def rows(self, i):
        m = self.models[i]
        return self.batch.probe(i) if self.batch is not None else m.graph_probe()

    def close(self):
        if self.batch is not None:
            self.batch.close()
        for m in self.models:
            m.close()


def steps_bucket(nsteps: int) -> int:
    """Runs whose step counts share a bucket (within a factor of two) share a
    batch: a batch runs as long as its longest member, and a shorter member is
    reported when the batch ends."""
    return max(0, int(math.ceil(math.log2(max(nsteps, 1) / 500.0))))


def batch_key(req: Request) -> tuple:
    """Requests with equal keys run the same kernel sequence (a batch stops each
    run at its own step count)."""
    p = req.p
    return (p.nx, p.ny, p.nz, p.kind, p.closure, p.tau_model, p.vext, p.dt <= 0.0, steps_bucket(req.nsteps),
            tuple(req.mode), req.kind)


class Server:
    def __init__(self, hbm_budget: float = 200e9, max_slots: int = 64, chunk: int = DT_EVERY,
                 precision: str = "mixed", cache=None, batch: bool = False, max_batch: int = 256,
                 hnc_device: bool = False):
        self.hbm_budget = float(hbm_budget)
        self.max_slots = int(max_slots)
        self.chunk = int(chunk)
        self.precision = precision
        self.cache = cache if cache is not None else _hnc.StructureCache()
        self.batch = bool(batch)
        self.max_batch = int(max_batch)
        self.hnc_device = bool(hnc_device)
        self.peak_active = 0      # runs in flight
        self.peak_bytes = 0
        self.batches_launched = 0  # slots: batches and single runs
        # finished models are destroyed only when no run is in flight (or when
        # the budget needs their memory): freeing device memory waits for all
        # queued work on the device and would stall the other runs' launches
        self._retired = []
        self._failed = []  # (request, admission time, message, code) of setups that failed

    def _admit(self, reqs, now: float) -> list:
        """Start `reqs` (one batch key) as one batch, or each as a captured run."""
        # a request whose setup fails (its initial density not positive, say)
        # fails alone: it is reported by serve() and the others start
        models, ok = [], []
        for r in reqs:
            try:
                models.append(make_model(r, self.precision, self.cache))
                ok.append(r)
            except (InitialError, _abi.HydftError) as e:
                self._failed.append((r, now, str(e), getattr(e, "code", _abi.HY_ERR_INVALID)))
        reqs = ok
        if not reqs:
            return []
        req = reqs[0]
        dt_every = DT_EVERY if req.p.dt <= 0.0 else 0
        # each run is charged the larger of the estimate it was admitted on and its actual bytes
        charge = [max(m.device_bytes(), estimate_bytes(r.p, self.precision)) for m, r in zip(models, reqs)]
        # same-key runs can still differ in their kept box (a 2/3-rule boundary
        # mode decided by the box length's rounding): one batch per kept box
        parts = {}
        for r, m, c in zip(reqs, models, charge):
            parts.setdefault(m.kept(), []).append((r, m, c))
        def rounded(r):
            return -(-r.nsteps // self.chunk) * self.chunk

        out = []
        for part in parts.values():
            if len(part) > 1:
                from .model import Batch
                from . import abi
                ms = [m for _, m, _ in part]
                steps = max(rounded(r) for r, _, _ in part)
                try:
                    b = Batch(ms, self.chunk, dt_every=dt_every, probe=req.mode, ring=steps)
                except abi.HydftError as e:
                    if e.code != abi.HY_ERR_INVALID:
                        raise
                else:
                    stops = [rounded(r) for r, _, _ in part]
                    if min(stops) < steps:
                        b.set_stops(stops)
                    b.launch(steps // self.chunk)
                    self.batches_launched += 1
                    out.append(_Slot([r for r, _, _ in part], ms, now, sum(c for _, _, c in part), steps, b))
                    continue
            for r, m, c in part:
                steps = rounded(r)
                m.graph_build(self.chunk, dt_every=dt_every, probe=r.mode, ring=steps)
                m.graph_launch(steps // self.chunk)
                self.batches_launched += 1
                out.append(_Slot([r], [m], now, c, steps))
        return out

    def _take(self, pending, now, realtime, room_bytes):
        """The next group to admit: the first arrived request and, when
        batching, the arrived requests of its key, within the budget."""
        head = pending[0]
        group = [head]
        if self.batch:
            key = batch_key(head)
            need = estimate_bytes(head.p, self.precision)
            for r in pending[1:]:
                if realtime and r.arrival > now:
                    break
                if len(group) >= self.max_batch or need * (len(group) + 1) > room_bytes:
                    break
                if batch_key(r) == key:
                    group.append(r)
        for r in group:
            pending.remove(r)
        return group

    def serve(self, requests, realtime: bool = True):
        """Run the requests; with `realtime`, a request is admitted no earlier
        than its arrival time.  Returns Results in completion order."""
        pending = sorted(requests, key=lambda r: r.arrival)
        active = []
        results = []
        used = 0
        t0 = time.perf_counter()
        while pending or active:
            now = time.perf_counter() - t0
            progressed = False
            if not active and self._retired:
                used -= self._release()
            if self.hnc_device and pending and (not realtime or pending[0].arrival <= now):
                self._prefetch([r for r in pending if not realtime or r.arrival <= now])
            while pending and len(active) < self.max_slots and (not realtime or pending[0].arrival <= now):
                need = estimate_bytes(pending[0].p, self.precision)
                if active and used + need > self.hbm_budget and self._retired:
                    used -= self._release()
                if active and used + need > self.hbm_budget:
                    break
                room = self.hbm_budget - used if active else max(self.hbm_budget - used, need)
                group = self._take(pending, now, realtime, room)
                for s in self._admit(group, now):
                    active.append(s)
                    used += s.bytes
                for r, at, msg, code in self._failed:
                    results.append(Result(r.rid, r.kind, r.arrival, at, time.perf_counter() - t0, 0,
                                          np.zeros((0, 6)), {"code": code}, msg))
                self._failed = []
                self.peak_active = max(self.peak_active, sum(len(s.models) for s in active))
                self.peak_bytes = max(self.peak_bytes, used)
                progressed = True
            for s in list(active):
                if not s.idle():
                    continue
                for i, (req, m) in enumerate(zip(s.reqs, s.models)):
                    try:
                        m.check()
                    except _abi.HydftError as e:  # this request failed; the others go on
                        results.append(Result(req.rid, req.kind, req.arrival, s.start, time.perf_counter() - t0,
                                              0, np.zeros((0, 6)), {"code": e.code}, str(e)))
                        continue
                    rows = s.rows(i)[: req.nsteps]
                    done = time.perf_counter() - t0
                    results.append(Result(req.rid, req.kind, req.arrival, s.start, done, len(rows), rows,
                                          analyse(req, rows, linear_for_model(m))))
                active.remove(s)
                self._retired.append(s)
                progressed = True
            if not progressed:
                time.sleep(0.0002)
        used -= self._release()
        return results

    def _prefetch(self, reqs):
        """c(k) of the requests that need an HNC solve, in one device batch."""
        from .model import hnc_request
        want = [hnc_request(r.p) for r in reqs if r.p.structure == "hnc"]
        if want:
            self.cache.solve_many(want, device=True)

    def _release(self) -> int:
        freed = 0
        for s in self._retired:
            s.close()
            freed += s.bytes
        self._retired = []
        return freed


def one_at_a_time(requests, precision: str = "mixed", cache=None):
    """The same requests, one model at a time, steps by hy_model_advance."""
    cache = cache if cache is not None else _hnc.StructureCache()
    results = []
    t0 = time.perf_counter()
    for req in sorted(requests, key=lambda r: r.arrival):
        start = time.perf_counter() - t0
        m = None
        try:
            m = make_model(req, precision, cache)
            taken, rows = m.advance(req.nsteps, dt_every=DT_EVERY if req.p.dt <= 0.0 else 0, probe=req.mode)
        except (InitialError, _abi.HydftError) as e:  # this request failed; the others go on
            results.append(Result(req.rid, req.kind, req.arrival, start, time.perf_counter() - t0, 0,
                                  np.zeros((0, 6)), {"code": getattr(e, "code", _abi.HY_ERR_INVALID)}, str(e)))
            if m is not None:
                m.close()
            continue
        done = time.perf_counter() - t0
        results.append(Result(req.rid, req.kind, req.arrival, start, done, taken, rows,
                              analyse(req, rows, linear_for_model(m))))
        m.close()
    return results


def summary(results, wall: float) -> dict:
    lat = np.array([r.latency for r in results])
    out = {"runs": len(results), "wall_s": wall, "runs_per_s": len(results) / wall,
           "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
           "failed": sum(1 for r in results if r.error)}
    dev = [r.value["chi_rel_dev"] for r in results if "chi_rel_dev" in r.value]
    if dev:
        out["chi_vs_linear_median"] = float(np.median(dev))
        out["chi_vs_linear_max"] = float(np.max(dev))
    return out


def stream_requests(n: int, seed: int = 5, grids=(32, 64, 96, 128), weights=(0.4, 0.3, 0.2, 0.1),
                    steps=(500, 5000), arrival_rate=None) -> list:
    """Requests of the serving stream: grid drawn from `grids` with `weights`;
    species ions (Gamma log-uniform in [1, 150], kappa uniform in [0.5, 3],
    constant eta = 0.1) or electrons (r_s log-uniform in [0.5, 4], theta
    log-uniform in [0.25, 4], electron_fit viscosity, terms
    tf,gradient,hartree,ry), 50/50; closure Newtonian or Maxwell (tau = 1),
    50/50; kind "chi" (drive 1e-4 cos(q x - w t), w uniform in [0.1, 3]) or
    "dispersion" (single mode of amplitude 1e-4), 50/50; q uniform in [0.2, 2]
    (box 2 pi/q); step count uniform in `steps`, a multiple of 20; CFL time
    step every 20 steps.  `arrival_rate`: Poisson arrivals at that rate (1/s),
    drawn from their own stream (the same runs with or without them)."""
    rng = np.random.default_rng(seed)
    arr = np.random.default_rng([seed, 1])
    cw = np.cumsum(np.asarray(weights, float) / float(np.sum(weights)))
    lo, hi = int(steps[0]) // DT_EVERY, int(steps[1]) // DT_EVERY
    out = []
    t = 0.0
    for i in range(n):
        N = int(grids[int(np.searchsorted(cw, rng.random(), side="right"))])
        electron = rng.random() < 0.5
        maxwell = rng.random() < 0.5
        chi = rng.random() < 0.5
        q = float(rng.uniform(0.2, 2.0))
        p = Params()
        if electron:
            p.kind = "electron"
            p.rs = float(np.exp(rng.uniform(np.log(0.5), np.log(4.0))))
            p.theta = float(np.exp(rng.uniform(np.log(0.25), np.log(4.0))))
            p.transport = "electron_fit"
        else:
            p.kind = "ion"
            p.gamma = float(np.exp(rng.uniform(0.0, np.log(150.0))))
            p.kappa = float(rng.uniform(0.5, 3.0))
            p.eta = 0.1
        p.nx = p.ny = p.nz = N
        p.lx = p.ly = p.lz = TWOPI / q
        if maxwell:
            p.closure, p.tau_model, p.tau = "maxwell", "constant", 1.0
        w = float(rng.uniform(0.1, 3.0))
        nsteps = DT_EVERY * int(rng.integers(lo, hi + 1))
        if chi:
            p.vext, p.vext_amplitude, p.vext_mode = "driven", 1.0e-4, (1, 0, 0)
            p.vext_omega, p.vext_ramp = w, 1.0
        else:
            p.ic, p.ic_amplitude, p.ic_mode = "mode", 1.0e-4, (1, 0, 0)
        p.dt, p.cfl = 0.0, 0.4
        p.finalize()
        if arrival_rate:
            t += float(arr.exponential(1.0 / arrival_rate))
        out.append(Request(i, p, nsteps, "chi" if chi else "dispersion", (1, 0, 0), t))
    return out


def chi_requests(n: int, seed: int = 7, grid=(32,), steps=400, arrival_rate=None, kind="ion") -> list:
    """Driven-mode scan requests in the spirit of the batched configuration:
    Yukawa ions, Gamma log-uniform in [1, 100], kappa uniform in [0.5, 3],
    Newtonian/Maxwell alternating, drive 1e-4 cos(q x - w t) on mode 1 with
    q uniform in [0.2, 2] (box L = 2 pi/q), w uniform in [0.1, 3] w_p
    (w_p = 1 in these units: n0 = 3/(4 pi), e^2 = 1/3); CFL time step
    recomputed every 20 steps, as the reference's driver does; seeded.
    `arrival_rate`: Poisson arrivals at that rate (1/s), else all at t = 0."""
    rng = np.random.default_rng(seed)
    arr = np.random.default_rng([seed, 1])  # arrivals: their own stream (same runs with or without them)
    out = []
    t = 0.0
    for i in range(n):
        N = int(grid[i % len(grid)])
        q = float(rng.uniform(0.2, 2.0))
        p = Params()
        p.kind = kind
        p.gamma = float(np.exp(rng.uniform(np.log(1.0), np.log(100.0))))
        p.kappa = float(rng.uniform(0.5, 3.0))
        p.nx = p.ny = p.nz = N
        p.lx = p.ly = p.lz = TWOPI / q
        p.eta = 0.1
        if i % 2:
            p.closure, p.tau_model, p.tau = "maxwell", "constant", 1.0
        else:
            p.closure = "newtonian"
        p.vext, p.vext_amplitude, p.vext_mode = "driven", 1.0e-4, (1, 0, 0)
        p.vext_omega = float(rng.uniform(0.1, 3.0))
        p.vext_ramp = 1.0
        p.dt, p.cfl = 0.0, 0.4
        p.finalize()
        if arrival_rate:
            t += float(arr.exponential(1.0 / arrival_rate))
        out.append(Request(i, p, steps, "chi", (1, 0, 0), t))
    return out
