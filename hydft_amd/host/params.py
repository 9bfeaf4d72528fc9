"""Run-time parameters from a namelist file (params.f90): the nine groups
&grid &species &hnc &time &initial &external &probe &linear &output with the
reference defaults, plus &gpu (precision) for this build.  Absent groups keep
their defaults; strings are lower-cased; electron kind overrides Gamma from
theta and sets kappa = 0 (params.f90:161-172)."""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field


class ParamsError(RuntimeError):
    pass


@dataclass
class Params:
    # &grid
    nx: int = 64
    ny: int = 1
    nz: int = 1
    lx: float = 20.0
    ly: float = 1.0
    lz: float = 1.0
    dealias: bool = True
    # &species
    kind: str = "ion"
    gamma: float = 2.0
    kappa: float = 0.1
    rs: float = 1.86
    theta: float = -1.0
    mass: float = 1.0
    terms: str = ""
    tfk_gamma: float = 1.0 / 9.0
    closure: str = "newtonian"
    transport: str = "constant"
    eta: float = 0.0
    xi: float = 0.0
    tau: float = 0.0
    tau_model: str = "ichimaru"
    adiabatic_index: float = 5.0 / 3.0
    structure: str = "hnc"
    structure_file: str = ""
    structure_kind: str = "s"
    # &hnc
    hnc_nr: int = 8192
    hnc_maxiter: int = 20000
    hnc_rmax: float = 80.0
    hnc_tol: float = 1.0e-9
    hnc_mix: float = 0.2
    # &time
    dt: float = 0.0
    cfl: float = 0.4
    tend: float = 50.0
    nsteps_max: int = 10000000
    iout: int = 0
    iprobe: int = 1
    idiag: int = 10
    # &initial
    ic: str = "mode"
    ic_amplitude: float = 1.0e-4
    ic_mode: tuple = (1, 0, 0)
    ic_width: float = 1.0
    ic_seed: int = 1234
    # &external
    vext: str = "none"
    vext_amplitude: float = 0.0
    vext_mode: tuple = (1, 0, 0)
    vext_omega: float = 0.0
    vext_width: float = 1.0
    vext_ramp: float = 0.0
    # &probe
    probe_mode: tuple = (1, 0, 0)
    # &linear
    qmin: float = 0.05
    qmax: float = 2.5
    nq: int = 50
    omega_max: float = 3.0
    nomega: int = 600
    q_dsf: tuple = (0.5, 1.0, 1.5, -1.0, -1.0, -1.0, -1.0, -1.0)
    # &output
    prefix: str = "out"
    # &gpu (this build): 'fp64' | 'mixed'
    precision: str = "fp64"

    def finalize(self):
        """params.f90:161-172."""
        if not self.terms.strip():
            self.terms = "tf,gradient,hartree,ry" if self.kind == "electron" else "ideal,hartree,ry"
        if self.kind == "electron" and self.theta > 0.0:
            self.gamma = 2.0 * self.rs / (self.theta * (9.0 * math.pi / 4.0) ** (2.0 / 3.0))
        if self.kind == "electron":
            self.kappa = 0.0
        return self


# namelist variable -> Params field, per group (params.f90:105-114)
_GROUPS = {
    "grid": {k: k for k in ("nx", "ny", "nz", "lx", "ly", "lz", "dealias")},
    "species": {k: k for k in ("kind", "gamma", "kappa", "rs", "theta", "mass", "terms", "tfk_gamma",
                               "closure", "transport", "eta", "xi", "tau", "tau_model",
                               "adiabatic_index", "structure", "structure_file", "structure_kind")},
    "hnc": {"nr": "hnc_nr", "rmax": "hnc_rmax", "tol": "hnc_tol", "mix": "hnc_mix", "maxiter": "hnc_maxiter"},
    "time": {k: k for k in ("dt", "cfl", "tend", "nsteps_max", "iout", "iprobe", "idiag")},
    "initial": {"type": "ic", "amplitude": "ic_amplitude", "mode": "ic_mode", "width": "ic_width",
                "seed": "ic_seed"},
    "external": {"type": "vext", "amplitude": "vext_amplitude", "mode": "vext_mode", "omega": "vext_omega",
                 "width": "vext_width", "ramp": "vext_ramp"},
    "probe": {"mode": "probe_mode"},
    "linear": {k: k for k in ("qmin", "qmax", "nq", "omega_max", "nomega", "q_dsf")},
    "output": {"prefix": "prefix"},
    "gpu": {"precision": "precision"},
}
_LOWER = {"kind", "terms", "closure", "transport", "tau_model", "structure", "structure_kind", "ic", "vext",
          "precision"}


def _scalar(tok: str):
    t = tok.strip()
    if not t:
        return None
    if t[0] in "'\"":
        return t[1:-1] if len(t) >= 2 and t[-1] == t[0] else t[1:]
    lt = t.lower()
    if lt in (".true.", "t", ".t.", "true"):
        return True
    if lt in (".false.", "f", ".f.", "false"):
        return False
    try:
        if re.fullmatch(r"[+-]?\d+", t):
            return int(t)
        return float(lt.replace("d", "e"))
    except ValueError:
        raise ParamsError(f"cannot parse namelist value {tok!r}")


def _split_values(s: str):
    out, cur, q = [], "", None
    for ch in s:
        if q:
            cur += ch
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
            cur += ch
        elif ch in ", \t\n":
            if cur.strip():
                out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return out


def parse_namelists(text: str):
    """Return {group: {var: value or [values]}} for every &group ... / block."""
    # strip comments (! outside quotes)
    lines = []
    for ln in text.splitlines():
        q, buf = None, ""
        for ch in ln:
            if q:
                buf += ch
                if ch == q:
                    q = None
            elif ch in "'\"":
                q = ch
                buf += ch
            elif ch == "!":
                break
            else:
                buf += ch
        lines.append(buf)
    text = "\n".join(lines)
    groups = {}
    for m in re.finditer(r"&(\w+)(.*?)(?:^\s*/|\s/\s*$|/\s*\n|/\s*$)", text, re.S | re.M):
        name = m.group(1).lower()
        body = m.group(2)
        vals = {}
        parts = re.split(r"(\b[A-Za-z_]\w*\s*(?:\(\s*\d+\s*\))?\s*=)", body)
        i = 1
        while i < len(parts) - 1:
            key = parts[i].rstrip("=").strip().lower()
            key = re.sub(r"\s+", "", key)
            raw = _split_values(parts[i + 1])
            v = [_scalar(r) for r in raw]
            vals[key] = v[0] if len(v) == 1 else v
            i += 2
        groups[name] = vals
    return groups


def read_params(path: str) -> Params:
    """params_read (params.f90:82-180)."""
    with open(path) as fh:
        text = fh.read()
    p = Params()
    for gname, vals in parse_namelists(text).items():
        if gname not in _GROUPS:
            raise ParamsError(f"unknown namelist group &{gname} in {path}")
        table = _GROUPS[gname]
        for key, v in vals.items():
            idx = None
            mm = re.fullmatch(r"(\w+)\((\d+)\)", key)
            if mm:
                key, idx = mm.group(1), int(mm.group(2)) - 1
            if key not in table:
                raise ParamsError(f"error reading namelist &{gname} in {path}: unknown variable {key}")
            fname = table[key]
            cur = getattr(p, fname)
            if isinstance(cur, tuple):
                seq = list(cur)
                if idx is not None:
                    seq[idx] = v
                else:
                    vv = v if isinstance(v, list) else [v]
                    seq[:len(vv)] = vv
                v = tuple(seq)
            elif isinstance(cur, bool):
                v = bool(v)
            elif isinstance(cur, int):
                v = int(v)
            elif isinstance(cur, float):
                v = float(v)
            elif isinstance(cur, str):
                v = str(v)
                if fname in _LOWER:
                    v = v.lower()
            setattr(p, fname, v)
    return p.finalize()
