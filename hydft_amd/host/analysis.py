"""Host post-processing of probe series (analysis.f90): frequency and damping of
a free oscillation, and the complex amplitude of a driven response."""
from __future__ import annotations

import math

import numpy as np


def fit_damped_mode(t, y, t_start):
    """y ~ A e^{gamma t} cos(omega t + phi) for t >= t_start (analysis.f90:21-74):
    omega from the mean spacing of linearly interpolated zero crossings, gamma
    from a least-squares line through ln|peak| between crossings.
    Returns (omega, gamma, nzeros)."""
    t = np.asarray(t, float)
    y = np.asarray(y, float)
    n = len(t)
    i0 = 0
    while i0 < n - 1 and t[i0] < t_start:
        i0 += 1
    tz = []
    for i in range(i0, n - 1):
        if y[i] == 0.0:
            continue
        if y[i] * y[i + 1] < 0.0:
            w = y[i] / (y[i] - y[i + 1])
            tz.append(t[i] + w * (t[i + 1] - t[i]))
    nz = len(tz)
    if nz < 3:
        return 0.0, 0.0, nz
    omega = math.pi * (nz - 1) / (tz[-1] - tz[0])
    tp, lp = [], []
    ay = np.abs(y)
    for i in range(nz - 1):
        sel = (t > tz[i]) & (t < tz[i + 1])
        if not sel.any():
            continue
        idx = np.nonzero(sel)[0]
        j = idx[np.argmax(ay[idx])]
        if ay[j] > 0.0:
            tp.append(t[j])
            lp.append(math.log(ay[j]))
    if len(tp) < 2:
        return omega, 0.0, nz
    tp = np.array(tp)
    lp = np.array(lp)
    npk = len(tp)
    sx, sy, sxx, sxy = tp.sum(), lp.sum(), (tp * tp).sum(), (tp * lp).sum()
    gamma = (npk * sxy - sx * sy) / (npk * sxx - sx * sx)
    return omega, gamma, nz


def fit_driven_response(t, z, w, t_start):
    """a = <z e^{i w t}> over t >= t_start and its relative scatter
    (analysis.f90:78-98).  For a drive A cos(k x - w t), chi = 2 a / A."""
    t = np.asarray(t, float)
    z = np.asarray(z, complex)
    sel = t >= t_start
    rot = z[sel] * np.exp(1j * w * t[sel])
    cnt = max(int(sel.sum()), 1)
    a = complex(rot.sum() / cnt)
    dev = float(np.max(np.abs(rot - a))) if sel.any() else 0.0
    scatter = dev / max(abs(a), np.finfo(float).tiny)
    return a, scatter
