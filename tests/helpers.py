"""Comparison helpers shared by the parity tests."""
from __future__ import annotations

import numpy as np

# Per-particle position/velocity agreement required after many MVS steps:
# rounding-order differences only (libm cos/pow inside the Kepler solver's
# starting guess are the only non-bitwise operations).
REL_TOL = 1e-11


def rel_err(a, b):
    """Max over bodies of |a-b| / |b| (vector norm per body), and the count of bodies
    that differ at all."""
    a = np.asarray(a); b = np.asarray(b)
    num = np.linalg.norm(a - b, axis=-1)
    den = np.maximum(np.linalg.norm(b, axis=-1), 1e-300)
    r = num / den
    return float(r.max()) if r.size else 0.0, int(np.count_nonzero(num))


def hip_event_rows(events, step):
    """MercuryEvent tuples of one step -> the oracle's flattened rows:
    close-encounter list, then near-miss list, then central impacts,
    each in the library's (min, max) index order (gpu_find_events,
    mercury_gpu.f90:206-234; gpu_central :235-247)."""
    clo, hit, cen = [], [], []
    for (i, j, kind, now, dist, time, xi, xj) in events:
        if kind & 1:
            clo.append((step, 1, i, j, dist, time))
        if kind & 2:
            hit.append((step, 6 if kind & 4 else 2, i, j, dist, time))
        if kind & 8:
            cen.append((step, 8, i, j, dist, time))
    return clo + hit + cen


def adv_event_rows(events):
    """mercury_hip_advance events (system, step, MercuryEvent...) -> rows per step."""
    out = []
    by_step = {}
    for ev in events:
        by_step.setdefault(ev[1], []).append(ev[2:])
    for s in sorted(by_step):
        out += hip_event_rows(by_step[s], s)
    return out


def oracle_rows(evi, evd):
    return [(int(a[0]), int(a[1]), int(a[2]), int(a[3]), float(d[0]), float(d[1]))
            for a, d in zip(evi, evd)]


def compare_rows(got, want, rtol=1e-9):
    """Identical (step, kind, i, j) lists; distance and time within rtol."""
    keys_g = [r[:4] for r in got]
    keys_w = [r[:4] for r in want]
    assert keys_g == keys_w, f"event lists differ:\n got {keys_g[:20]}\nwant {keys_w[:20]}"
    for g, w in zip(got, want):
        for k in (4, 5):
            assert abs(g[k] - w[k]) <= rtol * max(abs(w[k]), 1.0), (g, w)
