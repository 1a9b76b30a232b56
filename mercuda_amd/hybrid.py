"""HYBRID steps over libmercury_hip as the MERCURY6 driver takes them.

For each step the driver's device branch of mdt_hy (mercury6_2.for:3650-3660)
calls mercury_hip_hybrid_begin; when encounter pairs exist it integrates the
members with its own mdt_hkce loop (mercury6_2.for:3798-3960), whose BS2
substeps run in the library's compact encounter context; then
mercury_hip_hybrid_finish.  This module restates that host-side sequence for
callers without the Fortran driver (benchmarks, tests).  Collisions are not
resolved here (the driver's mce_coll/mxx_elim); an optional callback receives
each BS2 substep's start and end states so a caller can list encounter minima
as the driver's mce_stat does.
"""
from __future__ import annotations

import math

import numpy as np


def mdt_hkce(hip, time, h0, hrec, tol, s, x, v, ce, pairs, on_substep=None, device_loop=True):
    """Integrate the encounter members over h0 with BS2 substeps; returns the
    BS2 step recommendation kept for the next encounter (hrec).  Without an
    on_substep callback the loop runs on the device (mercury_hip_encounter_advance)
    unless device_loop is False; with one, each substep is a mercury_hip_step
    call whose states the callback receives."""
    m = s["m"]
    nbig = s["nbig"]
    index = [0]
    iback = {}
    nbsbig = 0
    for j in range(1, len(m)):
        if ce[j]:
            index.append(j)
            iback[j] = len(index) - 1
            if j < nbig:
                nbsbig = len(index)
    sel = np.array(index)
    mbs = m[sel].copy()
    xbs = x[sel].copy(); vbs = v[sel].copy()
    xbs[0] = 0.0; vbs[0] = 0.0
    rcb = s["rcrit"][sel].copy(); rceb = s["rce"][sel].copy(); rphb = s["rphys"][sel].copy()
    rcb[0] = rceb[0] = rphb[0] = 0.0
    bp = [(iback[i - 1] + 1, iback[j - 1] + 1) for (i, j) in pairs]
    hip.encounter_enter(mbs, xbs, vbs, rcb, rceb, rphb, nbsbig, bp)
    substeps = 0
    try:
        if on_substep is None and device_loop:
            # no per-substep listing: the whole loop with its step control on the device
            hrec, substeps, _, _ = hip.encounter_advance(h0, hrec, tol)
            mdt_hkce.substeps = getattr(mdt_hkce, "substeps", 0) + substeps
            xbs, vbs = hip.encounter_download()
            for k in range(1, len(index)):
                x[index[k]] = xbs[k]
                v[index[k]] = vbs[k]
            return hrec
        tlocal = 0.0
        hlocal = math.copysign(hrec, h0)
        while True:
            tmp0 = abs(h0) - abs(tlocal)
            hrec = hlocal
            if abs(hlocal) > tmp0:
                hlocal = math.copysign(tmp0, h0)
            hlocal, hdid, _, _ = hip.step(tlocal, hlocal, tol)
            tlocal = tlocal + hdid
            substeps += 1
            if on_substep is not None:
                # the driver pulls every substep for its encounter listing
                x0 = xbs; v0 = vbs
                xbs, vbs = hip.encounter_download()
                on_substep(time + tlocal, hdid, mbs, x0, v0, xbs, vbs, rceb, rphb, nbsbig, index)
            if not ((tlocal - h0) * h0 < 0.0):
                break
        if on_substep is None:
            xbs, vbs = hip.encounter_download()
    finally:
        hip.encounter_exit()
    mdt_hkce.substeps = getattr(mdt_hkce, "substeps", 0) + substeps
    for k in range(1, len(index)):
        x[index[k]] = xbs[k]
        v[index[k]] = vbs[k]
    return hrec


def hy_run(hip, s, nsteps, tol=1e-12, cap=None, time0=0.0, on_substep=None, on_step=None, upload=True,
           progress=None, device_loop=True):
    """nsteps HYBRID steps of system s (democratic heliocentric x, v; keys m,
    x, v, nbig, h, rcrit, rce, rphys, rcen).  Returns the final x, v and the
    largest number of encounter pairs met in one step.  on_step(step, time)
    runs after each step (the driver's event screening point)."""
    h = s["h"]
    if cap is None:      # every particle with every Big body, as the driver sizes CMAX
        cap = max(5000, len(s["m"]) * max(s["nbig"] - 1, 1))
    if upload:
        hip.configure(10)
        hip.upload_system(s)
    x = np.ascontiguousarray(s["x"], np.float64).copy()
    v = np.ascontiguousarray(s["v"], np.float64).copy()
    time = time0
    dtflag = 0
    hrec = h
    most = 0
    for step in range(nsteps):
        if dtflag == 0:
            hrec = h
        ce, pairs, _ = hip.hybrid_begin(h, s["rcrit"], dtflag, cap, x, v)
        dtflag = 2
        most = max(most, len(pairs))
        if pairs:
            sub = None
            if on_substep is not None:
                sub = lambda *a, _s=step: on_substep(_s, *a)
            hrec = mdt_hkce(hip, time, h, hrec, tol, s, x, v, ce, pairs, sub, device_loop)
        hip.hybrid_finish(h, s["m"], x, v)
        time = time + h
        if on_step is not None:
            on_step(step, time)
        if progress is not None:
            progress(step, len(pairs))
    xf, vf = hip.download()
    return xf, vf, most
