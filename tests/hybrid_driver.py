"""The MERCURY6 driver's HYBRID step sequence over the C ABI, for parity
tests: mdt_hy's device branch (mercury6_2.for:3650-3660) with mdt_hkce
(mercury6_2.for:3798-3960, collisions not allowed) run here on the host,
integrating the encounter members through the library's compact context and
listing encounter minima with the oracle's mce_stat, then the central-impact
check the driver makes after every step (mercury6_2.for:803-819)."""
import math

import numpy as np


def mdt_hkce(hip, oracle, time, h0, hrec, tol, s, x, v, ce, pairs, rows, step):
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
    nbs = len(index)
    sel = np.array(index)
    mbs = m[sel].copy()
    xbs = x[sel].copy(); vbs = v[sel].copy()
    xbs[0] = 0.0; vbs[0] = 0.0
    rcb = s["rcrit"][sel].copy(); rceb = s["rce"][sel].copy(); rphb = s["rphys"][sel].copy()
    rcb[0] = rceb[0] = rphb[0] = 0.0
    bp = [(iback[i - 1] + 1, iback[j - 1] + 1) for (i, j) in pairs]
    hip.encounter_enter(mbs, xbs, vbs, rcb, rceb, rphb, nbsbig, bp)
    tlocal = 0.0
    hlocal = math.copysign(hrec, h0)
    while True:
        tmp0 = abs(h0) - abs(tlocal)
        hrec = hlocal
        if abs(hlocal) > tmp0:
            hlocal = math.copysign(tmp0, h0)
        x0 = xbs.copy(); v0 = vbs.copy()
        hip.reset(1)
        hlocal, hdid, _, _ = hip.step(tlocal, hlocal, tol)
        xbs, vbs = hip.encounter_download()
        tlocal = tlocal + hdid
        ev = oracle.mce_stat(time + tlocal, hdid, mbs, x0, v0, xbs, vbs, rceb, rphb, nbsbig)
        for (i, j, d, t) in ev.clo():
            rows.append((step, 1, index[i - 1] + 1, index[j - 1] + 1, d, t))
        if not ((tlocal - h0) * h0 < 0.0):
            break
    hip.encounter_exit()
    for k in range(1, nbs):
        x[index[k]] = xbs[k]
        v[index[k]] = vbs[k]
    return hrec


def hy_run_device(hip, oracle, s, nsteps, tol=1e-12, cap=20000):
    """nsteps HYBRID steps through the library as the unchanged driver makes
    them.  Returns final x, v (democratic heliocentric), event rows and the
    largest number of encounter pairs in one step."""
    h = s["h"]
    hip.configure(10)
    hip.upload_system(s)
    x = np.ascontiguousarray(s["x"], np.float64).copy()
    v = np.ascontiguousarray(s["v"], np.float64).copy()
    rows = []
    time = 0.0
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
            hrec = mdt_hkce(hip, oracle, time, h, hrec, tol, s, x, v, ce, pairs, rows, step)
        hip.hybrid_finish(h, s["m"], x, v)
        time = time + h
        for (i, j, kind, now, dist, t, xi, xj) in hip.events(time, h, s["rcen"]):
            if kind & 8:
                rows.append((step, 8, i, j, dist, t))
    xf, vf = hip.download()
    return xf, vf, rows, most
