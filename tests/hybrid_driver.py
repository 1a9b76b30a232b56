"""HYBRID runs through the library with the driver's event listing, for
parity tests: mercuda_amd.hybrid's restatement of the driver's step sequence,
plus what the driver lists after each BS2 substep (the oracle's mce_stat on
the encounter set, mdt_hkce mercury6_2.for:3912-3916) and after each step
(the central impacts of the device's event screen, mercury6_2.for:803-819)."""
from mercuda_amd.hybrid import hy_run


def hy_run_device(hip, oracle, s, nsteps, tol=1e-12, cap=None):
    rows = []

    def on_substep(step, t, hdid, mbs, x0, v0, xbs, vbs, rceb, rphb, nbsbig, index):
        ev = oracle.mce_stat(t, hdid, mbs, x0, v0, xbs, vbs, rceb, rphb, nbsbig)
        for (i, j, d, tt) in ev.clo():
            rows.append((step, 1, index[i - 1] + 1, index[j - 1] + 1, d, tt))

    def on_step(step, time):
        for (i, j, kind, now, dist, t, xi, xj) in hip.events(time, s["h"], s["rcen"]):
            if kind & 8:
                rows.append((step, 8, i, j, dist, t))

    x, v, most = hy_run(hip, s, nsteps, tol=tol, cap=cap, on_substep=on_substep, on_step=on_step)
    return x, v, rows, most
