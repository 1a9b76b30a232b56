"""ctypes bindings for the checker libraries -- TEST INFRASTRUCTURE ONLY.

``Oracle`` wraps liborc.so, the C restatement of the reference hot path
(mercury_oracle.c).  ``Reference`` wraps _ref/libmercury_ref.so, the reference's
own Fortran routines compiled from mercury6_2.for by oracle/Makefile, called the
way tests/force_library.py in the reference calls them (by reference, with the
/mercury_sizes/ common block set first, mercury.inc:17-21).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use these.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORC_PATH = HERE / "liborc.so"
ORC_FMA_PATH = HERE / "liborc_fma.so"   # the same restatement with multiply-adds fused
REF_PATH = HERE / "_ref" / "libmercury_ref.so"
REF_BIN = HERE / "_ref" / "mercury6"
REFERENCE_SRC = Path("/root/reference")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def build(reference: bool = True) -> None:
    """Compile liborc.so, and _ref/ when the reference sources are present."""
    subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)
    if reference and REFERENCE_SRC.is_dir():
        subprocess.run(["make", "-s", "-C", str(HERE), "-j4", "ref"], check=True)
        # The unchanged driver over libmercury_hip (needs the library built first).
        subprocess.run(["make", "-s", "-C", str(HERE), "-j4", "dropin"], check=True)


class _Events(C.Structure):
    _fields_ = [("cap", C.c_int), ("nclo", C.c_int),
                ("iclo", _ip), ("jclo", _ip), ("dclo", _dp), ("tclo", _dp),
                ("ixvclo", _dp), ("jxvclo", _dp),
                ("nhit", C.c_int), ("ihit", _ip), ("jhit", _ip), ("chit", _ip),
                ("dhit", _dp), ("thit", _dp), ("thit1", C.c_double),
                ("nowflag", C.c_int), ("overflow", C.c_int)]


class EventLists:
    """Owner of the arrays an _Events struct points into."""

    def __init__(self, cap=4096):
        self.cap = cap
        self.iclo = np.zeros(cap, np.int32); self.jclo = np.zeros(cap, np.int32)
        self.dclo = np.zeros(cap); self.tclo = np.zeros(cap)
        self.ixv = np.zeros((cap, 6)); self.jxv = np.zeros((cap, 6))
        self.ihit = np.zeros(cap, np.int32); self.jhit = np.zeros(cap, np.int32)
        self.chit = np.zeros(cap, np.int32)
        self.dhit = np.zeros(cap); self.thit = np.zeros(cap)
        self.s = _Events(cap, 0, _i(self.iclo), _i(self.jclo), _d(self.dclo), _d(self.tclo),
                         _d(self.ixv), _d(self.jxv), 0, _i(self.ihit), _i(self.jhit),
                         _i(self.chit), _d(self.dhit), _d(self.thit), 0.0, 0, 0)

    def clo(self):
        n = self.s.nclo
        return [(int(self.iclo[k]), int(self.jclo[k]), float(self.dclo[k]), float(self.tclo[k]))
                for k in range(n)]

    def hit(self):
        n = self.s.nhit
        return [(int(self.ihit[k]), int(self.jhit[k]), int(self.chit[k]), float(self.dhit[k]),
                 float(self.thit[k])) for k in range(n)]


class Oracle:
    """The C restatement (liborc.so)."""

    def __init__(self, path: Path = ORC_PATH):
        if not Path(path).exists():
            raise FileNotFoundError(f"{path} not built; run make -C oracle")
        L = self.L = C.CDLL(str(path))
        L.orc_drift_one.restype = C.c_int
        L.orc_drift_one.argtypes = [C.c_double, _dp, C.c_double]
        L.orc_drift_branch.restype = C.c_int
        L.orc_drift_branch.argtypes = [C.c_double, _dp, C.c_double]
        L.orc_mco_sine.argtypes = [C.c_double, _dp, _dp]
        L.orc_drift_kepu_stumpff.argtypes = [C.c_double, _dp]
        L.orc_mco_h2j.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_mco_j2h.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_mfo_mvs.argtypes = [_dp, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_mfo_obl.argtypes = [_dp, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_mce_min.argtypes = [C.c_double] * 5 + [_dp, _dp]
        L.orc_mce_stat.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                   _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.POINTER(_Events)]
        L.orc_mce_cent.restype = C.c_int
        L.orc_mce_cent.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                   _dp, _dp, _dp, _dp, _dp, C.c_int, _ip, _dp, _dp]
        L.orc_mdt_mvs.argtypes = [C.c_double, C.c_double, C.c_double, _dp, C.c_int, C.c_int,
                                  _dp, _dp, _dp, _dp, _ip, _dp, _dp, C.POINTER(_Events)]
        L.orc_mco_mvs2h.argtypes = [C.c_double, _dp, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_mfo_all.argtypes = [_dp, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, _dp]
        L.orc_ra15_run.restype = C.c_long
        L.orc_ra15_run.argtypes = [C.c_double, C.c_double, C.c_double, _dp, C.c_int, C.c_int, _dp, _dp, _dp,
                                   C.c_int, C.c_int, _dp, _dp]
        L.orc_mvs_run.restype = C.c_int
        L.orc_mvs_run.argtypes = [C.c_double, C.c_double, C.c_double, _dp, C.c_int, C.c_int,
                                  _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, _ip, _dp]
        L.orc_hy_run.restype = C.c_int
        L.orc_hy_run.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, _dp, C.c_int, C.c_int,
                                 _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_int, _ip, _dp, C.POINTER(C.c_long)]
        L.orc_mfo_hy.argtypes = [_dp, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_mco_dh2h.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_mfo_hkce.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int, _ip, _ip]
        L.orc_mdt_bs2.restype = C.c_int
        L.orc_mdt_bs2.argtypes = [C.c_double, _dp, _dp, C.c_double, C.c_int, _dp, _dp, _dp, _dp, C.c_int, _ip, _ip,
                                  C.POINTER(C.c_long)]
        L.orc_bs_run.restype = C.c_long
        L.orc_bs_run.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _dp, C.c_int, C.c_int, _dp, _dp,
                                 _dp, C.c_int, C.c_int, _dp, _dp, C.POINTER(C.c_long)]
        L.orc_set_ngf.argtypes = [C.c_int, _dp, C.c_int]
        L.orc_mfo_ngf.restype = C.c_int
        L.orc_mfo_ngf.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.orc_mfo_pr.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_mce_snif.restype = C.c_int
        L.orc_mce_snif.argtypes = [C.c_double, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _ip, C.c_int, _ip, _ip]

    def drift_one(self, mu, xv, dt):
        w = np.array(xv, dtype=np.float64).copy()
        flag = self.L.orc_drift_one(mu, _d(w), dt)
        return w, flag

    def drift_branch(self, mu, xv, dt):
        w = np.ascontiguousarray(xv, dtype=np.float64)
        return self.L.orc_drift_branch(mu, _d(w), dt)

    def mco_sine(self, x):
        s = np.zeros(1); c = np.zeros(1)
        self.L.orc_mco_sine(x, _d(s), _d(c))
        return s[0], c[0]

    def stumpff(self, x):
        c = np.zeros(4)
        self.L.orc_drift_kepu_stumpff(x, _d(c))
        return c

    def mfo_mvs(self, m, x, nbig, jcen=(0.0, 0.0, 0.0)):
        m = np.ascontiguousarray(m, np.float64); x = np.ascontiguousarray(x, np.float64)
        n = len(m)
        xj = np.zeros_like(x); vj = np.zeros_like(x)
        self.L.orc_mco_h2j(nbig, _d(m), _d(x), _d(np.zeros_like(x)), _d(xj), _d(vj))
        for j in range(nbig, n):
            xj[j] = x[j]
        a = np.zeros_like(x)
        jc = np.ascontiguousarray(jcen, np.float64)
        self.L.orc_mfo_mvs(_d(jc), n, nbig, _d(m), _d(x), _d(xj), _d(a))
        return a

    @contextlib.contextmanager
    def ngf_model(self, sysd):
        """The system's non-gravitational model (sysd["ngflag"], sysd["ngf"][n, 5])
        for the calls inside the block (mfo_all, RA15, BS, MVS)."""
        flag = int(sysd.get("ngflag", 0))
        if not flag:
            yield
            return
        ngf = np.ascontiguousarray(sysd["ngf"], np.float64)
        self.L.orc_set_ngf(flag, _d(ngf), len(ngf))
        try:
            yield
        finally:
            self.L.orc_set_ngf(0, None, 0)

    def mfo_ngf(self, x, v, ngf):
        x = np.ascontiguousarray(x, np.float64); v = np.ascontiguousarray(v, np.float64)
        ngf = np.ascontiguousarray(ngf, np.float64)
        a = np.zeros_like(x)
        bad = self.L.orc_mfo_ngf(len(x), _d(x), _d(v), _d(a), _d(ngf))
        return a, bad

    def mfo_pr(self, m, x, v, ngf):
        m = np.ascontiguousarray(m, np.float64)
        x = np.ascontiguousarray(x, np.float64); v = np.ascontiguousarray(v, np.float64)
        ngf = np.ascontiguousarray(ngf, np.float64)
        a = np.zeros_like(x)
        self.L.orc_mfo_pr(len(x), _d(m), _d(x), _d(v), _d(a), _d(ngf))
        return a

    def mce_min(self, d0, d1, d0t, d1t, h):
        a = np.zeros(1); b = np.zeros(1)
        self.L.orc_mce_min(d0, d1, d0t, d1t, h, _d(a), _d(b))
        return a[0], b[0]

    def mce_stat(self, time, h, m, x0, v0, x1, v1, rce, rphys, nbig, cap=4096):
        ev = EventLists(cap)
        arrs = [np.ascontiguousarray(a, np.float64) for a in (m, x0, v0, x1, v1, rce, rphys)]
        self.L.orc_mce_stat(time, h, 0.0, len(arrs[0]), nbig, *[_d(a) for a in arrs], C.byref(ev.s))
        return ev

    def mce_stat_lists(self, time, h, m, x0, v0, x1, v1, rce, rphys, nbig, cap=4096):
        """mce_stat's outputs in the layout Reference.mce_stat returns (1-based
        indices as the Fortran's)."""
        ev = self.mce_stat(time, h, m, x0, v0, x1, v1, rce, rphys, nbig, cap)
        k, q = ev.s.nclo, ev.s.nhit
        return dict(iclo=ev.iclo[:k].copy(), jclo=ev.jclo[:k].copy(), dclo=ev.dclo[:k].copy(),
                    tclo=ev.tclo[:k].copy(), ixv=ev.ixv[:k].copy(), jxv=ev.jxv[:k].copy(),
                    ihit=ev.ihit[:q].copy(), jhit=ev.jhit[:q].copy(), chit=ev.chit[:q].copy(),
                    dhit=ev.dhit[:q].copy(), thit=ev.thit[:q].copy(), thit1=ev.s.thit1,
                    nowflag=ev.s.nowflag)

    def mce_cent(self, time, h, rcen, m, x0, v0, x1, v1, i0=2):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (m, x0, v0, x1, v1)]
        n = len(arrs[0])
        j = np.zeros(max(n, 1), np.int32); t = np.zeros(max(n, 1)); d = np.zeros(max(n, 1))
        k = self.L.orc_mce_cent(time, h, rcen, i0, n, *[_d(a) for a in arrs], n, _i(j), _d(t), _d(d))
        return [(int(j[q]), float(t[q]), float(d[q])) for q in range(k)]

    def mvs_run(self, sysd, nsteps, time0=0.0, screen=True, cap=200000):
        """nsteps MVS steps of a whole system dict (see mercuda_amd.systems)."""
        m = np.ascontiguousarray(sysd["m"], np.float64)
        x = np.ascontiguousarray(sysd["x"], np.float64).copy()
        v = np.ascontiguousarray(sysd["v"], np.float64).copy()
        rce = np.ascontiguousarray(sysd["rce"], np.float64)
        rphys = np.ascontiguousarray(sysd["rphys"], np.float64)
        jc = np.ascontiguousarray(sysd.get("jcen", (0.0, 0.0, 0.0)), np.float64)
        evi = np.zeros((cap, 4), np.int32); evd = np.zeros((cap, 2))
        with self.ngf_model(sysd):
            k = self.L.orc_mvs_run(time0, sysd["h"], sysd["rcen"], _d(jc), len(m), sysd["nbig"],
                                   _d(m), _d(x), _d(v), _d(rce), _d(rphys), nsteps, int(screen),
                                   cap, _i(evi), _d(evd))
        if k < 0:
            raise RuntimeError("oracle event capacity exceeded")
        return x, v, evi[:k].copy(), evd[:k].copy()

    def mfo_all(self, sysd, x=None, v=None, pn=0):
        m = np.ascontiguousarray(sysd["m"], np.float64)
        x = np.ascontiguousarray(sysd["x"] if x is None else x, np.float64)
        v = np.ascontiguousarray(sysd["v"] if v is None else v, np.float64)
        jc = np.ascontiguousarray(sysd.get("jcen", (0.0, 0.0, 0.0)), np.float64)
        a = np.zeros_like(x)
        with self.ngf_model(sysd):
            self.L.orc_mfo_all(_d(jc), len(m), sysd["nbig"], _d(m), _d(x), _d(v), int(pn), _d(a))
        return a

    def ra15_run(self, sysd, nsteps, h0=None, tol=1e-12, pn=0, time0=0.0):
        """nsteps RA15 sequences of a whole system: final x, v, the sizes done,
        the next proposal and the force-evaluation count."""
        m = np.ascontiguousarray(sysd["m"], np.float64)
        x = np.ascontiguousarray(sysd["x"], np.float64).copy()
        v = np.ascontiguousarray(sysd["v"], np.float64).copy()
        jc = np.ascontiguousarray(sysd.get("jcen", (0.0, 0.0, 0.0)), np.float64)
        hdid = np.zeros(max(nsteps, 1)); hn = np.zeros(1)
        with self.ngf_model(sysd):
            nf = self.L.orc_ra15_run(time0, sysd["h"] if h0 is None else h0, tol, _d(jc), len(m), sysd["nbig"],
                                     _d(m), _d(x), _d(v), int(pn), nsteps, _d(hdid), _d(hn))
        return x, v, hdid[:nsteps].copy(), float(hn[0]), int(nf)

    def bs_run(self, algor, sysd, nsteps, h0=None, tol=1e-12, pn=0, time0=0.0):
        """nsteps steps of algorithm 2 (mdt_bs1) or 3 (mdt_bs2) with mfo_all on a
        whole system: final x, v, the sizes done, the next proposal, the halvings
        and the force-evaluation count."""
        m = np.ascontiguousarray(sysd["m"], np.float64)
        x = np.ascontiguousarray(sysd["x"], np.float64).copy()
        v = np.ascontiguousarray(sysd["v"], np.float64).copy()
        jc = np.ascontiguousarray(sysd.get("jcen", (0.0, 0.0, 0.0)), np.float64)
        hdid = np.zeros(max(nsteps, 1)); hn = np.zeros(1); rej = C.c_long(0)
        with self.ngf_model(sysd):
            nf = self.L.orc_bs_run(int(algor), time0, sysd["h"] if h0 is None else h0, tol, _d(jc), len(m),
                                   sysd["nbig"], _d(m), _d(x), _d(v), int(pn), nsteps, _d(hdid), _d(hn),
                                   C.byref(rej))
        if nf < 0:
            raise RuntimeError(f"oracle BS run failed ({nf})")
        return x, v, hdid[:nsteps].copy(), float(hn[0]), int(rej.value), int(nf)

    def hy_run(self, sysd, nsteps, time0=0.0, tol=1e-12, cap=200000):
        """nsteps HYBRID steps (mdt_hy + the driver's mce_cent on physical
        coordinates), collisions not allowed.  Returns x, v (democratic
        heliocentric), event rows, force-evaluation count."""
        m = np.ascontiguousarray(sysd["m"], np.float64)
        x = np.ascontiguousarray(sysd["x"], np.float64).copy()
        v = np.ascontiguousarray(sysd["v"], np.float64).copy()
        jc = np.ascontiguousarray(sysd.get("jcen", (0.0, 0.0, 0.0)), np.float64)
        arrs = [np.ascontiguousarray(sysd[k], np.float64) for k in ("rcrit", "rce", "rphys")]
        evi = np.zeros((cap, 4), np.int32); evd = np.zeros((cap, 2)); nf = C.c_long(0)
        with self.ngf_model(sysd):
            k = self.L.orc_hy_run(time0, sysd["h"], tol, sysd["rcen"], _d(jc), len(m), sysd["nbig"], _d(m), _d(x),
                                  _d(v), *[_d(a) for a in arrs], nsteps, cap, _i(evi), _d(evd), C.byref(nf))
        if k < 0:
            raise RuntimeError("oracle HYBRID run failed (capacity or BS2 stall)")
        return x, v, evi[:k].copy(), evd[:k].copy(), nf.value

    def mfo_hkce(self, m, x, rcrit, pairs):
        m = np.ascontiguousarray(m, np.float64); x = np.ascontiguousarray(x, np.float64)
        rc = np.ascontiguousarray(rcrit, np.float64)
        pi = np.ascontiguousarray([p[0] for p in pairs], np.int32); pj = np.ascontiguousarray([p[1] for p in pairs], np.int32)
        a = np.zeros_like(x)
        self.L.orc_mfo_hkce(len(m), _d(m), _d(x), _d(rc), _d(a), len(pairs), _i(pi), _i(pj))
        return a

    def bs2(self, m, x, v, rcrit, pairs, h0, tol=1e-12, time=0.0):
        """One mdt_bs2 call: (x, v, h done, next h, halvings, force evaluations)."""
        m = np.ascontiguousarray(m, np.float64)
        x = np.ascontiguousarray(x, np.float64).copy(); v = np.ascontiguousarray(v, np.float64).copy()
        rc = np.ascontiguousarray(rcrit, np.float64)
        pi = np.ascontiguousarray([p[0] for p in pairs] or [0], np.int32)
        pj = np.ascontiguousarray([p[1] for p in pairs] or [0], np.int32)
        h = C.c_double(h0); hd = C.c_double(0.0); nf = C.c_long(0)
        r = self.L.orc_mdt_bs2(time, C.byref(h), C.byref(hd), tol, len(m), _d(m), _d(x), _d(v), _d(rc), len(pairs),
                               _i(pi), _i(pj), C.byref(nf))
        return x, v, hd.value, h.value, r, nf.value

    def mfo_hy(self, sysd, x=None):
        m = np.ascontiguousarray(sysd["m"], np.float64)
        x = np.ascontiguousarray(sysd["x"] if x is None else x, np.float64)
        jc = np.ascontiguousarray(sysd.get("jcen", (0.0, 0.0, 0.0)), np.float64)
        a = np.zeros_like(x)
        self.L.orc_mfo_hy(_d(jc), len(m), sysd["nbig"], _d(m), _d(x), _d(np.ascontiguousarray(sysd["rcrit"])), _d(a))
        return a

    def dh2h(self, sysd, x, v):
        m = np.ascontiguousarray(sysd["m"], np.float64)
        x = np.ascontiguousarray(x, np.float64); v = np.ascontiguousarray(v, np.float64)
        xh = np.zeros_like(x); vh = np.zeros_like(v)
        self.L.orc_mco_dh2h(len(m), _d(m), _d(x), _d(v), _d(xh), _d(vh))
        return xh, vh

    def mvs2h(self, sysd, x, v):
        m = np.ascontiguousarray(sysd["m"], np.float64)
        jc = np.ascontiguousarray(sysd.get("jcen", (0.0, 0.0, 0.0)), np.float64)
        x = np.ascontiguousarray(x, np.float64); v = np.ascontiguousarray(v, np.float64)
        xh = np.zeros_like(x); vh = np.zeros_like(v)
        with self.ngf_model(sysd):
            self.L.orc_mco_mvs2h(sysd["h"], _d(jc), len(m), sysd["nbig"], _d(m), _d(x), _d(v),
                                 _d(xh), _d(vh))
        return xh, vh


class Reference:
    """The reference's own Fortran routines (oracle/_ref/libmercury_ref.so)."""

    NMESS = 200

    def __init__(self, path: Path = REF_PATH, nmax=20000, cmax=20000):
        if not Path(path).exists():
            raise FileNotFoundError(f"{path} not built; run make -C oracle ref")
        L = self.L = C.CDLL(str(path))
        sizes = (C.c_int * 2).in_dll(L, "mercury_sizes_")
        sizes[0] = nmax
        sizes[1] = cmax
        self.nmax, self.cmax = nmax, cmax

    @staticmethod
    def _r(v):
        return C.byref(C.c_double(v))

    def drift_one(self, mu, xv, dt):
        w = [C.c_double(float(q)) for q in xv]
        fl = C.c_int(0)
        self.L.drift_one_(self._r(mu), *[C.byref(q) for q in w], self._r(dt), C.byref(fl))
        return np.array([q.value for q in w]), fl.value

    def mco_sine(self, x):
        xx = C.c_double(x); s = C.c_double(); c = C.c_double()
        self.L.mco_sine_(C.byref(xx), C.byref(s), C.byref(c))
        return s.value, c.value

    def stumpff(self, x):
        xx = C.c_double(x); c = [C.c_double() for _ in range(4)]
        self.L.drift_kepu_stumpff_(C.byref(xx), *[C.byref(q) for q in c])
        return np.array([q.value for q in c])

    def mfo_mvs(self, m, x, nbig, jcen=(0.0, 0.0, 0.0)):
        m = np.ascontiguousarray(m, np.float64); x = np.ascontiguousarray(x, np.float64)
        n = len(m)
        xj = np.zeros_like(x); vj = np.zeros_like(x); v = np.zeros_like(x)
        ngf = np.zeros((n, 5)); opt = np.zeros(8, np.int32)
        jc = np.ascontiguousarray(jcen, np.float64)
        z = C.c_double(0.0); ng = C.c_int(0)
        self.L.mco_h2j_(C.byref(z), _d(jc), C.byref(C.c_int(nbig)), C.byref(C.c_int(nbig)),
                        C.byref(z), _d(m), _d(x), _d(v), _d(xj), _d(vj), _d(ngf), C.byref(ng), _i(opt))
        for j in range(nbig, n):
            xj[j] = x[j]
        a = np.zeros_like(x); stat = np.zeros(n, np.int32)
        self.L.mfo_mvs_(_d(jc), C.byref(C.c_int(n)), C.byref(C.c_int(nbig)), _d(m), _d(x),
                        _d(xj), _d(a), _i(stat))
        return a

    def mfo_all(self, sysd, x=None, v=None, pn=0):
        m = np.ascontiguousarray(sysd["m"], np.float64)
        x = np.ascontiguousarray(sysd["x"] if x is None else x, np.float64)
        v = np.ascontiguousarray(sysd["v"] if v is None else v, np.float64)
        n = len(m)
        jc = np.ascontiguousarray(sysd.get("jcen", (0.0, 0.0, 0.0)), np.float64)
        a = np.zeros_like(x); s = np.zeros((n, 3)); rc = np.zeros(n); st = np.zeros(n, np.int32)
        ngf = np.ascontiguousarray(sysd.get("ngf", np.zeros((n, 5))), np.float64)
        ngflag = C.c_int(int(sysd.get("ngflag", 0)))
        opt = np.array([0, 1, 1, 2, 0, 1, int(pn), 0], np.int32)
        self.L.mfo_all_(self._r(0.0), _d(jc), C.byref(C.c_int(n)), C.byref(C.c_int(sysd["nbig"])), _d(m),
                        _d(x), _d(v), _d(s), _d(rc), _d(a), _i(st), _d(ngf), C.byref(ngflag), _i(opt),
                        C.byref(C.c_int(0)), _i(np.zeros(1, np.int32)), _i(np.zeros(1, np.int32)))
        return a

    def ra15_run(self, sysd, nsteps, h0=None, tol=1e-12, pn=0, time0=0.0, until=None):
        """nsteps calls of the reference mdt_ra15 with mfo_all as the force (or,
        with `until`, calls until that time, the last size clipped to it)."""
        m = np.ascontiguousarray(sysd["m"], np.float64)
        n = len(m)
        x = np.ascontiguousarray(sysd["x"], np.float64).copy()
        v = np.ascontiguousarray(sysd["v"], np.float64).copy()
        jc = np.ascontiguousarray(sysd.get("jcen", (0.0, 0.0, 0.0)), np.float64)
        s = np.zeros((n, 3)); rphys = np.zeros(n); rc = np.zeros(n); st = np.zeros(n, np.int32)
        ngf = np.ascontiguousarray(sysd.get("ngf", np.zeros((n, 5))), np.float64)
        ngflag = C.c_int(int(sysd.get("ngflag", 0)))
        opt = np.array([0, 1, 1, 2, 0, 1, int(pn), 0], np.int32)
        force = C.cast(self.L.mfo_all_, C.c_void_p)
        t = C.c_double(sysd["h"] if h0 is None else h0); tdid = C.c_double(0.0)
        dtflag = C.c_int(0)
        hd = []
        time = time0
        k = 0
        while (k < nsteps) if until is None else (time < until):
            k += 1
            if until is not None and abs(t.value) > until - time:
                t.value = until - time
            self.L.mdt_ra15_(self._r(time), C.byref(t), C.byref(tdid), self._r(tol), _d(jc), C.byref(C.c_int(n)),
                             C.byref(C.c_int(sysd["nbig"])), _d(m), _d(x), _d(v), _d(s), _d(rphys), _d(rc),
                             _d(ngf), _i(st), C.byref(dtflag), C.byref(ngflag), _i(opt),
                             C.byref(C.c_int(0)), _i(np.zeros(1, np.int32)), _i(np.zeros(1, np.int32)), force)
            hd.append(tdid.value)
            time = time + tdid.value
        return x, v, np.array(hd), t.value

    def bs_run(self, algor, sysd, nsteps, h0=None, tol=1e-12, pn=0, time0=0.0):
        """nsteps calls of the reference mdt_bs1 (algor 2) or mdt_bs2 (algor 3)
        with mfo_all as the force, as mal_hvar makes them."""
        m = np.ascontiguousarray(sysd["m"], np.float64)
        n = len(m)
        x = np.ascontiguousarray(sysd["x"], np.float64).copy()
        v = np.ascontiguousarray(sysd["v"], np.float64).copy()
        jc = np.ascontiguousarray(sysd.get("jcen", (0.0, 0.0, 0.0)), np.float64)
        s = np.zeros((n, 3)); rphys = np.zeros(n); rc = np.zeros(n); st = np.zeros(n, np.int32)
        ngf = np.ascontiguousarray(sysd.get("ngf", np.zeros((n, 5))), np.float64)
        ngflag = C.c_int(int(sysd.get("ngflag", 0)))
        opt = np.array([0, 1, 1, 2, 0, 1, int(pn), 0], np.int32)
        force = C.cast(self.L.mfo_all_, C.c_void_p)
        fn = self.L.mdt_bs1_ if algor == 2 else self.L.mdt_bs2_
        t = C.c_double(sysd["h"] if h0 is None else h0); tdid = C.c_double(0.0)
        dtflag = C.c_int(2)
        hd = []
        time = time0
        for _ in range(nsteps):
            fn(self._r(time), C.byref(t), C.byref(tdid), self._r(tol), _d(jc), C.byref(C.c_int(n)),
               C.byref(C.c_int(sysd["nbig"])), _d(m), _d(x), _d(v), _d(s), _d(rphys), _d(rc), _d(ngf), _i(st),
               C.byref(dtflag), C.byref(ngflag), _i(opt), C.byref(C.c_int(0)), _i(np.zeros(1, np.int32)),
               _i(np.zeros(1, np.int32)), force)
            hd.append(tdid.value)
            time = time + tdid.value
        return x, v, np.array(hd), t.value

    def mce_min(self, d0, d1, d0t, d1t, h):
        a = C.c_double(); b = C.c_double()
        self.L.mce_min_(self._r(d0), self._r(d1), self._r(d0t), self._r(d1t), self._r(h),
                        C.byref(a), C.byref(b))
        return a.value, b.value

    def mdt_hy_run(self, sysd, nsteps, time0=0.0, tol=1e-12):
        """nsteps calls of the reference mdt_hy (collisions not allowed) plus the
        driver's mce_cent on physical coordinates (mco_dh2h) after each step, as
        mal_hcon does for algorithm 10.  Rows as Oracle.hy_run."""
        m = np.ascontiguousarray(sysd["m"], np.float64)
        n = len(m)
        nbig = sysd["nbig"]
        x = np.ascontiguousarray(sysd["x"], np.float64).copy()
        v = np.ascontiguousarray(sysd["v"], np.float64).copy()
        rce = np.ascontiguousarray(sysd["rce"], np.float64)
        rphys = np.ascontiguousarray(sysd["rphys"], np.float64)
        rcrit = np.ascontiguousarray(sysd["rcrit"], np.float64)
        jc = np.ascontiguousarray(sysd.get("jcen", (0.0, 0.0, 0.0)), np.float64)
        h = sysd["h"]; rcen = sysd["rcen"]
        cm = self.cmax
        s = np.zeros((n, 3)); stat = np.zeros(n, np.int32)
        ngf = np.ascontiguousarray(sysd.get("ngf", np.zeros((n, 5))), np.float64)
        ngflag = int(sysd.get("ngflag", 0))
        opt = np.array([0, 0, 0, 2, 0, 1, 0, 0], np.int32)
        ids = C.create_string_buffer(b" " * (25 * n))
        outfile = C.create_string_buffer(b"/dev/null".ljust(80) * 3)
        mem = C.create_string_buffer(b" " * (80 * self.NMESS))
        lmem = np.ones(self.NMESS, np.int32)
        iclo = np.zeros(cm, np.int32); jclo = np.zeros(cm, np.int32)
        dclo = np.zeros(cm); tclo = np.zeros(cm)
        ixv = np.zeros((cm, 6)); jxv = np.zeros((cm, 6))
        nclo = C.c_int(0); colflag = C.c_int(0); dtflag = C.c_int(0)
        en = np.zeros(3); am = np.zeros(3)
        rows_i = []; rows_d = []
        time = time0
        xh0 = np.zeros_like(x); vh0 = np.zeros_like(v); xh = np.zeros_like(x); vh = np.zeros_like(v)
        zero = C.c_int(0)

        def dh2h(xa, va, xo, vo):
            self.L.mco_dh2h_(self._r(time), _d(jc), C.byref(C.c_int(n)), C.byref(C.c_int(nbig)), self._r(h), _d(m),
                             _d(xa), _d(va), _d(xo), _d(vo), _d(ngf), C.byref(zero), _i(opt))

        for step in range(nsteps):
            dh2h(x, v, xh0, vh0)
            self.L.mdt_hy_(self._r(time), self._r(time0), self._r(h), self._r(tol),
                           self._r(100.0), _d(en), _d(am), _d(jc), self._r(rcen),
                           C.byref(C.c_int(n)), C.byref(C.c_int(nbig)), _d(m), _d(x), _d(v),
                           _d(s), _d(rphys), _d(rcrit), _d(rce), _i(stat), ids,
                           _d(ngf), C.byref(C.c_int(10)), _i(opt), C.byref(dtflag),
                           C.byref(C.c_int(ngflag)), C.byref(C.c_int(-1)), C.byref(colflag),
                           C.byref(nclo), _i(iclo), _i(jclo), _d(dclo), _d(tclo), _d(ixv), _d(jxv),
                           outfile, mem, _i(lmem),
                           C.c_size_t(25), C.c_size_t(80), C.c_size_t(80))
            time = time + h
            for k in range(nclo.value):
                rows_i.append((step, 1, int(iclo[k]), int(jclo[k])))
                rows_d.append((float(dclo[k]), float(tclo[k])))
            dh2h(x, v, xh, vh)
            xh0[0] = vh0[0] = xh[0] = vh[0] = 0.0
            for (j, t, d) in self.mce_cent(time, h, rcen, m, xh0, vh0, xh, vh):
                rows_i.append((step, 8, 1, j)); rows_d.append((d, t))
        return x, v, np.array(rows_i, np.int32).reshape(-1, 4), np.array(rows_d).reshape(-1, 2)

    def mvs2h(self, sysd, x, v):
        """The reference's symplectic corrector mco_mvs2h (mercury6_2.for:2107-2219)."""
        m = np.ascontiguousarray(sysd["m"], np.float64)
        n = len(m)
        jc = np.ascontiguousarray(sysd.get("jcen", (0.0, 0.0, 0.0)), np.float64)
        x = np.ascontiguousarray(x, np.float64); v = np.ascontiguousarray(v, np.float64)
        xh = np.zeros_like(x); vh = np.zeros_like(v)
        ngf = np.ascontiguousarray(sysd.get("ngf", np.zeros((n, 5))), np.float64)
        opt = np.zeros(8, np.int32)
        self.L.mco_mvs2h_(self._r(0.0), _d(jc), C.byref(C.c_int(n)), C.byref(C.c_int(sysd["nbig"])),
                          self._r(sysd["h"]), _d(m), _d(x), _d(v), _d(xh), _d(vh), _d(ngf),
                          C.byref(C.c_int(int(sysd.get("ngflag", 0)))), _i(opt))
        return xh, vh

    def mce_cent(self, time, h, rcen, m, x0, v0, x1, v1, i0=2):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (m, x0, v0, x1, v1)]
        n = len(arrs[0])
        jhit = np.zeros(self.cmax, np.int32); thit = np.zeros(self.cmax); dhit = np.zeros(self.cmax)
        nhit = C.c_int(0); jc = np.zeros(3); ngf = np.zeros((max(n, 1), 5))
        self.L.mce_cent_(self._r(time), self._r(h), self._r(rcen), _d(jc), C.byref(C.c_int(i0)),
                         C.byref(C.c_int(n)), C.byref(C.c_int(n)), _d(arrs[0]), _d(arrs[1]),
                         _d(arrs[2]), _d(arrs[3]), _d(arrs[4]), C.byref(nhit), _i(jhit),
                         _d(thit), _d(dhit), C.byref(C.c_int(1)), _d(ngf), C.byref(C.c_int(0)))
        return [(int(jhit[q]), float(thit[q]), float(dhit[q])) for q in range(nhit.value)]

    def mfo_ngf(self, x, v, ngf):
        """The reference mfo_ngf (mercury6_2.for:5106-5152) on every body."""
        x = np.ascontiguousarray(x, np.float64); v = np.ascontiguousarray(v, np.float64)
        ngf = np.ascontiguousarray(ngf, np.float64)
        a = np.zeros_like(x)
        self.L.mfo_ngf_(C.byref(C.c_int(len(x))), _d(x), _d(v), _d(a), _d(ngf))
        return a

    def mfo_pr(self, m, nbig, x, v, ngf):
        """The reference mfo_pr (mercury6_2.for:5307-5357); ngf[n, 5] as the
        driver holds it (the routine reads the first four of each body's
        five, so the array is repacked to ngf(4,n))."""
        m = np.ascontiguousarray(m, np.float64)
        x = np.ascontiguousarray(x, np.float64); v = np.ascontiguousarray(v, np.float64)
        ng4 = np.ascontiguousarray(np.asarray(ngf, np.float64)[:, :4])
        a = np.zeros_like(x)
        self.L.mfo_pr_(C.byref(C.c_int(len(x))), C.byref(C.c_int(nbig)), _d(m), _d(x), _d(v), _d(a), _d(ng4))
        return a

    def mce_stat(self, time, h, m, x0, v0, x1, v1, rce, rphys, nbig, nclo0=0):
        """The reference mce_stat (mercury6_2.for:1765-1931) on one step's
        start (x0, v0) and end (x1, v1) states: the close-encounter list (appended
        after nclo0 entries) and the near-miss/collision list, with thit1 and
        nowflag."""
        arrs = [np.ascontiguousarray(a, np.float64) for a in (m, x0, v0, x1, v1, rce, rphys)]
        n = len(arrs[0])
        cm = self.cmax
        iclo = np.zeros(cm, np.int32); jclo = np.zeros(cm, np.int32)
        dclo = np.zeros(cm); tclo = np.zeros(cm); ixv = np.zeros((cm, 6)); jxv = np.zeros((cm, 6))
        ihit = np.zeros(cm, np.int32); jhit = np.zeros(cm, np.int32); chit = np.zeros(cm, np.int32)
        dhit = np.zeros(cm); thit = np.zeros(cm)
        nclo = C.c_int(nclo0); nhit = C.c_int(0); thit1 = C.c_double(0.0); nowflag = C.c_int(0)
        stat = np.zeros(n, np.int32)
        outfile = C.create_string_buffer(b"/dev/null".ljust(80))
        mem = C.create_string_buffer(b" " * (80 * self.NMESS))
        lmem = np.ones(self.NMESS, np.int32)
        self.L.mce_stat_(self._r(time), self._r(h), self._r(0.0), C.byref(C.c_int(n)), C.byref(C.c_int(nbig)),
                         *[_d(a) for a in arrs], C.byref(nclo), _i(iclo), _i(jclo), _d(dclo), _d(tclo),
                         _d(ixv), _d(jxv), C.byref(nhit), _i(ihit), _i(jhit), _i(chit), _d(dhit), _d(thit),
                         C.byref(thit1), C.byref(nowflag), _i(stat), outfile, mem, _i(lmem),
                         C.c_size_t(80), C.c_size_t(80))
        k, q = nclo.value, nhit.value
        return dict(iclo=iclo[nclo0:k].copy(), jclo=jclo[nclo0:k].copy(), dclo=dclo[nclo0:k].copy(),
                    tclo=tclo[nclo0:k].copy(), ixv=ixv[nclo0:k].copy(), jxv=jxv[nclo0:k].copy(),
                    ihit=ihit[:q].copy(), jhit=jhit[:q].copy(), chit=chit[:q].copy(), dhit=dhit[:q].copy(),
                    thit=thit[:q].copy(), thit1=thit1.value, nowflag=nowflag.value)

    def mdt_mvs_run(self, sysd, nsteps, time0=0.0, algor=1):
        """nsteps calls of the reference mdt_mvs on a whole system, plus the
        driver's mce_cent after each step (mal_hcon mercury6_2.for:811-841).
        Returns final (x, v) and event rows in Oracle.mvs_run's layout; the
        near-miss list of mce_stat stays inside mdt_mvs, so only close-encounter
        (kind 1) and central-impact (kind 8) rows are returned."""
        m = np.ascontiguousarray(sysd["m"], np.float64)
        n = len(m)
        nbig = sysd["nbig"]
        x = np.ascontiguousarray(sysd["x"], np.float64).copy()
        v = np.ascontiguousarray(sysd["v"], np.float64).copy()
        rce = np.ascontiguousarray(sysd["rce"], np.float64)
        rphys = np.ascontiguousarray(sysd["rphys"], np.float64)
        jc = np.ascontiguousarray(sysd.get("jcen", (0.0, 0.0, 0.0)), np.float64)
        h = sysd["h"]; rcen = sysd["rcen"]
        cm = self.cmax
        s = np.zeros((n, 3)); rcrit = np.zeros(n); stat = np.zeros(n, np.int32)
        # the cometary model (mdt_mvs keeps its accelerations in a saved array
        # that a later call with ngflag 0 does not clear: one model per process)
        ngf = np.ascontiguousarray(sysd.get("ngf", np.zeros((n, 5))), np.float64)
        ngflag = int(sysd.get("ngflag", 0))
        opt = np.array([0, 1, 1, 2, 0, 1, 0, 0], np.int32)
        ids = C.create_string_buffer(b" " * (25 * n))
        outfile = C.create_string_buffer(b"/dev/null".ljust(80) * 3)
        mem = C.create_string_buffer(b" " * (80 * self.NMESS))
        lmem = np.ones(self.NMESS, np.int32)
        iclo = np.zeros(cm, np.int32); jclo = np.zeros(cm, np.int32)
        dclo = np.zeros(cm); tclo = np.zeros(cm)
        ixv = np.zeros((cm, 6)); jxv = np.zeros((cm, 6))
        nclo = C.c_int(0); colflag = C.c_int(0); dtflag = C.c_int(0)
        en = np.zeros(3); am = np.zeros(3)
        rows_i = []; rows_d = []
        time = time0
        for step in range(nsteps):
            xs = x.copy(); vs = v.copy()
            self.L.mdt_mvs_(self._r(time), self._r(time0), self._r(h), self._r(1e-12),
                            self._r(100.0), _d(en), _d(am), _d(jc), self._r(rcen),
                            C.byref(C.c_int(n)), C.byref(C.c_int(nbig)), _d(m), _d(x), _d(v),
                            _d(s), _d(rphys), _d(rcrit), _d(rce), _i(stat), ids,
                            _d(ngf), C.byref(C.c_int(algor)), _i(opt), C.byref(dtflag),
                            C.byref(C.c_int(ngflag)), C.byref(C.c_int(-1)), C.byref(colflag),
                            C.byref(nclo), _i(iclo), _i(jclo), _d(dclo), _d(tclo), _d(ixv), _d(jxv),
                            outfile, mem, _i(lmem),
                            C.c_size_t(25), C.c_size_t(80), C.c_size_t(80))
            time = time + h
            for k in range(nclo.value):
                rows_i.append((step, 1, int(iclo[k]), int(jclo[k])))
                rows_d.append((float(dclo[k]), float(tclo[k])))
            for (j, t, d) in self.mce_cent(time, h, rcen, m, xs, vs, x, v):
                rows_i.append((step, 8, 1, j)); rows_d.append((d, t))
        return x, v, np.array(rows_i, np.int32).reshape(-1, 4), np.array(rows_d).reshape(-1, 2)
