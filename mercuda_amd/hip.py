"""ctypes front end of libmercury_hip (include/mercury_hip.h).

``MercuryHip`` mirrors the calls the reference's Fortran binding makes over
the backend ABI (mercury_gpu.f90: gpu_select/push/advance/pull/find_events/
central) with the same argument meaning and the same failure behaviour: any
nonzero return raises ``MercuryHipError`` -- there is no CPU fallback.  Arrays
use MERCURY's layout, x[n,3] (C order == Fortran x(3,n)), body 0 central.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libmercury_hip.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)


class MercuryHipError(RuntimeError):
    pass


class MercuryEvent(C.Structure):
    _fields_ = [("i", C.c_int), ("j", C.c_int), ("kind", C.c_int), ("now", C.c_int),
                ("distance", C.c_double), ("time", C.c_double),
                ("xi", C.c_double * 6), ("xj", C.c_double * 6)]


class MercuryHipEvent(C.Structure):
    _fields_ = [("e", MercuryEvent), ("step", C.c_int), ("system", C.c_int)]


# Every function the header declares, with its ctypes signature.
SIGNATURES = {
    "mercury_hip_available": (C.c_int, []),
    "mercury_hip_configure": (C.c_int, [C.c_int]),
    "mercury_hip_reset": (None, [C.c_int]),
    "mercury_hip_limits": (C.c_int, [_dp]),
    "mercury_hip_export": (C.c_int, [C.c_double, C.c_int, _dp, _dp]),
    "mercury_hip_upload": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "mercury_hip_step": (C.c_int, [C.c_double, _dp, _dp, C.c_double, _i64p, _i64p]),
    "mercury_hip_download": (C.c_int, [_dp, _dp, C.c_int]),
    "mercury_hip_events": (C.c_int, [C.c_double, C.c_double, C.c_double,
                                     C.POINTER(C.POINTER(MercuryEvent)), _ip]),
    "mercury_hip_force": (C.c_int, [_dp]),
    "mercury_hip_force_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                            _dp, C.c_void_p, C.c_longlong]),
    "mercury_hip_hybrid_begin": (C.c_int, [C.c_double, _dp, C.c_int, C.c_int, _ip, _ip, _ip, _ip,
                                           _dp, _dp, _i64p]),
    "mercury_hip_hybrid_finish": (C.c_int, [C.c_double, _dp, _dp, _dp, _i64p]),
    "mercury_hip_encounter_enter": (C.c_int, [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                                              C.c_int, _ip, _ip]),
    "mercury_hip_encounter_update": (C.c_int, [_dp, _dp, _dp]),
    "mercury_hip_encounter_exit": (None, []),
    "mercury_hip_encounter_advance": (C.c_int, [C.c_double, _dp, C.c_double, _i64p, _i64p, _i64p]),
    "mercury_hip_free": (None, []),
    "mercury_hip_advance": (C.c_int, [C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, _i64p]),
    "mercury_hip_advance_events": (C.c_int, [C.POINTER(C.POINTER(MercuryHipEvent)), _ip]),
    "mercury_hip_set_tuning": (C.c_int, [C.c_int, C.c_int, C.c_int]),
    "mercury_hip_timing": (C.c_int, [C.c_int]),
    "mercury_hip_timing_read": (C.c_int, [_dp, _ip, _dp, _ip, _dp, _ip]),
    "mercury_hip_parked": (C.c_int64, []),
    "mercury_hip_work_read": (C.c_int, [_i64p, _i64p]),
    "mercury_hip_check_arith": (C.c_int, [C.c_int64, C.c_uint64, _i64p]),
    "mercury_hip_ensemble_upload": (C.c_int, [C.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "mercury_hip_ensemble_advance": (C.c_int, [C.c_double, C.c_int, C.c_int, _i64p]),
    "mercury_hip_ensemble_download": (C.c_int, [_dp, _dp]),
    "mercury_hip_set_graph": (C.c_int, [C.c_int]),
    "mercury_hip_graph_stats": (C.c_int, [_ip, _ip]),
    "mercury_hip_autotune": (C.c_int, [C.c_double, C.c_int, _ip, _ip, _ip, _dp]),
    "mercury_hip_adjoint_forward": (C.c_int, [C.c_int, C.c_double, C.c_int, _dp, _dp]),
    "mercury_hip_adjoint_backward": (C.c_int, [_dp, _dp, _dp, _dp, _dp]),
    "mercury_hip_tangent": (C.c_int, [_dp, _dp, _dp, _dp, _dp]),
}


def load_library(path: Path = LIB_PATH) -> C.CDLL:
    """Load libmercury_hip.so and bind every declared symbol (raises if any is missing)."""
    if not Path(path).exists():
        raise MercuryHipError(f"{path} is not built; run python -m mercuda_amd.build")
    lib = C.CDLL(str(path))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def event_tuple(e: MercuryEvent):
    return (e.i, e.j, e.kind, e.now, e.distance, e.time, tuple(e.xi), tuple(e.xj))


class MercuryHip:
    """The HIP backend for one process (the library holds one global context)."""

    def __init__(self, path: Path = LIB_PATH):
        self.L = load_library(path)
        self.n = 0
        self.nbig = 0

    def _ok(self, rc, what):
        if rc != 0:
            raise MercuryHipError(f"{what} failed")

    def available(self) -> bool:
        return bool(self.L.mercury_hip_available())

    def configure(self, algorithm: int):
        self._ok(self.L.mercury_hip_configure(algorithm), f"configure({algorithm})")

    def reset(self, flag: int = 0):
        self.L.mercury_hip_reset(flag)

    def upload(self, m, x, v, nbig, *, pn=0, ngflag=0, ngf=None, jcen=(0.0, 0.0, 0.0), rce=None, rphys=None):
        m = _f64(m); n = len(m)
        x = _f64(x, (n, 3)); v = _f64(v, (n, 3))
        ngf = _f64(np.zeros((n, 5)) if ngf is None else ngf, (n, 5))
        jc = _f64(jcen, (3,))
        rce = _f64(np.zeros(n) if rce is None else rce, (n,))
        rphys = _f64(np.zeros(n) if rphys is None else rphys, (n,))
        self._ok(self.L.mercury_hip_upload(n, nbig, pn, ngflag, _d(m), _d(x), _d(v), _d(ngf), _d(jc),
                                           _d(rce), _d(rphys)), "upload")
        self.n, self.nbig = n, nbig

    def upload_system(self, sysd, **kw):
        """Upload a system dict (mercuda_amd.systems), its non-gravitational
        model (sysd["ngflag"], sysd["ngf"][n, 5]) included when present."""
        if sysd.get("ngflag", 0):
            kw.setdefault("ngflag", int(sysd["ngflag"]))
            kw.setdefault("ngf", sysd["ngf"])
        self.upload(sysd["m"], sysd["x"], sysd["v"], sysd["nbig"], jcen=sysd.get("jcen", (0, 0, 0)),
                    rce=sysd["rce"], rphys=sysd["rphys"], **kw)

    def step(self, time, h, tol=1e-12):
        hh = C.c_double(h); hdid = C.c_double(0.0)
        nf = C.c_int64(0); nr = C.c_int64(0)
        self._ok(self.L.mercury_hip_step(time, C.byref(hh), C.byref(hdid), tol, C.byref(nf), C.byref(nr)),
                 "step")
        return hh.value, hdid.value, nf.value, nr.value

    def download(self, previous=False):
        x = np.zeros((self.n, 3)); v = np.zeros((self.n, 3))
        self._ok(self.L.mercury_hip_download(_d(x), _d(v), int(previous)), "download")
        return x, v

    def export(self, h, physical=True):
        x = np.zeros((self.n, 3)); v = np.zeros((self.n, 3))
        self._ok(self.L.mercury_hip_export(h, int(physical), _d(x), _d(v)), "export")
        return x, v

    def events(self, time, h, rcen):
        ptr = C.POINTER(MercuryEvent)(); cnt = C.c_int(0)
        self._ok(self.L.mercury_hip_events(time, h, rcen, C.byref(ptr), C.byref(cnt)), "events")
        return [event_tuple(ptr[k]) for k in range(cnt.value)]

    def limits(self, rce):
        rce = _f64(rce, (self.n,))
        self._ok(self.L.mercury_hip_limits(_d(rce)), "limits")

    def force(self):
        a = np.zeros((self.n, 3))
        self._ok(self.L.mercury_hip_force(_d(a)), "force")
        return a

    def force_noalloc(self, n, nbig, m_dev, x_dev, v_dev, a_dev, ld, *, pn=0, jcen=(0.0, 0.0, 0.0)):
        """mercury_hip_force_noalloc on device buffers (addresses as ints):
        m[n], x/v/a component-major with leading dimension ld."""
        jc = _f64(jcen, (3,))
        self._ok(self.L.mercury_hip_force_noalloc(n, nbig, pn, C.c_void_p(m_dev), C.c_void_p(x_dev),
                                                  C.c_void_p(v_dev), _d(jc), C.c_void_p(a_dev), ld),
                 "force_noalloc")

    def advance(self, time, h, nsteps, rcen=0.005, screen=True):
        nf = C.c_int64(0)
        self._ok(self.L.mercury_hip_advance(time, h, nsteps, rcen, int(screen), C.byref(nf)), "advance")
        return nf.value

    def advance_events(self):
        ptr = C.POINTER(MercuryHipEvent)(); cnt = C.c_int(0)
        self._ok(self.L.mercury_hip_advance_events(C.byref(ptr), C.byref(cnt)), "advance_events")
        return [(ptr[k].system, ptr[k].step) + event_tuple(ptr[k].e) for k in range(cnt.value)]

    def set_tuning(self, steps_per_launch=0, block_size=0, particles_per_lane=0):
        self._ok(self.L.mercury_hip_set_tuning(steps_per_launch, block_size, particles_per_lane), "set_tuning")

    def timing(self, enable=True, fast_only=False):
        """Kernel timing on (every kernel kind, or the fast particle kernel
        only) or off."""
        self._ok(self.L.mercury_hip_timing(0 if not enable else (2 if fast_only else 1)), "timing")

    def timing_read(self):
        pm = C.c_double(); pl = C.c_int(); qm = C.c_double(); ql = C.c_int(); gm = C.c_double(); gl = C.c_int()
        self._ok(self.L.mercury_hip_timing_read(C.byref(pm), C.byref(pl), C.byref(qm), C.byref(ql),
                                                C.byref(gm), C.byref(gl)), "timing_read")
        return dict(particle_ms=pm.value, particle_launches=pl.value, planet_ms=qm.value,
                    planet_launches=ql.value, general_ms=gm.value, general_launches=gl.value)

    def work_read(self):
        f = C.c_int64(); g = C.c_int64()
        self._ok(self.L.mercury_hip_work_read(C.byref(f), C.byref(g)), "work_read")
        return f.value, g.value

    def check_arith(self, n: int, seed: int = 1) -> dict:
        """Bit mismatches of the fast kernel's divide, reciprocal, square root
        and constant divide against the IEEE expansions over n random draws."""
        out = (C.c_int64 * 4)()
        self._ok(self.L.mercury_hip_check_arith(int(n), int(seed) & (2**64 - 1), out), "check_arith")
        return dict(div=out[0], rcp=out[1], sqrt=out[2], div_const=out[3])

    def parked(self) -> int:
        return int(self.L.mercury_hip_parked())

    # ---- batched ensembles
    def ensemble_upload(self, systems_list):
        """Upload a list of system dicts (mercuda_amd.systems) as one ensemble."""
        ns = len(systems_list)
        n = np.array([len(q["m"]) for q in systems_list], np.int32)
        nbig = np.array([q["nbig"] for q in systems_list], np.int32)
        cat = lambda key, shape: np.ascontiguousarray(np.concatenate([np.asarray(q[key], np.float64).reshape(shape)
                                                                      for q in systems_list]))
        m = cat("m", (-1,)); x = cat("x", (-1, 3)); v = cat("v", (-1, 3))
        rce = cat("rce", (-1,)); rphys = cat("rphys", (-1,))
        h = np.array([q["h"] for q in systems_list], np.float64)
        rcen = np.array([q["rcen"] for q in systems_list], np.float64)
        ip = lambda a: a.ctypes.data_as(_ip)
        self._ok(self.L.mercury_hip_ensemble_upload(ns, ip(n), ip(nbig), _d(m), _d(x), _d(v), _d(rce), _d(rphys),
                                                    _d(h), _d(rcen)), "ensemble_upload")
        self.n = int(n.sum()); self.nbig = 0
        self.ens_offsets = np.concatenate([[0], np.cumsum(n)])

    def ensemble_advance(self, time, nsteps, screen=True):
        nf = C.c_int64(0)
        self._ok(self.L.mercury_hip_ensemble_advance(time, nsteps, int(screen), C.byref(nf)), "ensemble_advance")
        return nf.value

    def ensemble_download(self):
        x = np.zeros((self.n, 3)); v = np.zeros((self.n, 3))
        self._ok(self.L.mercury_hip_ensemble_download(_d(x), _d(v)), "ensemble_download")
        o = self.ens_offsets
        return [(x[o[k]:o[k + 1]], v[o[k]:o[k + 1]]) for k in range(len(o) - 1)]

    def set_graph(self, enable=True):
        self._ok(self.L.mercury_hip_set_graph(int(enable)), "set_graph")

    def graph_stats(self):
        a = C.c_int(0); b = C.c_int(0)
        self._ok(self.L.mercury_hip_graph_stats(C.byref(a), C.byref(b)), "graph_stats")
        return {"captures": a.value, "replays": b.value}

    def autotune(self, h=0.0, probe_steps=32):
        spl = C.c_int(0); blk = C.c_int(0); ppl = C.c_int(0); rate = C.c_double(0.0)
        self._ok(self.L.mercury_hip_autotune(float(h), int(probe_steps), C.byref(spl), C.byref(blk), C.byref(ppl),
                                             C.byref(rate)), "autotune")
        return {"steps_per_launch": spl.value, "block": blk.value, "particles_per_lane": ppl.value,
                "particle_steps_per_s": rate.value}

    # ---- HYBRID (algorithm 10) and the compact encounter context ----
    def hybrid_begin(self, h, rcrit, flag, cap, x, v):
        """First half of a HYBRID step; x, v (n x 3) receive the encounter
        members' pre-drift states.  Returns ce, pairs (i, j) 1-based, forces."""
        rcrit = _f64(rcrit, (self.n,))
        ce = np.zeros(self.n, np.int32)
        pi = np.zeros(max(cap, 1), np.int32); pj = np.zeros(max(cap, 1), np.int32)
        nce = C.c_int(0); nf = C.c_int64(0)
        assert x.flags.c_contiguous and v.flags.c_contiguous and x.dtype == np.float64
        self._ok(self.L.mercury_hip_hybrid_begin(float(h), _d(rcrit), int(flag), int(cap), ce.ctypes.data_as(_ip),
                                                 C.byref(nce), pi.ctypes.data_as(_ip), pj.ctypes.data_as(_ip),
                                                 _d(x), _d(v), C.byref(nf)), "hybrid_begin")
        k = nce.value
        return ce, list(zip(pi[:k].tolist(), pj[:k].tolist())), nf.value

    def hybrid_finish(self, h, m, x, v):
        m = _f64(m, (self.n,)); x = _f64(x, (self.n, 3)); v = _f64(v, (self.n, 3))
        nf = C.c_int64(0)
        self._ok(self.L.mercury_hip_hybrid_finish(float(h), _d(m), _d(x), _d(v), C.byref(nf)), "hybrid_finish")
        return nf.value

    def encounter_enter(self, m, x, v, rcrit, rce, rphys, nbig, pairs):
        n = len(m)
        arr = [_f64(a) for a in (m, x, v, rcrit, rce, rphys)]
        pi = np.ascontiguousarray([p[0] for p in pairs] or [0], np.int32)
        pj = np.ascontiguousarray([p[1] for p in pairs] or [0], np.int32)
        self._ok(self.L.mercury_hip_encounter_enter(n, int(nbig), *[_d(a) for a in arr], len(pairs),
                                                    pi.ctypes.data_as(_ip), pj.ctypes.data_as(_ip)),
                 "encounter_enter")
        self.enc_n = n

    def encounter_update(self, m, x, v):
        self._ok(self.L.mercury_hip_encounter_update(_d(_f64(m)), _d(_f64(x)), _d(_f64(v))), "encounter_update")

    def encounter_download(self):
        x = np.zeros((self.enc_n, 3)); v = np.zeros((self.enc_n, 3))
        self._ok(self.L.mercury_hip_download(_d(x), _d(v), 0), "download (encounter)")
        return x, v

    def encounter_exit(self):
        self.L.mercury_hip_encounter_exit()

    def encounter_advance(self, h0, hrec, tol):
        """mdt_hkce's whole BS2 loop over h0 with the step control on the device.
        Returns (hrec, substeps, forces, rejected)."""
        hr = C.c_double(hrec)
        ns = C.c_int64(); nf = C.c_int64(); nr = C.c_int64()
        self._ok(self.L.mercury_hip_encounter_advance(float(h0), C.byref(hr), float(tol), C.byref(ns), C.byref(nf),
                                                       C.byref(nr)), "encounter_advance")
        return hr.value, ns.value, nf.value, nr.value

    # ---- differentiation (tape of the last adjoint_forward) ----
    def adjoint_forward(self, nsteps, h, segment=0):
        """Final (x, v) after nsteps MVS steps of the uploaded state; keeps the tape."""
        x = np.zeros((self.n, 3)); v = np.zeros((self.n, 3))
        self._ok(self.L.mercury_hip_adjoint_forward(int(nsteps), float(h), int(segment), _d(x), _d(v)),
                 "adjoint_forward")
        return x, v

    def adjoint_backward(self, gx, gv):
        """Gradient of <gx, x_T> + <gv, v_T> w.r.t. initial x, v and masses."""
        gx = _f64(gx, (self.n, 3)); gv = _f64(gv, (self.n, 3))
        gx0 = np.zeros((self.n, 3)); gv0 = np.zeros((self.n, 3)); gm = np.zeros(self.n)
        self._ok(self.L.mercury_hip_adjoint_backward(_d(gx), _d(gv), _d(gx0), _d(gv0), _d(gm)), "adjoint_backward")
        return gx0, gv0, gm

    def tangent(self, dx0, dv0, dm):
        """Derivative of the final (x, v) along (dx0, dv0, dm)."""
        dx0 = _f64(dx0, (self.n, 3)); dv0 = _f64(dv0, (self.n, 3)); dm = _f64(dm, (self.n,))
        dx = np.zeros((self.n, 3)); dv = np.zeros((self.n, 3))
        self._ok(self.L.mercury_hip_tangent(_d(dx0), _d(dv0), _d(dm), _d(dx), _d(dv)), "tangent")
        return dx, dv

    def free(self):
        self.L.mercury_hip_free()
