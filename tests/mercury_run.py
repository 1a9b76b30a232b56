"""Run directories for the unchanged MERCURY6 driver: input writers and dump
readers (the file formats are the driver's: param.in lines `label = value`,
big.in/small.in in Cartesian style, *.dmp read back by name).  The layout
follows what the reference's own test helper writes (tests/cases.py:20-51)."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import numpy as np

from mercuda_amd import systems

REF_DIR = Path(__file__).resolve().parents[1] / "oracle" / "_ref"
CPU_BIN = REF_DIR / "mercury6"
HIP_BIN = REF_DIR / "mercury6_hip"

PARAM_LABELS = [
    "algorithm", "start time (days)", "stop time (days)", "output interval (days)", "timestep (days)",
    "accuracy parameter", "stop integration after a close encounter", "allow collisions to occur",
    "include collisional fragmentation", "express time in days or years",
    "express time relative to integration start time", "output precision", "< not used at present >",
    "include relativity in integration", "include user-defined force", "ejection distance (AU)",
    "radius of central body (AU)", "central mass (solar)", "central J2", "central J4", "central J6",
    "< not used at present >", "< not used at present >", "Hybrid integrator changeover (Hill radii)",
    "number of timesteps between data dumps", "number of timesteps between periodic effects",
]


def write_run(path, sysd, *, algorithm="MVS", stop=None, nsteps=100, interval=None, backend="cpu",
              rceh=1.0, tol=1e-12, rmax=100.0, dump_every=5000, periodic=100, changeover=3.0, pn=False):
    """A complete run directory for one system dict (masses in G*M units)."""
    path = Path(path)
    path.mkdir(parents=True, exist_ok=True)
    for name in ("message.in", "files.in"):
        shutil.copyfile(REF_DIR / name, path / name)
    h = sysd["h"]
    stop = nsteps * h if stop is None else stop
    interval = stop if interval is None else interval
    m0 = sysd["m"][0] / systems.K2
    values = [algorithm, 0.0, stop, interval, h, tol, "no", "no", "no", "days", "no", "high", "unused",
              "yes" if pn else "no", "no", rmax, sysd["rcen"], m0, 0.0, 0.0, 0.0, "unused", "unused",
              changeover, dump_every, periodic]
    fmt = lambda v: f"{v:.17e}" if isinstance(v, float) else str(v)
    (path / "param.in").write_text(")O+_06 Integration parameters\n"
                                   + "\n".join(f" {lab} = {fmt(v)}" for lab, v in zip(PARAM_LABELS, values))
                                   + f"\n execution backend = {backend}\n")
    nb = sysd["nbig"]
    for fname, rng_, big in (("big.in", range(1, nb), True), ("small.in", range(nb, len(sysd["m"])), False)):
        lines = [")O+_06 Initial data", " style = Cartesian"]
        if big:
            lines.append(f" epoch = {0.0:.17e}")
        for j in rng_:
            name = ("B" if big else "P") + str(j)
            mass = sysd["m"][j] / systems.K2
            head = f"{name} m={mass:.17e}"
            if big:
                head += f" r={rceh:.17e}"
            if "ngf" in sysd:   # non-gravitational parameters as the driver reads them (a1 a2 a3 b yar)
                for key, q in (("a1", 0), ("a2", 1), ("a3", 2), ("b", 3), ("yar", 4)):
                    if sysd["ngf"][j, q] != 0.0:
                        head += f" {key}={sysd['ngf'][j, q]:.17e}"
            lines.append(head)
            lines.append(" ".join(f"{q:.17e}" for q in list(sysd["x"][j]) + list(sysd["v"][j]) + [0.0] * 3))
        (path / fname).write_text("\n".join(lines) + "\n")
    return path


def _big_stack():
    """The driver keeps body-count-sized arrays on the stack: at 1e5 bodies
    they exceed the usual 8 MiB limit."""
    import resource
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    want = resource.RLIM_INFINITY if hard == resource.RLIM_INFINITY else hard
    resource.setrlimit(resource.RLIMIT_STACK, (want, hard))


def run(path, binary, timeout=600):
    r = subprocess.run([str(binary)], cwd=path, capture_output=True, text=True, timeout=timeout,
                       preexec_fn=_big_stack)
    return r


def read_dump(path, filename):
    """name -> (x[3], v[3]) from big.dmp / small.dmp."""
    lines = [ln.strip() for ln in (Path(path) / filename).read_text().splitlines()
             if ln.strip() and not ln.startswith(")")]
    lines = lines[2 if filename.startswith("big") else 1:]
    out = {}
    for k in range(0, len(lines), 4):
        name = lines[k].split()[0]
        vals = [float(t.replace("D", "E")) for ln in lines[k + 1:k + 3] for t in ln.split()[:3]]
        out[name] = (np.array(vals[:3]), np.array(vals[3:6]))
    return out


def info_lines(path):
    return (Path(path) / "info.out").read_text().splitlines()


def body(name="PARTICLE", mass=0.0, a=1.0, e=0.1, phase=0.0, **params):
    """A body at the pericentre of a coplanar orbit about one solar mass (the
    reference's test bodies, tests/cases.py:12-17): mass in solar masses, extra
    big.in/small.in parameters (d=, r=, a1=, a2=, a3=, b=, yar=) as keywords."""
    r = a * (1 - e)
    speed = np.sqrt(systems.K2 * (1 + e) / r)
    c, s = np.cos(phase), np.sin(phase)
    return dict(name=name, mass=mass, x=[r * c, r * s, 0.0], v=[-speed * s, speed * c, 0.0], params=params)


def write_case(path, small=None, big=None, *, algorithm="BS", stop=100.0, step=1.0, interval=7.3, tol=1e-12,
               backend="cpu", collisions=False, pn=False, rcen=0.005, rmax=100.0, periodic=100,
               dump_every=5000, changeover=3.0, jcen=(0.0, 0.0, 0.0), start=0.0, epoch=0.0):
    """A run directory from named bodies (body()), with the driver's parameters
    as the reference's own test cases set them (tests/cases.py:20-48); start
    and epoch are the integration start time and the Big bodies' epoch."""
    path = Path(path)
    path.mkdir(parents=True, exist_ok=True)
    for name in ("message.in", "files.in"):
        shutil.copyfile(REF_DIR / name, path / name)
    values = [algorithm, start, stop, interval, step, tol, "no", "yes" if collisions else "no", "no", "days", "no",
              "high", "unused", "yes" if pn else "no", "no", rmax, rcen, 1.0, *jcen, "unused", "unused",
              changeover, dump_every, periodic]
    fmt = lambda v: f"{v:.17e}" if isinstance(v, float) else str(v)
    (path / "param.in").write_text(")O+_06 Integration parameters\n"
                                   + "\n".join(f" {lab} = {fmt(v)}" for lab, v in zip(PARAM_LABELS, values))
                                   + f"\n execution backend = {backend}\n")
    for fname, objs, is_big in (("big.in", big or [], True), ("small.in", small if small is not None else [body()],
                                                             False)):
        lines = [")O+_06 Initial data", " style = Cartesian"]
        if is_big:
            lines.append(f" epoch = {epoch:.17e}")
        for o in objs:
            params = {"m": o["mass"], **o.get("params", {})}
            lines.append(o["name"] + " " + " ".join(f"{k}={v:.17e}" for k, v in params.items()))
            lines.append(" ".join(f"{q:.17e}" for q in list(o["x"]) + list(o["v"]) + [0.0] * 3))
        (path / fname).write_text("\n".join(lines) + "\n")
    return path


def read_dump_full(path, filename="small.dmp"):
    """name -> dict(mass, x[3], v[3], params) from a dump file."""
    lines = [ln.strip() for ln in (Path(path) / filename).read_text().splitlines()
             if ln.strip() and not ln.startswith(")")]
    lines = lines[2 if filename.startswith("big") else 1:]
    out = {}
    for k in range(0, len(lines), 4):
        tok = lines[k].replace("=", " = ").split()
        params = {tok[i - 1]: float(tok[i + 1].replace("D", "E")) for i, t in enumerate(tok) if t == "="}
        vals = [float(t.replace("D", "E")) for ln in lines[k + 1:k + 3] for t in ln.split()[:3]]
        out[tok[0]] = dict(mass=params.pop("m", 0.0), x=np.array(vals[:3]), v=np.array(vals[3:6]), params=params)
    return out
