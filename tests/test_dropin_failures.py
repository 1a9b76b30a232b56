"""The drop-in's failure contract and restarts with a backend switch, through
the unchanged driver (reference tests/test_edges.py:34-67,100-123 and
tests/test_events.py:110-113), with real failure conditions rather than
injected return codes:

- `auto` with an upload the device path refuses (a Small body with mass)
  falls back to the CPU and says so in info.out; `hip` fails;
- a device step that fails (a particle on a planet: non-finite force) ends the
  run with an error under both `hip` and `auto`, never with a fallback;
- a non-finite input state never runs on the device;
- a run started on one backend and continued from its dumps on the other
  ends where a CPU-only run ends."""
import re

import numpy as np
import pytest

import mercury_run as R

needs_dropin = pytest.mark.skipif(not (R.HIP_BIN.exists() and R.CPU_BIN.exists()),
                                  reason="drop-in or reference driver not built (make -C oracle ref dropin)")
pytestmark = [needs_dropin, pytest.mark.gpu]


def many(nparticles=4096, massive=False):
    """Enough particles for `auto` to choose the device (n - nbig >= 4096)."""
    small = [R.body(f"P{j}", a=2.0 + j * 1e-3, phase=0.001 * j) for j in range(nparticles)]
    if massive:
        small[7]["mass"] = 1e-12
    return small


def fallback_lines(path):
    if not (path / "info.out").exists():
        return []
    return [ln for ln in R.info_lines(path) if "fallback" in ln]


@pytest.mark.parametrize("algorithm", ["BS", "RADAU"])
def test_auto_upload_failure_falls_back_logged(tmp_path, algorithm):
    small = many(massive=True)
    case = dict(algorithm=algorithm, small=small, big=[R.body("PLANET", mass=1e-6, a=5.0)], stop=2.0, step=0.5,
                interval=2.0)
    print(f"case: {algorithm} auto fallback on an upload the device refuses", flush=True)
    ref = R.write_case(tmp_path / "cpu", backend="cpu", **case)
    assert R.run(ref, R.CPU_BIN).returncode == 0
    auto = R.write_case(tmp_path / "auto", backend="auto", **case)
    r = R.run(auto, R.HIP_BIN)
    assert r.returncode == 0, r.stdout + r.stderr
    assert any("CPU (HIP initialization failed; auto fallback)" in ln for ln in fallback_lines(auto))
    assert R.read_dump_full(auto) .keys() == R.read_dump_full(ref).keys()
    assert (auto / "small.dmp").read_text() == (ref / "small.dmp").read_text()
    hip = R.write_case(tmp_path / "hip", backend="hip", **case)
    r = R.run(hip, R.HIP_BIN)
    assert r.returncode != 0
    assert "HIP state upload failed" in r.stdout + r.stderr
    assert not fallback_lines(hip)


@pytest.mark.parametrize("algorithm", ["TEST", "BS", "RADAU"])
@pytest.mark.parametrize("backend", ["hip", "auto"])
def test_device_step_failure_never_falls_back(tmp_path, algorithm, backend):
    """TEST is the MVS step without the symplectic corrector's input transform
    (which under MVS moves the particle off its 1e-160 AU offset)."""
    planet = R.body("PLANET", mass=1e-6, a=5.0, e=0.0)
    small = many()
    # 1e-160 AU from the planet (a distinct position, so the driver's input
    # check passes): the squared separation is subnormal and the force is not
    # finite on the first step
    small[3]["x"] = [planet["x"][0], planet["x"][1], 1e-160]
    small[3]["v"] = [0.0, 0.0, 1e-3]
    p = R.write_case(tmp_path / backend, backend=backend, algorithm=algorithm, small=small, big=[planet],
                     stop=2.0, step=0.5, interval=2.0)
    print(f"case: {algorithm} step failure under {backend}", flush=True)
    r = R.run(p, R.HIP_BIN)
    assert r.returncode != 0
    assert "HIP integration step failed" in r.stdout + r.stderr, r.stdout + r.stderr
    assert any("Execution backend: HIP" in ln for ln in R.info_lines(p))
    assert not fallback_lines(p)


def test_nonfinite_input_never_runs_on_the_device(tmp_path):
    small = [R.body("OK", a=1.5), R.body("BAD", a=2.0)]
    small[1]["v"][1] = float("nan")
    p = R.write_case(tmp_path / "nan", backend="hip", algorithm="MVS", small=small, stop=1.0, step=0.5,
                     interval=1.0)
    r = R.run(p, R.HIP_BIN)
    print(f"case: non-finite input state, exit {r.returncode}", flush=True)
    assert r.returncode != 0
    assert "Nonfinite" in r.stdout + r.stderr
    assert not fallback_lines(p)


@pytest.mark.parametrize("algorithm", ["BS", "BS2", "RADAU", "MVS", "HYBRID"])
@pytest.mark.parametrize("first,second", [("cpu", "hip"), ("hip", "cpu")])
def test_restart_switches_backend(tmp_path, algorithm, first, second):
    """Stop at 8 d, raise the stop time in param.dmp, switch the backend in
    param.in and continue from the dumps (reference tests/test_edges.py:
    107-123); the result agrees with a CPU-only run to 16 d."""
    small = [R.body(), R.body("Q", a=1.7, e=0.3, phase=1.0)]
    big = [R.body("PLANET", mass=1e-6, a=3.0, e=0.02)]
    case = dict(algorithm=algorithm, small=small, big=big, step=0.5, interval=2.0)
    p = R.write_case(tmp_path / "run", backend=first, stop=8.0, **case)
    assert R.run(p, R.HIP_BIN).returncode == 0
    dmp = (p / "param.dmp").read_text()
    (p / "param.dmp").write_text(re.sub(r"(stop time.*?=)\s*[^\n]+", r"\g<1> 16", dmp, flags=re.I))
    (p / "param.in").write_text((p / "param.in").read_text().replace(f"backend = {first}", f"backend = {second}"))
    r = R.run(p, R.HIP_BIN)
    assert r.returncode == 0, r.stdout + r.stderr
    ref = R.write_case(tmp_path / "ref", backend="cpu", stop=16.0, **case)
    assert R.run(ref, R.CPU_BIN).returncode == 0
    a, b = R.read_dump_full(p), R.read_dump_full(ref)
    assert a.keys() == b.keys()
    worst = max(np.linalg.norm(a[k][q] - b[k][q]) / np.linalg.norm(b[k][q]) for k in a for q in ("x", "v"))
    print(f"case: restart {algorithm} {first}->{second}: max relative difference {worst:.3e}", flush=True)
    assert worst <= 1e-10
    backends = [ln for ln in R.info_lines(p) if "Execution backend:" in ln]
    assert ("HIP" in backends[0]) == (first == "hip") and ("HIP" in backends[-1]) == (second == "hip")
