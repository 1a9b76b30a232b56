"""The drop-in: the unchanged MERCURY6 driver linked against libmercury_hip
through mercuda_amd/fortran/mercury_gpu.f90 (built by `make -C oracle dropin`).

CPU part: with `execution backend = cpu` the drop-in binary is the reference
CPU path (identical dumps); requesting the device without one is a hard
failure, never a silent fallback (reference tests/test_edges.py:59-67).
GPU part: with the device selected, the dumps, encounter file and compressed
output agree with the reference CPU run on identical inputs."""
import numpy as np
import pytest

import mercury_run as R
from helpers import REL_TOL
from mercuda_amd import systems

needs_dropin = pytest.mark.skipif(not (R.HIP_BIN.exists() and R.CPU_BIN.exists()),
                                  reason="drop-in or reference driver not built (make -C oracle ref dropin)")


def _pair(tmp_path, sysd, nsteps, backend_b, **kw):
    a = R.write_run(tmp_path / "cpu", sysd, nsteps=nsteps, backend="cpu", **kw)
    b = R.write_run(tmp_path / "dev", sysd, nsteps=nsteps, backend=backend_b, **kw)
    ra = R.run(a, R.CPU_BIN)
    assert ra.returncode == 0, ra.stdout + ra.stderr
    rb = R.run(b, R.HIP_BIN)
    return a, b, rb


def _compare_dumps(a, b):
    worst = 0.0
    for f in ("big.dmp", "small.dmp"):
        da, db = R.read_dump(a, f), R.read_dump(b, f)
        assert da.keys() == db.keys()
        for k in da:
            for u, w in zip(da[k], db[k]):
                worst = max(worst, np.linalg.norm(u - w) / max(np.linalg.norm(u), 1e-300))
    return worst


@needs_dropin
def test_dropin_cpu_backend_is_the_reference(tmp_path):
    s = systems.config_small(seed=6, nparticles=200)
    a, b, rb = _pair(tmp_path, s, 200, "cpu")
    assert rb.returncode == 0, rb.stdout + rb.stderr
    assert _compare_dumps(a, b) == 0.0
    assert (a / "ce.out").read_bytes() == (b / "ce.out").read_bytes()
    assert any("Execution backend: CPU" in ln for ln in R.info_lines(b))


@needs_dropin
def test_dropin_device_request_without_device_fails(tmp_path):
    from mercuda_amd.hip import MercuryHip
    if MercuryHip().available():
        pytest.skip("a HIP device is present")
    s = systems.config_small(seed=6, nparticles=20)
    for spelling in ("hip", "cuda"):
        a, b, rb = _pair(tmp_path / spelling, s, 20, spelling)
        assert rb.returncode != 0
        assert "HIP was requested but no HIP device" in rb.stdout + rb.stderr


@needs_dropin
@pytest.mark.gpu
@pytest.mark.parametrize("algorithm", ["MVS", "TEST"])
def test_dropin_on_device_matches_reference(tmp_path, algorithm):
    s = systems.config_small(seed=1, nparticles=1000)
    print(f"case: drop-in {algorithm}, 8 planets + 1000 particles, 456 steps", flush=True)
    a, b, rb = _pair(tmp_path, s, 456, "hip", algorithm=algorithm, interval=8.0 * 57)
    assert rb.returncode == 0, rb.stdout + rb.stderr
    assert any("Execution backend: HIP" in ln for ln in R.info_lines(b)), R.info_lines(b)
    worst = _compare_dumps(a, b)
    print(f"max relative dump difference {worst:.3e}", flush=True)
    assert worst <= REL_TOL
    assert (a / "ce.out").read_bytes() == (b / "ce.out").read_bytes()
    assert (a / "xv.out").read_bytes() == (b / "xv.out").read_bytes()


@needs_dropin
@pytest.mark.gpu
@pytest.mark.parametrize("pn", [False, True])
def test_dropin_radau_on_device_matches_reference(tmp_path, pn):
    """RADAU through the unchanged adaptive driver (mal_hvar): clipped steps at
    every output time reset the RA15 history; dumps, encounters and output
    agree with the reference CPU run."""
    s = systems.config_encounters(seed=3, nparticles=600)
    s["rce"][1:s["nbig"]] *= 3.0
    print(f"case: drop-in RADAU pn={pn}, 8 planets + 600 particles, 600 d, outputs every 50 d", flush=True)
    a, b, rb = _pair(tmp_path, s, 300, "hip", algorithm="RADAU", interval=50.0, pn=pn)
    assert rb.returncode == 0, rb.stdout + rb.stderr
    assert any("Execution backend: HIP" in ln for ln in R.info_lines(b)), R.info_lines(b)
    worst = _compare_dumps(a, b)
    print(f"max relative dump difference {worst:.3e}", flush=True)
    assert worst <= REL_TOL
    assert (a / "ce.out").read_bytes() == (b / "ce.out").read_bytes()
    assert (a / "xv.out").read_bytes() == (b / "xv.out").read_bytes()


@needs_dropin
@pytest.mark.gpu
@pytest.mark.parametrize("algorithm", ["BS", "BS2"])
def test_dropin_bs_on_device_matches_reference(tmp_path, algorithm):
    """Bulirsch-Stoer (mdt_bs1 / mdt_bs2 with mfo_all) through the unchanged
    adaptive driver (mal_hvar), steps clipped at every output time; dumps,
    encounters and output agree with the reference CPU run."""
    s = systems.config_encounters(seed=3, nparticles=600)
    s["rce"][1:s["nbig"]] *= 3.0
    print(f"case: drop-in {algorithm}, 8 planets + 600 particles, 600 d, outputs every 50 d", flush=True)
    a, b, rb = _pair(tmp_path, s, 300, "hip", algorithm=algorithm, interval=50.0)
    assert rb.returncode == 0, rb.stdout + rb.stderr
    assert any("Execution backend: HIP" in ln for ln in R.info_lines(b)), R.info_lines(b)
    worst = _compare_dumps(a, b)
    print(f"max relative dump difference {worst:.3e}", flush=True)
    assert worst <= REL_TOL
    assert (a / "ce.out").read_bytes() == (b / "ce.out").read_bytes()
    assert (a / "xv.out").read_bytes() == (b / "xv.out").read_bytes()


@needs_dropin
@pytest.mark.gpu
def test_dropin_hybrid_on_device_matches_reference(tmp_path):
    """HYBRID through the unchanged driver: regular map and changeover scan on
    the device, the driver's mdt_hkce integrating encounter members through
    the library's BS2 context, central impacts from the device's events."""
    s = systems.config_encounters(seed=3, nparticles=600)
    print("case: drop-in HYBRID, 8 planets + 600 particles, 300 steps of 2 d", flush=True)
    a, b, rb = _pair(tmp_path, s, 300, "hip", algorithm="HYBRID", interval=100.0)
    assert rb.returncode == 0, rb.stdout + rb.stderr
    assert any("Execution backend: HIP" in ln for ln in R.info_lines(b)), R.info_lines(b)
    worst = _compare_dumps(a, b)
    print(f"max relative dump difference {worst:.3e}", flush=True)
    assert worst <= REL_TOL
    assert (a / "ce.out").read_bytes() == (b / "ce.out").read_bytes()
    assert (a / "xv.out").read_bytes() == (b / "xv.out").read_bytes()
