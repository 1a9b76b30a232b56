"""The drop-in driver with non-gravitational forces on the device path: the
unchanged MERCURY6 driver reads a1/a2/a3/b/yar from big.in and small.in
(mercury6_2.for:6279-6290), sets ngflag (:6360-6370) and hands the parameters
to mercury_hip_upload; every run is compared with the reference CPU driver on
identical inputs, as the reference's own CPU-vs-GPU cases
(tests/test_gpu_algorithms.py:89-100).

RADAU with 1PN, radiation pressure, Poynting-Robertson drag and Yarkovsky
(the configuration of the reference's published RADAU+1PN+PR+Yarkovsky
timing) uses no pow, so its dumps, encounter file and compressed output are
byte-identical.  Cometary A1-A3 scale by Marsden's g(r), computed correctly
rounded on the device and with libm's pow on the CPU: those runs meet the
1e-11 bar on the dumps and compare the encounter and output files as well."""
import numpy as np
import pytest

import mercury_run as R
from helpers import REL_TOL
from mercuda_amd import systems

needs_dropin = pytest.mark.skipif(not (R.HIP_BIN.exists() and R.CPU_BIN.exists()),
                                  reason="drop-in or reference driver not built (make -C oracle ref dropin)")
pytestmark = [needs_dropin, pytest.mark.gpu]


def run_pair(tmp_path, sysd, nsteps, **kw):
    a = R.write_run(tmp_path / "cpu", sysd, nsteps=nsteps, backend="cpu", **kw)
    b = R.write_run(tmp_path / "dev", sysd, nsteps=nsteps, backend="hip", **kw)
    ra = R.run(a, R.CPU_BIN)
    assert ra.returncode == 0, ra.stdout + ra.stderr
    rb = R.run(b, R.HIP_BIN)
    assert rb.returncode == 0, rb.stdout + rb.stderr
    assert any("Execution backend: HIP" in ln for ln in R.info_lines(b)), R.info_lines(b)
    return a, b


def worst_dump_difference(a, b):
    worst = 0.0
    for f in ("big.dmp", "small.dmp"):
        da, db = R.read_dump_full(a, f), R.read_dump_full(b, f)
        assert da.keys() == db.keys()
        for k in da:
            assert da[k]["params"] == db[k]["params"]
            for q in ("x", "v"):
                worst = max(worst, np.linalg.norm(da[k][q] - db[k][q]) / max(np.linalg.norm(da[k][q]), 1e-300))
    return worst


def encounter_system(seed, nparticles, kinds, big):
    s = systems.config_encounters(seed=seed, nparticles=nparticles)
    s["rce"][1:s["nbig"]] *= 3.0
    return systems.add_nongrav(s, seed=seed + 50, fraction=0.6, kinds=kinds, big=big)


@pytest.mark.parametrize("algorithm", ["RADAU", "BS"])
def test_dropin_pn_pr_yarkovsky_byte_identical(tmp_path, algorithm):
    s = encounter_system(3, 600, ("yar", "beta"), big=True)
    nb = int(np.sum(s["ngf"][:, 3] != 0)); ny = int(np.sum(s["ngf"][:, 4] != 0))
    print(f"case: drop-in {algorithm} 1PN + PR ({nb} bodies) + Yarkovsky ({ny} bodies), 8 planets + 600 "
          "particles, 600 d, outputs every 50 d", flush=True)
    a, b = run_pair(tmp_path, s, 300, algorithm=algorithm, interval=50.0, pn=True)
    assert any("1PN" in ln for ln in R.info_lines(b))
    for f in ("big.dmp", "small.dmp", "ce.out", "xv.out"):
        assert (a / f).read_bytes() == (b / f).read_bytes(), f
    assert worst_dump_difference(a, b) == 0.0


@pytest.mark.parametrize("algorithm,kinds", [("RADAU", ("a1", "a2", "a3", "yar", "beta")),
                                             ("BS", ("a1", "a2", "a3", "yar", "beta"))])
def test_dropin_cometary_matches_reference(tmp_path, algorithm, kinds):
    """(The driver refuses any non-gravitational force with BS2,
    mercury6_2.for:6547; the library's BS2 takes them, tests/test_gpu_nongrav.py.)"""
    s = encounter_system(4, 600, kinds, big=True)
    print(f"case: drop-in {algorithm} ngflag {s['ngflag']} (cometary on planets and particles), 600 d", flush=True)
    a, b = run_pair(tmp_path, s, 300, algorithm=algorithm, interval=50.0)
    worst = worst_dump_difference(a, b)
    print(f"max relative dump difference {worst:.3e}", flush=True)
    assert worst <= REL_TOL
    assert (a / "ce.out").read_bytes() == (b / "ce.out").read_bytes()
    assert (a / "xv.out").read_bytes() == (b / "xv.out").read_bytes()


def test_dropin_mvs_comets_match_reference(tmp_path):
    """MVS with a1 and a3 on every small body (the reference's case,
    tests/test_gpu_algorithms.py:92-93, over 600 particles)."""
    s = systems.config_small(seed=1, nparticles=600)
    n, nb = len(s["m"]), s["nbig"]
    s["ngf"] = np.zeros((n, 5))
    s["ngf"][nb:, 0] = 1e-9
    s["ngf"][nb:, 2] = 1e-9
    s["ngflag"] = 1
    print("case: drop-in MVS with a1 = a3 = 1e-9 on 600 particles, 456 steps of 8 d", flush=True)
    a, b = run_pair(tmp_path, s, 456, algorithm="MVS", interval=8.0 * 57)
    worst = worst_dump_difference(a, b)
    print(f"max relative dump difference {worst:.3e}", flush=True)
    assert worst <= REL_TOL
    assert (a / "ce.out").read_bytes() == (b / "ce.out").read_bytes()
    assert (a / "xv.out").read_bytes() == (b / "xv.out").read_bytes()


def test_dropin_published_extra_force_case(tmp_path):
    """The reference's published RADAU + 1PN + PR + Yarkovsky case at its full
    size (docs/benchmarks.md:36-44: 8 planets + 100,000 particles with
    beta 1e-4 and yar 1e-12 AU/day^2, 365 d, tol 1e-11, step 1 d, output at
    the final epoch): the drop-in run is byte-identical to the CPU driver's.
    Both whole-program wall times are printed (the reference publishes 6.04 s
    on an RTX 4070 and 33.4 s for its own CPU path)."""
    import time
    s = systems.config_extra_forces(nparticles=100_000)
    print("case: drop-in RADAU 1PN + PR + Yarkovsky, 8 planets + 100000 particles, 365 d", flush=True)
    a = R.write_run(tmp_path / "cpu", s, nsteps=365, backend="cpu", algorithm="RADAU", tol=1e-11, pn=True)
    b = R.write_run(tmp_path / "dev", s, nsteps=365, backend="hip", algorithm="RADAU", tol=1e-11, pn=True)
    t0 = time.perf_counter()
    ra = R.run(a, R.CPU_BIN)
    t1 = time.perf_counter()
    rb = R.run(b, R.HIP_BIN)
    t2 = time.perf_counter()
    assert ra.returncode == 0, ra.stdout + ra.stderr
    assert rb.returncode == 0, rb.stdout + rb.stderr
    print(f"whole program: CPU driver {t1 - t0:.2f} s, drop-in on the device {t2 - t1:.2f} s", flush=True)
    assert any("Execution backend: HIP" in ln for ln in R.info_lines(b))
    for f in ("big.dmp", "small.dmp", "ce.out", "xv.out"):
        assert (a / f).read_bytes() == (b / f).read_bytes(), f


def test_dropin_hybrid_comets_match_reference(tmp_path):
    """HYBRID with a1 and a3 on every small body (the reference's case,
    tests/test_gpu_algorithms.py:92-93), over 600 particles half near planet
    orbits, so that encounter members pass through mdt_hkce."""
    s = systems.config_encounters(seed=3, nparticles=600)
    n, nb = len(s["m"]), s["nbig"]
    s["ngf"] = np.zeros((n, 5))
    s["ngf"][nb:, 0] = 1e-9
    s["ngf"][nb:, 2] = 1e-9
    s["ngflag"] = 1
    print("case: drop-in HYBRID with a1 = a3 = 1e-9 on 600 particles, 300 steps of 2 d", flush=True)
    a, b = run_pair(tmp_path, s, 300, algorithm="HYBRID", interval=100.0)
    worst = worst_dump_difference(a, b)
    print(f"max relative dump difference {worst:.3e}", flush=True)
    assert worst <= REL_TOL
    assert (a / "ce.out").read_bytes() == (b / "ce.out").read_bytes()
    assert (a / "xv.out").read_bytes() == (b / "xv.out").read_bytes()
