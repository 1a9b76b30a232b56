"""The drop-in driver where its event consumers fire on the device path:
mergers of Big bodies (mce_coll, collisions on), central-body impacts (the
step redo through download(previous=1)) and ejections (mxx_ejec), each run
with `execution backend = hip` and with the reference CPU driver on the same
inputs.  The cases are the reference's own (tests/test_events.py:55-102):
the surviving bodies, their masses and states, the encounter file, the
compressed output and the event lines of info.out must agree."""
import numpy as np
import pytest

import mercury_run as R
from helpers import REL_TOL

needs_dropin = pytest.mark.skipif(not (R.HIP_BIN.exists() and R.CPU_BIN.exists()),
                                  reason="drop-in or reference driver not built (make -C oracle ref dropin)")
pytestmark = [needs_dropin, pytest.mark.gpu]


def event_lines(path, names):
    """info.out lines that name a body (collisions, impacts, ejections)."""
    return [ln for ln in R.info_lines(path) if any(n in ln for n in names)]


def run_both(tmp_path, label, names, **case):
    a = R.write_case(tmp_path / "cpu", backend="cpu", **case)
    b = R.write_case(tmp_path / "hip", backend="hip", **case)
    print(f"case: {label}", flush=True)
    ra = R.run(a, R.CPU_BIN)
    assert ra.returncode == 0, ra.stdout + ra.stderr
    rb = R.run(b, R.HIP_BIN)
    assert rb.returncode == 0, rb.stdout + rb.stderr
    assert any("Execution backend: HIP" in ln for ln in R.info_lines(b)), R.info_lines(b)
    worst = 0.0
    out = {}
    for f in ("big.dmp", "small.dmp"):
        da, db = R.read_dump_full(a, f), R.read_dump_full(b, f)
        assert da.keys() == db.keys(), (f, sorted(da), sorted(db))
        for k in da:
            assert da[k]["mass"] == db[k]["mass"], (k, da[k]["mass"], db[k]["mass"])
            for q in ("x", "v"):
                worst = max(worst, np.linalg.norm(da[k][q] - db[k][q]) / max(np.linalg.norm(da[k][q]), 1e-300))
        out[f] = db
    print(f"{label}: max relative dump difference {worst:.3e}", flush=True)
    assert worst <= REL_TOL
    assert (a / "ce.out").read_bytes() == (b / "ce.out").read_bytes()
    assert (a / "xv.out").read_bytes() == (b / "xv.out").read_bytes()
    assert event_lines(a, names) == event_lines(b, names)
    return out, event_lines(b, names)


@pytest.mark.parametrize("algorithm", ["BS", "BS2", "RADAU", "HYBRID"])
def test_dropin_merge(tmp_path, algorithm):
    """Two Big bodies of low density overlap and merge into the more massive
    one; the integration continues with the merged body."""
    a = R.body("A", mass=1e-8, e=0.0, d=1e-6)
    b = R.body("B", mass=2e-8, e=0.0, d=1e-6)
    b["x"][0] += 3e-4
    out, ev = run_both(tmp_path, f"{algorithm} merger", ["A", "B"], algorithm=algorithm, big=[a, b],
                       small=[R.body("P", a=2.0)], stop=0.2, step=0.01, interval=0.05, collisions=True)
    assert set(out["big.dmp"]) == {"B"}
    assert abs(out["big.dmp"]["B"]["mass"] - 3e-8) <= 1e-20


def test_dropin_hybrid_merge_and_central_impact(tmp_path):
    a = R.body("A", mass=1e-8, e=0.0, d=1e-6)
    b = R.body("B", mass=2e-8, e=0.0, d=1e-6)
    b["x"][0] += 3e-4
    impact = R.body("IMPACT")
    impact["x"] = [0.00505, 0.0, 0.0]; impact["v"] = [-0.1, 0.0, 0.0]
    out, ev = run_both(tmp_path, "HYBRID merger and central impact", ["A", "B", "IMPACT"], algorithm="HYBRID",
                       big=[a, b], small=[impact, R.body("KEEP", a=2.0)], collisions=True, stop=0.01,
                       step=0.001, interval=0.001)
    assert set(out["big.dmp"]) == {"B"} and set(out["small.dmp"]) == {"KEEP"}


@pytest.mark.parametrize("algorithm", ["BS", "BS2", "RADAU", "MVS", "HYBRID"])
def test_dropin_central_impact_and_ejection(tmp_path, algorithm):
    """A particle falls into the central body (the driver redoes the step
    from the saved start state) and another passes the ejection distance."""
    infall = R.body("IMPACT")
    infall["x"] = [0.006, 0.0, 0.0]; infall["v"] = [-0.1, 0.0, 0.0]
    escaping = R.body("ESCAPE")
    escaping["x"] = [99.9, 0.0, 0.0]; escaping["v"] = [10.0, 0.0, 0.0]
    # the survivor carries cometary A2, Yarkovsky and radiation parameters
    # where the driver allows them (tests/test_events.py:92-98)
    params = {"a2": 2e-11, "yar": 3e-12, "b": 1e-4} if algorithm in ("BS", "RADAU") else {}
    out, ev = run_both(tmp_path, f"{algorithm} central impact and ejection", ["IMPACT", "ESCAPE"],
                       algorithm=algorithm, small=[infall, escaping, R.body("KEEP", **params)], stop=0.3,
                       step=0.001, interval=0.05, periodic=10)
    assert set(out["small.dmp"]) == {"KEEP"}
    for key, value in params.items():
        assert out["small.dmp"]["KEEP"]["params"][key] == value
    assert any("IMPACT" in ln for ln in ev) and any("ESCAPE" in ln for ln in ev)


def test_dropin_hybrid_encounter_at_large_epoch(tmp_path):
    """Two Big bodies inside each other's changeover radius at an epoch of
    1e15 d, where one ulp of the epoch (0.125 d) exceeds the encounter set's
    BS2 substeps: both backends must measure the substeps' progress on the
    encounter's local clock (mdt_hkce, mercury6_2.for:3883-3905).  The states
    after 8 days agree with the same run at epoch 0, on the CPU and on the
    device (the reference's case, tests/test_gpu_algorithms.py:49-65)."""
    big = [R.body("A", mass=1e-4, a=1.0), R.body("B", mass=1e-4, a=1.005, phase=0.005)]
    states = []
    for k, (epoch, backend, binary) in enumerate([(0.0, "cpu", R.CPU_BIN), (1e15, "cpu", R.CPU_BIN),
                                                  (1e15, "hip", R.HIP_BIN)]):
        print(f"case: HYBRID encounter pair, epoch {epoch:g}, backend {backend}", flush=True)
        p = R.write_case(tmp_path / str(k), small=[], big=big, algorithm="HYBRID", backend=backend,
                         epoch=epoch, start=epoch, stop=epoch + 8.0, interval=8.0, step=0.5)
        r = R.run(p, binary)
        assert r.returncode == 0, r.stdout + r.stderr
        if backend == "hip":
            assert any("Execution backend: HIP" in ln for ln in R.info_lines(p)), R.info_lines(p)
        states.append(R.read_dump_full(p, "big.dmp"))
    for st in states[1:]:
        assert st.keys() == states[0].keys()
        for name in st:
            for q in ("x", "v"):
                d = np.max(np.abs(st[name][q] - states[0][name][q]))
                print(f"{name} {q}: max |difference| from the epoch-0 run {d:.3e}", flush=True)
                assert d < 1e-10
