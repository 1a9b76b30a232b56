"""hipGraph replay and the autotuner for batched ensembles and single systems.

A captured advance call is replayed for later calls of the same shape; every
replay must give bitwise the states and events of the eager launches, with
the program's own eager work (downloads, synchronisations) in between.  The
autotuner must leave the state untouched, and its choice must not change
results (launch boundaries never do)."""
import numpy as np
import pytest

from mercuda_amd import systems

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def run_calls(hip, upload, calls, advance, graph):
    hip.set_graph(False)
    upload()
    hip.set_graph(graph)
    out = []
    for (t, n) in calls:
        advance(t, n)
        x, v = hip.download()
        out.append((x, v, sorted(hip.advance_events())))
    hip.set_graph(False)
    return out


def test_graph_replays_equal_eager_ensemble(hip):
    ens = systems.ensemble(seed=12, nsystems=12, nmin=50, nmax=4000)
    for q in ens[::2]:
        q["rce"][1:q["nbig"]] *= 4.0
    calls = [(0.0, 40), (160.0, 40), (320.0, 40), (480.0, 40), (640.0, 40), (800.0, 24)]
    say(f"case: graph replay, ensemble of {len(ens)} systems, {len(calls)} calls")
    hip.configure(1)
    hip.set_tuning(steps_per_launch=16, block_size=128)

    def upload():
        hip.configure(1)
        hip.ensemble_upload(ens)

    adv = lambda t, n: hip.ensemble_advance(t, n, screen=True)
    eager = run_calls(hip, upload, calls, adv, False)
    before = hip.graph_stats()
    graph = run_calls(hip, upload, calls, adv, True)
    after = hip.graph_stats()
    say(f"graph stats before {before} after {after}")
    assert after["captures"] - before["captures"] >= 2      # the reset call, then the steady shape
    assert after["replays"] - before["replays"] >= 3
    for k, (a, b) in enumerate(zip(eager, graph)):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), f"call {k}"
        assert a[2] == b[2], f"events of call {k}"


def test_graph_replays_equal_eager_single(hip):
    s = systems.config_encounters(seed=3, nparticles=3000)
    s["rce"][1:s["nbig"]] *= 3.0
    h = s["h"]
    calls = [(k * 48 * h, 48) for k in range(5)]
    say("case: graph replay, one system of 3000 particles, 5 calls of 48 steps")
    hip.configure(1)
    hip.set_tuning(steps_per_launch=32, block_size=128)

    def upload():
        hip.configure(1)
        hip.upload_system(s)

    adv = lambda t, n: hip.advance(t, h, n, s["rcen"], screen=True)
    eager = run_calls(hip, upload, calls, adv, False)
    graph = run_calls(hip, upload, calls, adv, True)
    for k, (a, b) in enumerate(zip(eager, graph)):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), f"call {k}"
        assert a[2] == b[2], f"events of call {k}"
    assert sum(len(c[2]) for c in eager) > 0


def test_autotune_leaves_state_and_results(hip):
    ens = systems.ensemble(seed=13, nsystems=10, nmin=100, nmax=5000)
    say(f"case: autotune on an ensemble of {len(ens)} systems")
    hip.configure(1)
    hip.ensemble_upload(ens)
    hip.set_tuning(steps_per_launch=16, block_size=128)
    hip.ensemble_advance(0.0, 20, screen=True)
    ref = hip.download()
    choice = hip.autotune(probe_steps=16)
    say(f"autotune choice {choice}")
    assert choice["steps_per_launch"] in (8, 16, 32, 64) and choice["block"] in (64, 128, 256)
    assert choice["particles_per_lane"] in (1, 2, 4)
    after = hip.download()
    assert np.array_equal(ref[0], after[0]) and np.array_equal(ref[1], after[1])
    hip.ensemble_advance(80.0, 50, screen=True)
    tuned = hip.download()
    hip.configure(1)
    hip.ensemble_upload(ens)
    hip.set_tuning(steps_per_launch=16, block_size=128, particles_per_lane=1)
    hip.ensemble_advance(0.0, 20, screen=True)
    hip.ensemble_advance(80.0, 50, screen=True)
    plain = hip.download()
    assert np.array_equal(tuned[0], plain[0]) and np.array_equal(tuned[1], plain[1])


def test_graph_after_reupload_with_more_particles(hip):
    """A captured graph is not replayed for a different upload: re-uploading a
    system with a few more particles (same capacity and tiles) integrates all
    of them, as the eager launches do."""
    s = systems.config_small(seed=8, nparticles=1000)
    more = systems.config_small(seed=8, nparticles=1040)
    h = s["h"]
    say("case: graph capture on 1000 particles, then replay attempt after re-upload of 1040")

    def run(graph):
        hip.set_graph(False)
        hip.configure(1)
        hip.set_tuning(steps_per_launch=16, block_size=128)
        hip.upload_system(s)
        hip.set_graph(graph)
        hip.advance(0.0, h, 32, s["rcen"], screen=True)
        hip.advance(32 * h, h, 32, s["rcen"], screen=True)
        hip.upload_system(more)
        hip.advance(0.0, h, 32, more["rcen"], screen=True)
        hip.advance(32 * h, h, 32, more["rcen"], screen=True)
        out = hip.download()
        hip.set_graph(False)
        return out

    x1, v1 = run(False)
    x2, v2 = run(True)
    assert np.array_equal(x1, x2) and np.array_equal(v1, v2)
    assert not np.array_equal(x1[-40:], more["x"][-40:])
