"""The hand-written adjoint and tangent-linear MVS step on the device.

Checked four ways:
* the tape's forward pass is the MVS map itself (bitwise equal to
  mercury_hip_advance);
* gradients agree with central finite differences of the oracle's CPU MVS
  integrator (bit-identical to the reference's mdt_mvs), and the
  tangent-linear pass with central differences of every final state;
* the adjoint is the transpose of the tangent-linear pass: <grad, d> equals
  the tangent's dL for random directions d (at the full workload size);
* the checkpoint interval changes nothing (bitwise).
"""
import numpy as np
import pytest

from mercuda_amd import systems

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def fit_system(seed, nparticles, nplanets=8):
    """8 planets + particles as the orbit-fitting workload (massless test
    particles a U[0.5, 40] AU), jcen = 0."""
    s = systems.make_system(seed, nparticles, nplanets=nplanets, a_range=(0.5, 40.0), emax=0.3,
                            imax_deg=20.0, h=4.0)
    s["jcen"] = np.zeros(3)
    return s


def loss_and_cotangent(xT, target, nb):
    d = xT[nb:] - target
    gx = np.zeros_like(xT)
    gx[nb:] = 2.0 * d
    return float(np.sum(d * d)), gx


def test_tape_forward_is_the_mvs_map(hip):
    s = fit_system(41, 512)
    say("case: tape forward vs advance, 8 planets + 512 particles, 300 steps")
    hip.configure(1)
    hip.upload_system(s)
    x0, v0 = hip.download()
    xT, vT = hip.adjoint_forward(300, s["h"], segment=17)
    xa, va = hip.download()
    assert np.array_equal(xa, x0) and np.array_equal(va, v0), "the uploaded state must stay unchanged"
    hip.advance(0.0, s["h"], 300, s["rcen"], screen=False)
    xb, vb = hip.download()
    assert np.array_equal(xT, xb) and np.array_equal(vT, vb)


def test_gradient_matches_central_differences(hip, oracle):
    nsteps = 400
    s = fit_system(42, 64)
    nb = s["nbig"]
    rng = np.random.default_rng(5)
    target = s["x"][nb:] + rng.normal(0, 1e-3, (len(s["m"]) - nb, 3))
    say(f"case: gradient vs central differences, 8 planets + 64 particles, {nsteps} steps of 4 d")
    hip.configure(1)
    hip.upload_system(s)
    xT, vT = hip.adjoint_forward(nsteps, s["h"])
    L, gx = loss_and_cotangent(xT, target, nb)
    gx0, gv0, gm = hip.adjoint_backward(gx, np.zeros_like(vT))

    def loss_cpu(sd):
        x, v, _, _ = oracle.mvs_run(sd, nsteps, screen=False)
        return float(np.sum((x[nb:] - target) ** 2))

    # the oracle's forward is the same map: loss equal to rounding
    assert abs(loss_cpu(s) - L) <= 1e-12 * L
    def central(kind, j, c, step):
        vals = []
        for sign in (1, -1):
            sd = {k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v) for k, v in s.items()}
            if c is None:
                sd[kind][j] += sign * step
            else:
                sd[kind][j, c] += sign * step
            vals.append(loss_cpu(sd))
        return (vals[0] - vals[1]) / (2 * step)

    # Particle parameters: central differences at relative step 1e-7.
    # Planet parameters move the loss ~1e3 times less per unit of relative
    # step, so at 1e-7 the rounding of L dominates; there the difference is
    # Richardson-extrapolated from relative steps 5e-5 and 1e-4 (fourth
    # order), whose remaining noise is ~1e-4 relative.
    picks = [("x", nb + 3, 0, 2e-6), ("x", nb + 17, 2, 2e-6), ("v", nb + 40, 1, 2e-6), ("v", nb + 5, 0, 2e-6),
             ("x", 1, 0, 1e-3), ("x", 4, 1, 1e-3), ("x", 8, 2, 1e-3), ("v", 2, 1, 1e-3), ("v", 6, 0, 1e-3),
             ("v", 8, 2, 1e-3), ("m", 1, None, 1e-3), ("m", 5, None, 1e-3), ("m", 8, None, 1e-3)]
    for kind, j, c, tol in picks:
        base = s[kind][j] if c is None else s[kind][j, c]
        scale = abs(base) if kind == "m" else np.linalg.norm(s[kind][j])
        if j >= nb:
            fd = central(kind, j, c, 1e-7 * scale)
        else:
            h = 5e-5 * scale
            fd = (4.0 * central(kind, j, c, h) - central(kind, j, c, 2 * h)) / 3.0
        ad = {"x": gx0, "v": gv0}[kind][j, c] if c is not None else gm[j]
        rel = abs(ad - fd) / abs(fd)
        say(f"  d L / d {kind}[{j}{'' if c is None else ',' + str(c)}]: adjoint {ad:+.10e} fd {fd:+.10e} rel {rel:.2e}")
        assert rel <= tol


def test_planet_final_state_loss_matches_central_differences(hip, oracle):
    """Cotangents on the Big bodies' final states (positions and velocities of
    three planets) against central differences (relative step 1e-7) of the
    oracle's CPU integrator: planet initial states and masses; particle
    initial states get no gradient from it."""
    nsteps = 300
    s = fit_system(45, 16)
    nb = s["nbig"]
    rng = np.random.default_rng(6)
    sel = [2, 5, 8]
    wx = rng.normal(size=(len(sel), 3))
    wv = rng.normal(size=(len(sel), 3)) * 100.0
    say(f"case: loss on planet final states vs central differences, 8 planets + 16 particles, {nsteps} steps")
    hip.configure(1)
    hip.upload_system(s)
    xT, vT = hip.adjoint_forward(nsteps, s["h"])
    gx = np.zeros_like(xT); gv = np.zeros_like(vT)
    gx[sel] = 2.0 * wx * xT[sel]
    gv[sel] = wv
    gx0, gv0, gm = hip.adjoint_backward(gx, gv)

    def loss_cpu(sd):
        x, v, _, _ = oracle.mvs_run(sd, nsteps, screen=False)
        return float(np.sum(wx * x[sel] ** 2) + np.sum(wv * v[sel]))

    def central(kind, j, c, rel_step):
        base = s[kind][j] if c is None else s[kind][j, c]
        step = rel_step * (abs(base) if kind == "m" else np.linalg.norm(s[kind][j]))
        vals = []
        for sign in (1, -1):
            sd = {k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v) for k, v in s.items()}
            if c is None:
                sd[kind][j] += sign * step
            else:
                sd[kind][j, c] += sign * step
            vals.append(loss_cpu(sd))
        return (vals[0] - vals[1]) / (2 * step)

    # The loss's own planets' states: central differences at relative step
    # 1e-7.  Other planets' states reach the loss only through the planets'
    # mutual pull and move it ~1e3-1e6 times less: there the difference is
    # Richardson-extrapolated from relative steps 5e-5 and 1e-4.  The masses
    # (1e-10 to 1e-8 here) move the loss so little that below a relative step
    # of ~1e-4 its rounding scatters the differences by ~1e-3; they take a
    # plain central difference at 3e-4, whose truncation is ~1e-6.
    picks = [("x", 2, 0, "direct", 1e-5), ("x", 5, 1, "direct", 1e-5), ("x", 8, 2, "direct", 1e-5),
             ("v", 2, 2, "direct", 1e-5), ("v", 8, 0, "direct", 1e-5), ("x", 3, 1, "richardson", 1e-3),
             ("v", 1, 1, "richardson", 1e-3), ("m", 2, None, "wide", 1e-3), ("m", 5, None, "wide", 1e-3),
             ("m", 7, None, "wide", 1e-3)]
    for kind, j, c, mode, tol in picks:
        if mode == "direct":
            fd = central(kind, j, c, 1e-7)
        elif mode == "wide":
            fd = central(kind, j, c, 3e-4)
        else:
            fd = (4.0 * central(kind, j, c, 5e-5) - central(kind, j, c, 1e-4)) / 3.0
        ad = {"x": gx0, "v": gv0}[kind][j, c] if c is not None else gm[j]
        rel = abs(ad - fd) / abs(fd)
        say(f"  d L / d {kind}[{j}{'' if c is None else ',' + str(c)}]: adjoint {ad:+.10e} fd {fd:+.10e} rel {rel:.2e}")
        assert rel <= tol
    assert not np.any(gx0[nb:]) and not np.any(gv0[nb:])


def test_adjoint_is_transpose_of_tangent_with_planet_cotangents(hip):
    """<grad L, d> = dL along d from the tangent-linear pass when the
    cotangent covers Big bodies and particles together."""
    nsteps = 600
    s = fit_system(46, 300)
    nb = s["nbig"]
    n = len(s["m"])
    rng = np.random.default_rng(12)
    say(f"case: dot-product test with Big-body cotangents, 8 planets + 300 particles, {nsteps} steps")
    hip.configure(1)
    hip.upload_system(s)
    hip.adjoint_forward(nsteps, s["h"])
    gx = rng.normal(size=(n, 3)); gv = rng.normal(size=(n, 3)) * 50.0
    gx[0] = 0.0; gv[0] = 0.0
    gx0, gv0, gm = hip.adjoint_backward(gx, gv)
    for trial in range(3):
        dx0 = rng.normal(size=(n, 3)) * 1e-3
        dv0 = rng.normal(size=(n, 3)) * 1e-5
        dm = np.zeros(n)
        dm[1:nb] = rng.normal(size=nb - 1) * s["m"][1:nb] * 1e-2
        dx0[0] = 0.0; dv0[0] = 0.0
        if trial == 1:      # planets only
            dx0[nb:] = 0.0; dv0[nb:] = 0.0
        dx, dv = hip.tangent(dx0, dv0, dm)
        lin = float(np.sum(gx * dx) + np.sum(gv * dv))
        adj = float(np.sum(gx0 * dx0) + np.sum(gv0 * dv0) + np.sum(gm * dm))
        say(f"  trial {trial}: tangent {lin:+.12e} adjoint {adj:+.12e}")
        assert abs(lin - adj) <= 1e-9 * max(abs(lin), abs(adj))


def test_tangent_matches_differences_of_final_states(hip, oracle):
    """The tangent-linear pass against central differences (relative step
    1e-7 along the direction) of every body's final state from the oracle's
    CPU integrator: planets' own tangents (Jacobi transforms, drift with
    gm(m), Big forces) and the particles' response to them."""
    nsteps = 400
    s = fit_system(45, 48)
    nb = s["nbig"]
    n = len(s["m"])
    rng = np.random.default_rng(8)
    say(f"case: tangent vs differences of final states, 8 planets + 48 particles, {nsteps} steps")
    hip.configure(1)
    hip.upload_system(s)
    hip.adjoint_forward(nsteps, s["h"])
    for label, with_particles in (("planets and masses", False), ("all bodies", True)):
        dx0 = rng.normal(size=(n, 3)) * np.linalg.norm(s["x"], axis=1)[:, None]
        dv0 = rng.normal(size=(n, 3)) * np.linalg.norm(s["v"], axis=1)[:, None]
        dm = rng.normal(size=n) * s["m"]
        dx0[0] = dv0[0] = 0.0; dm[0] = 0.0; dm[nb:] = 0.0
        if not with_particles:
            dx0[nb:] = 0.0; dv0[nb:] = 0.0
        tx, tv = hip.tangent(dx0, dv0, dm)
        eps = 1e-7
        outs = []
        for sign in (1, -1):
            sd = dict(s)
            sd["x"] = s["x"] + sign * eps * dx0
            sd["v"] = s["v"] + sign * eps * dv0
            sd["m"] = s["m"] + sign * eps * dm
            x, v, _, _ = oracle.mvs_run(sd, nsteps, screen=False)
            outs.append((x, v))
        fx = (outs[0][0] - outs[1][0]) / (2 * eps)
        fv = (outs[0][1] - outs[1][1]) / (2 * eps)
        for name, rows in (("planets", slice(1, nb)), ("particles", slice(nb, n))):
            ex = np.linalg.norm(tx[rows] - fx[rows]) / np.linalg.norm(fx[rows])
            ev = np.linalg.norm(tv[rows] - fv[rows]) / np.linalg.norm(fv[rows])
            say(f"  {label}: {name} final x rel {ex:.2e}, v rel {ev:.2e}")
            # the differences themselves agree across steps 1e-6..1e-7 to
            # ~1e-8 for the planets and ~2e-6 for the particles' (small)
            # response to planet-only perturbations
            tol = 1e-6 if name == "planets" else 5e-6
            assert ex <= tol and ev <= tol


def test_adjoint_is_transpose_of_tangent_full_size(hip):
    """8 planets + 4,096 particles, 10^4 steps of 4 d (the orbit-fitting
    workload): <grad L, d> = dL/d along d from the tangent-linear pass."""
    nsteps = 10000
    s = fit_system(7, 4096)
    nb = s["nbig"]
    n = len(s["m"])
    rng = np.random.default_rng(11)
    target = s["x"][nb:] + rng.normal(0, 1e-3, (n - nb, 3))
    say("case: dot-product test, 8 planets + 4096 particles, 10^4 steps")
    hip.configure(1)
    hip.upload_system(s)
    xT, vT = hip.adjoint_forward(nsteps, s["h"])
    L, gx = loss_and_cotangent(xT, target, nb)
    gx0, gv0, gm = hip.adjoint_backward(gx, np.zeros_like(vT))
    assert np.all(np.isfinite(gx0)) and np.all(np.isfinite(gv0)) and np.all(np.isfinite(gm))
    for trial in range(3):
        dx0 = rng.normal(size=(n, 3)) * 1e-3
        dv0 = rng.normal(size=(n, 3)) * 1e-5
        dm = np.zeros(n)
        dm[1:nb] = rng.normal(size=nb - 1) * s["m"][1:nb] * 1e-2
        dx0[0] = 0.0; dv0[0] = 0.0
        if trial == 1:      # planets only
            dx0[nb:] = 0.0; dv0[nb:] = 0.0
        if trial == 2:      # particles only
            dx0[1:nb] = 0.0; dv0[1:nb] = 0.0; dm[:] = 0.0
        dx, dv = hip.tangent(dx0, dv0, dm)
        lin = float(np.sum(gx * dx))
        adj = float(np.sum(gx0 * dx0) + np.sum(gv0 * dv0) + np.sum(gm * dm))
        say(f"  trial {trial}: tangent {lin:+.12e} adjoint {adj:+.12e}")
        assert abs(lin - adj) <= 1e-9 * max(abs(lin), abs(adj))


def test_checkpoint_interval_changes_nothing(hip):
    s = fit_system(43, 300)
    nb = s["nbig"]
    say("case: checkpoint interval 1, 17, 250 over 250 steps")
    hip.configure(1)
    hip.upload_system(s)
    out = []
    for seg in (1, 17, 250):
        xT, vT = hip.adjoint_forward(250, s["h"], segment=seg)
        gx = np.zeros_like(xT); gx[nb:] = xT[nb:]
        gv = np.zeros_like(vT); gv[nb:] = 0.5 * vT[nb:]
        out.append(hip.adjoint_backward(gx, gv))
    for other in out[1:]:
        for a, b in zip(out[0], other):
            assert np.array_equal(a, b)


def test_torch_autograd_function(hip):
    import torch
    from mercuda_amd.adjoint import MVSFlow, library
    s = fit_system(44, 40, nplanets=3)
    nb = s["nbig"]
    x0 = torch.tensor(s["x"], requires_grad=True)
    v0 = torch.tensor(s["v"], requires_grad=True)
    m = torch.tensor(s["m"], requires_grad=True)
    xT, vT = MVSFlow.apply(x0, v0, m, nb, 120, s["h"])
    loss = (xT[nb:] ** 2).sum() + 3.0 * (vT[nb:, 2] ** 2).sum() + 0.5 * (xT[1:nb] ** 2).sum()
    loss.backward()
    lib = library()
    gx = np.zeros((len(s["m"]), 3)); gx[nb:] = 2 * xT.detach().numpy()[nb:]
    gx[1:nb] = xT.detach().numpy()[1:nb]
    gv = np.zeros((len(s["m"]), 3)); gv[nb:, 2] = 6 * vT.detach().numpy()[nb:, 2]
    xT2, _ = lib.adjoint_forward(120, s["h"])
    assert np.array_equal(xT2, xT.detach().numpy())
    gx0, gv0, gm = lib.adjoint_backward(gx, gv)
    assert np.array_equal(x0.grad.numpy(), gx0)
    assert np.array_equal(v0.grad.numpy(), gv0)
    assert np.array_equal(m.grad.numpy(), gm)
