"""The learned initial guess of the BiCGSTAB solve.

CPU: the oracle's restatement of the network, pinned by central finite differences of its loss
(gradient), by the closed form of a zero output head (the reference guess, loss = loss_ref) and
by the tangency of its prediction.
GPU: the engine's features against numpy recomputations from the assembled system, its forward
prediction and gradient against the oracle on the same batch, and training: on held-out applied
fields the trained network's guess error falls well below its untrained base's (the linear
extrapolation) and it takes no more iterations than the base, while the solve still meets its
tolerance.
"""
import math

import numpy as np
import pytest

import oracle
from feellgood_amd import native as nv
from feellgood_amd.mesh import HostMesh

NP = nv.FG_GUESS_NPARAM
_O = nv.GUESS_OFFSETS
W1, B1, W2, B2, W3, B3 = _O["W1"], _O["b1"], _O["W2"], _O["b2"], _O["W3"], _O["b3"]
NS = nv.FG_GUESS_NSCALE


def random_problem(n=40, seed=0):
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    h = rng.normal(size=(n, 3)) * 3e4
    v = rng.normal(size=(n, 3)) * 2e9
    d2 = rng.normal(size=(n, 3)) * 1e7
    feat = np.concatenate([u, h, v, d2], axis=1).astype(np.float32)
    nb = [set() for _ in range(n)]
    for i in range(n):
        for j in ((i + 1) % n, (i + 7) % n):
            nb[i].add(j)
            nb[j].add(i)
    ptr = np.zeros(n + 1, dtype=np.int64)
    idx = []
    for i in range(n):
        idx += sorted(nb[i] | {i})
        ptr[i + 1] = len(idx)
    mag = np.ones(n, dtype=np.int32)
    mag[5] = 0
    target = v + rng.normal(size=(n, 3)) * 1e8
    target -= np.sum(target * u, axis=1, keepdims=True) * u
    p = (rng.normal(size=NP) * 0.2).astype(np.float32)
    p[-NS:] = [1 / 3e4, 1 / 2e9, 1e8, 1 / 1e7]
    return p, feat, ptr, np.array(idx, dtype=np.int32), mag, target


def test_oracle_gradient_matches_finite_differences():
    p, feat, ptr, idx, mag, tgt = random_problem()
    _, L, _, g = oracle.guess_eval(p, feat, ptr, idx, mag, tgt, round_bf16=False)
    rng = np.random.default_rng(1)
    picks = np.concatenate([rng.choice(np.arange(a, b), 8, replace=False)
                            for a, b in ((W1, B1), (B1, W2), (W2, B2), (B2, W3), (W3, B3))] + [np.arange(B3, B3 + 3)])
    for k in picks:
        hstep = 1e-2 * max(abs(float(p[k])), 0.1)
        q = p.copy()
        q[k] = p[k] + hstep
        lp = oracle.guess_eval(q, feat, ptr, idx, mag, tgt, round_bf16=False, want_grad=False)[1]
        q[k] = p[k] - hstep
        lm = oracle.guess_eval(q, feat, ptr, idx, mag, tgt, round_bf16=False, want_grad=False)[1]
        hh = (np.float64(p[k] + hstep) - np.float64(p[k] - hstep))
        fd = (lp - lm) / hh
        assert abs(fd - g[k]) <= 2e-3 * max(abs(g[k]), 1e-3 * L), (k, fd, g[k])


def test_oracle_zero_head_is_the_reference_guess():
    p, feat, ptr, idx, mag, tgt = random_problem(seed=2)
    p[W3:B3 + 3] = 0
    vp, L, Lr, _ = oracle.guess_eval(p, feat, ptr, idx, mag, tgt)
    u, v = feat[:, :3].astype(np.float64), feat[:, 6:9].astype(np.float64)
    ref = v - np.sum(v * u, axis=1, keepdims=True) * u
    ok = mag.astype(bool)
    assert np.allclose(vp[ok], ref[ok], rtol=0, atol=1e-6 * np.abs(ref).max())
    assert L == pytest.approx(Lr, rel=1e-12)
    # with a base (the engine's 16-column export), the zero head predicts the projected base
    rng = np.random.default_rng(4)
    base = (v + rng.normal(size=v.shape) * 3e8).astype(np.float32)
    row = np.concatenate([feat, base, np.zeros((len(feat), 1), np.float32)], axis=1)
    vb, L2, Lr2, _ = oracle.guess_eval(p, row, ptr, idx, mag, tgt)
    b64 = base.astype(np.float64)
    refb = b64 - np.sum(b64 * u, axis=1, keepdims=True) * u
    assert np.allclose(vb[ok], refb[ok], rtol=0, atol=1e-6 * np.abs(refb).max())
    assert Lr2 == pytest.approx(Lr, rel=1e-12) and L2 != pytest.approx(L, rel=1e-3)


def test_oracle_prediction_is_tangent():
    p, feat, ptr, idx, mag, tgt = random_problem(seed=3)
    vp, *_ = oracle.guess_eval(p, feat, ptr, idx, mag, tgt)
    u = feat[:, :3].astype(np.float64)
    # u is stored in FP32 (|u| = 1 to 1e-7): tangent to that precision
    assert np.abs(np.sum(vp * u, axis=1)).max() <= 1e-6 * np.abs(vp).max()


# ---------------------------------------------------------------------------------------- GPU
def cube(n=8, seed=4):
    m = HostMesh.cuboid((40.0, 40.0, 40.0), (n, n, n), jitter=0.1, seed=seed)
    m.prepare({"volume": dict(Ms=8e5, Ae=1.3e-11, alpha=0.1)})
    return m


def adjacency(m):
    a = m.arrays()
    tn = a["tets"]
    n = m.n_nodes
    nb = [set([i]) for i in range(n)]
    for t in tn:
        for x in t:
            nb[x].update(int(y) for y in t)
    ptr = np.zeros(n + 1, dtype=np.int64)
    idx = []
    for i in range(n):
        idx += sorted(nb[i])
        ptr[i + 1] = len(idx)
    return ptr, np.array(idx, dtype=np.int32)


def lumped_mass(m):
    a = m.arrays()
    pos = a["pos"]
    tn = a["tets"]
    d = pos[tn[:, 1:]] - pos[tn[:, :1]]
    vol = np.abs(np.linalg.det(d)) / 6
    mass = np.zeros(m.n_nodes)
    np.add.at(mass, tn.ravel(), np.repeat(vol / 4, 4))
    return mass


def initial_u(m, seed=5):
    rng = np.random.default_rng(seed)
    u = np.array([1.0, 1.0, 0.2]) + rng.uniform(-0.3, 0.3, size=(m.n_nodes, 3))
    return u / np.linalg.norm(u, axis=1, keepdims=True)


def field(theta, b=0.03):
    return np.array([math.cos(theta), math.sin(theta), 0.3]) * b / nv.MU0


@pytest.mark.gpu
def test_features_forward_and_gradient_match_oracle():
    from feellgood_amd.engine import Engine
    m = cube()
    g = Engine(m.mesh, tol=1e-10, max_iter=1000, demag="direct")
    g.set_state(nv.FG_BOTH, u=initial_u(m))
    g.calc_demag()
    g.evolution()
    g.guess_init(seed=7)
    dt, pref, H = 2e-14, 1 + 0.01 * abs(math.log(0.01)), field(0.4)
    v_hist = []
    for k in range(4):
        u_before = g.get_state(nv.FG_CURRENT)[0]
        v_before = g.get_state(nv.FG_NEXT)[1]
        v_hist.append(v_before)
        g.base_projection(0.1 * k)
        g.prepare_elements(H, dt, pref)
        failed, info = g.solve(dt)
        assert not failed
        if k == 3:
            ep, eq = g.basis()
            _, _, _, rhs = g.system()
            break
        g.calc_demag()
        g.evolution()
    feat = g.guess_features()
    # features: u, tangent RHS over the lumped mass, previous v
    b = rhs.reshape(-1, 2)
    h = (b[:, 1:2] * ep + b[:, 0:1] * eq) / lumped_mass(m)[:, None]
    assert np.allclose(feat[:, :3], u_before, atol=1e-6)
    assert np.allclose(feat[:, 6:9], v_before, rtol=1e-6, atol=1e-6 * np.abs(v_before).max())
    assert np.allclose(feat[:, 3:6], h, rtol=1e-5, atol=1e-5 * np.abs(h).max())
    # the second difference of the last three solutions (all at the same dt: from the fourth
    # assembly after set_state on), and the base: the linear extrapolation of the last two
    d2 = v_hist[3] - 2 * v_hist[2] + v_hist[1]
    assert np.allclose(feat[:, 9:12], d2, rtol=1e-4, atol=1e-4 * np.abs(d2).max())
    base = 2 * v_hist[3] - v_hist[2]
    assert np.allclose(feat[:, 12:15], base, rtol=1e-5, atol=1e-5 * np.abs(base).max())
    # gradient on this batch (scales are set from it on the first call)
    grad, loss, loss_ref = g.guess_gradient()
    p = g.guess_get()
    # scales set from this batch: the untrained network (the base) scores 1
    assert p[-NS:].min() > 0 and loss == pytest.approx(1.0, rel=1e-4)
    vstar = g.get_state(nv.FG_NEXT)[1]
    ptr, idx = adjacency(m)
    mag = np.ones(m.n_nodes, dtype=np.int32)
    _, L, Lr, go = oracle.guess_eval(p, feat, ptr, idx, mag, vstar)
    assert loss == pytest.approx(L, rel=1e-4) and loss_ref == pytest.approx(Lr, rel=1e-4)
    # W3 = 0 at initialisation: only the head has a gradient, the hidden layers have none
    assert np.abs(grad[:W3]).max() == 0 and np.abs(go[:W3]).max() == 0
    assert np.allclose(grad, go, rtol=1e-3, atol=1e-4 * np.abs(go).max())
    # after a few updates every layer has a gradient; compare again
    for _ in range(3):
        g.guess_train(3e-3)
    grad, loss, _ = g.guess_gradient()
    p = g.guess_get()
    _, L, _, go = oracle.guess_eval(p, feat, ptr, idx, mag, vstar)
    assert loss == pytest.approx(L, rel=1e-4)
    assert np.abs(go[:W3]).max() > 0
    assert np.allclose(grad, go, rtol=2e-3, atol=2e-4 * np.abs(go).max())


@pytest.mark.gpu
def test_inference_pass_matches_oracle():
    """The matrix-core inference pass that seeds the solve: with a tolerance the initial residual
    already meets, the solve stops at iteration 0 and NEXT v is the guess, v_pred projected on the
    tangent basis; compared with the oracle's network on the recorded features."""
    from feellgood_amd.engine import Engine
    m = cube()
    u0 = initial_u(m)
    dt, pref, H = 2e-14, 1 + 0.01 * abs(math.log(0.01)), field(1.1)
    ref = Engine(m.mesh, tol=1e-10, max_iter=1000, demag="direct")
    ref.set_state(nv.FG_BOTH, u=u0)
    ref.calc_demag()
    ref.evolution()
    ref.guess_init(seed=3)
    ref.base_projection(0.2)
    ref.prepare_elements(H, dt, pref)
    assert not ref.solve(dt)[0]
    ref.guess_gradient()  # sets the feature and output scales from this solve
    p = ref.guess_get()
    rng = np.random.default_rng(8)
    p[W3:B3 + 3] = rng.normal(size=B3 + 3 - W3) * 0.2
    ref.close()
    g = Engine(m.mesh, tol=1e300, max_iter=10, demag="direct")
    g.set_state(nv.FG_BOTH, u=u0, v=rng.normal(size=(m.n_nodes, 3)) * 1e9)
    g.calc_demag()
    g.evolution()
    g.guess_init(seed=3)
    g.guess_set(p)
    g.guess_enable(True)
    g.base_projection(0.7)
    g.prepare_elements(H, dt, pref)
    failed, info = g.solve(dt)
    assert not failed and info.iterations == 0
    feat = g.guess_features()
    ptr, idx = adjacency(m)
    vp, *_ = oracle.guess_eval(p, feat, ptr, idx, np.ones(m.n_nodes, dtype=np.int32),
                               np.zeros((m.n_nodes, 3)), want_grad=False)
    ep, eq = g.basis()
    proj = np.sum(vp * ep, axis=1, keepdims=True) * ep + np.sum(vp * eq, axis=1, keepdims=True) * eq
    v = g.get_state(nv.FG_NEXT)[1]
    # the network's own contribution must be visible against its base
    prev = feat[:, 12:15].astype(np.float64)
    assert np.abs(vp - prev).max() > 1e-3 * np.abs(prev).max()
    assert np.allclose(v, proj, rtol=0, atol=2e-5 * np.abs(proj).max())
    g.close()


@pytest.mark.gpu
def test_training_beats_the_untrained_base_on_held_out_fields():
    """Trained on plain solves of five field directions, the network against its own untrained base
    (W3 = b3 = 0: the linear extrapolation of the last two solutions) and against the reference
    guess (previous v) on two held-out directions, from the same states and basis angles: its loss
    (the guess's error against the converged solution) falls below 0.7 of the base's, and it takes
    no more BiCGSTAB iterations than the base, which takes no more than the reference guess."""
    from feellgood_amd.engine import Engine
    m = cube(10)
    dt, pref = 2e-14, 1 + 0.01 * abs(math.log(0.01))
    g = Engine(m.mesh, tol=1e-10, max_iter=1000, demag="direct")
    g.guess_init(seed=11)
    u0 = initial_u(m)
    rng = np.random.default_rng(12)
    for ep in range(6):
        for theta in (0.0, 1.3, 2.6, 3.9, 5.2):
            g.set_state(nv.FG_BOTH, u=u0)
            g.calc_demag()
            g.evolution()
            H = field(theta + 0.1 * ep)
            for k in range(10):
                g.base_projection(rng.uniform(0, 2 / math.pi))
                g.prepare_elements(H, dt, pref)
                failed, _ = g.solve(dt)
                assert not failed
                if k >= 2:
                    for _ in range(4):
                        g.guess_train(2e-3)
                g.calc_demag()
                g.evolution()
    trained = g.guess_get()
    base = trained.copy()
    base[W3:B3 + 3] = 0
    its, loss = {}, {}
    for name, params, on in (("reference", trained, False), ("base", base, True), ("trained", trained, True)):
        g.guess_set(params)
        g.guess_enable(on)
        rng = np.random.default_rng(99)
        tot, ls = 0, []
        for theta in (0.65, 3.25):
            g.set_state(nv.FG_BOTH, u=u0)
            g.calc_demag()
            g.evolution()
            for k in range(10):
                g.base_projection(rng.uniform(0, 2 / math.pi))
                g.prepare_elements(field(theta), dt, pref)
                failed, info = g.solve(dt)
                assert not failed and info.residual <= 1e-10
                tot += info.iterations
                if k >= 3:  # the second difference is there from the fourth solve on
                    ls.append(g.guess_gradient()[1])
                g.calc_demag()
                g.evolution()
        its[name], loss[name] = tot, float(np.mean(ls))
    print(f"held-out: iterations {its}, loss trained / base {loss['trained'] / loss['base']:.3f}", flush=True)
    assert loss["trained"] < 0.7 * loss["base"], loss
    assert its["trained"] <= its["base"] <= its["reference"], its
