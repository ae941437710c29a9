"""Batched independent instances (fg_create_batch): several meshes in one context. Each instance
of the batch must step exactly as it does alone: same initial state, field and basis angle, then
the states after each step agree to the solver tolerance (direct and FMM demag), and the union
energies are the sums of the instances' energies."""
import math

import numpy as np
import pytest

from feellgood_amd import native as nv
from feellgood_amd.mesh import HostMesh


def dots(n=3, seed=4):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        a, b, c = rng.uniform(10, 22), rng.uniform(10, 22), rng.uniform(3, 8)
        m = HostMesh.ellipsoid((0.0, 0.0, 0.0), (a, b, c), 2.5)
        uk = rng.normal(size=3)
        m.prepare({"volume": dict(Ms=rng.uniform(6e5, 1.4e6), Ae=1.3e-11, alpha=0.1, K=rng.uniform(0, 5e4),
                                  uk=tuple(uk / np.linalg.norm(uk)))})
        u = rng.normal(size=(m.n_nodes, 3))
        out.append((m, u / np.linalg.norm(u, axis=1, keepdims=True)))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("demag", ["direct", "fmm"])
def test_batch_steps_like_the_instances_alone(demag):
    from feellgood_amd.engine import Engine
    items = dots()
    kw = dict(tol=1e-10, max_iter=2000, demag=demag, fmm_order=8, fmm_leaf=32)
    g = Engine.batch([m.mesh for m, _ in items], **kw)
    noff, toff = g.layout()
    assert len(noff) == len(items) + 1 and noff[-1] == sum(m.n_nodes for m, _ in items)
    alone = [Engine(m.mesh, **kw) for m, _ in items]
    u0 = np.concatenate([u for _, u in items])
    H = np.array([0.01, -0.02, 0.015]) / nv.MU0
    dt, pref = 1e-14, 1 + 0.01 * abs(math.log(0.01))
    g.set_state(nv.FG_BOTH, u=u0)
    for e, (_, u) in zip(alone, items):
        e.set_state(nv.FG_BOTH, u=u)
    for e in [g] + alone:
        e.calc_demag()
        e.evolution()
    for step in range(3):
        r = 0.17 * (step + 1)
        E_alone = np.zeros(4)
        for e in [g] + alone:
            e.base_projection(r)
            e.prepare_elements(H, dt, pref)
            failed, info = e.solve(dt)
            assert not failed
            e.calc_demag()
            E = e.energy(H)
            if e is not g:
                E_alone += E
            else:
                E_batch = E
            e.evolution()
        ub = g.get_state(nv.FG_CURRENT)[0]
        for k, e in enumerate(alone):
            ua = e.get_state(nv.FG_CURRENT)[0]
            assert np.abs(ub[noff[k]:noff[k + 1]] - ua).max() <= 1e-9, (step, k)
        assert np.allclose(E_batch, E_alone, rtol=1e-8, atol=1e-8 * np.abs(E_alone).max())


@pytest.mark.gpu
@pytest.mark.parametrize("block", [256, 1024])
def test_persistent_batch_solve_like_the_instances_alone(block):
    """FG_BATCH_PERSISTENT: one launch with a block per instance, each instance solved to its own
    criterion ||r_k|| <= tol ||b_k||, the criterion of the instance alone: three steps (direct
    demag) agree with the instances alone as above, the batch's iteration count is the largest
    instance count within one, and an iteration limit that stops the instances is reported as the
    overflow status."""
    from feellgood_amd.engine import Engine
    items = dots()
    kw = dict(tol=1e-10, max_iter=2000, demag="direct")
    g = Engine.batch([m.mesh for m, _ in items], batch_solver="persistent", batch_block=block, **kw)
    noff, _ = g.layout()
    alone = [Engine(m.mesh, **kw) for m, _ in items]
    u0 = np.concatenate([u for _, u in items])
    H = np.array([0.01, -0.02, 0.015]) / nv.MU0
    dt, pref = 1e-14, 1 + 0.01 * abs(math.log(0.01))
    g.set_state(nv.FG_BOTH, u=u0)
    for e, (_, u) in zip(alone, items):
        e.set_state(nv.FG_BOTH, u=u)
    for e in [g] + alone:
        e.calc_demag()
        e.evolution()
    for step in range(3):
        r = 0.23 * (step + 1)
        its = []
        for e in [g] + alone:
            e.base_projection(r)
            e.prepare_elements(H, dt, pref)
            failed, info = e.solve(dt)
            assert not failed and info.status == nv.FG_ITER_CONVERGED
            its.append(info.iterations)
            e.calc_demag()
            e.evolution()
        assert abs(its[0] - max(its[1:])) <= 1, its
        ub = g.get_state(nv.FG_CURRENT)[0]
        for k, e in enumerate(alone):
            ua = e.get_state(nv.FG_CURRENT)[0]
            assert np.abs(ub[noff[k]:noff[k + 1]] - ua).max() <= 1e-9, (step, k)
    g.close()
    g = Engine.batch([m.mesh for m, _ in items], batch_solver="persistent", batch_block=block, tol=1e-14,
                     max_iter=2, demag="direct")
    g.set_state(nv.FG_BOTH, u=u0)
    g.calc_demag()
    g.evolution()
    g.base_projection(0.3)
    g.prepare_elements(H, dt, pref)
    failed, info = g.solve(dt)
    assert failed and info.status == nv.FG_ITER_OVERFLOW and info.iterations == 2
    g.close()
