"""The octree FMM of the demag potential.

CPU: the host evaluation of the FMM (fg_fmm_reference, the same plan and operators as the device
pass) against exact direct sums of the oracle's charges: the error falls with the expansion order
and is bounded at each order.
GPU: the device FMM against the host FMM (same algorithm: 1e-10) and against the oracle's exact
potentials (bounded by the truncation error), then one LLG step with FMM demag against the exact
oracle step (|dm| <= 1e-6).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from feellgood_amd import native as nv
from feellgood_amd.mesh import HostMesh


def film(nx=20, ny=10, nz=3, seed=3):
    m = HostMesh.cuboid((10.0 * nx, 10.0 * ny, 10.0 * nz), (nx, ny, nz), jitter=0.1, seed=seed)
    m.prepare({"volume": dict(Ms=8e5, Ae=1.3e-11, alpha=0.02)})
    return m


def charged(m, seed=1):
    p = oracle.Problem(m.mesh)
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(m.n_nodes, 3)) * 0.3 + [1, 0.5, 0]
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    v = rng.normal(size=(m.n_nodes, 3)) * 1e9
    p.set_state(nv.FG_BOTH, u=u, v=v)
    p.calc_demag()
    return p, u, v


def fmm_ref(xyz, qu, qv, tgt, order, leaf, far=0):
    pu = np.zeros(len(tgt))
    pv = np.zeros(len(tgt))
    L = C.c_int()
    rc = nv.hip_lib().fg_fmm_reference(nv.dptr(np.ascontiguousarray(xyz.reshape(-1))), len(qu),
                                       nv.dptr(np.ascontiguousarray(tgt.reshape(-1))), len(tgt),
                                       nv.dptr(qu), nv.dptr(qv), order, far, leaf, nv.dptr(pu), nv.dptr(pv),
                                       C.byref(L))
    assert rc == 0
    return pu, pv, L.value


def test_oracle_fmm_converges_with_order():
    """The CPU oracle's FMM (the CPU baseline's demag: a classic octree FMM with M2L on every level
    and L2P) against the exact direct sums: the error falls with the order, by at least 3x per two
    orders, and stays below 1e-4 of the largest potential at order 8 (5e-4 for the random v charges)."""
    m = film()
    p, _, _ = charged(m)
    xyz, qu, qv = p.sources()
    cu, cv = p.corrections()
    _, _, phi, phiv = p.get_state(nv.FG_NEXT)
    tgt = m.arrays()["pos"]
    exact = (phi * 4 * np.pi - cu, phiv * 4 * np.pi - cv)
    errs = []
    for order in (4, 6, 8):
        pu, pv, L = oracle.fmm_potentials(xyz, qu, qv, tgt, order, 64)
        assert L >= 3
        errs.append([np.abs(a - b).max() / np.abs(b).max() for a, b in ((pu, exact[0]), (pv, exact[1]))])
    for k in range(2):
        assert errs[1][k] * 3 < errs[0][k] and errs[2][k] * 3 < errs[1][k], errs
    assert errs[2][0] < 1e-4 and errs[2][1] < 5e-4, errs
    # the problem-level switch: orc_calc_demag with the FMM against the exact potentials
    u, v, _, _ = p.get_state(nv.FG_NEXT)
    q = oracle.Problem(m.mesh)
    q.set_demag(8, 64)
    q.set_state(nv.FG_BOTH, u=u, v=v)
    q.calc_demag()
    _, _, fphi, fphiv = q.get_state(nv.FG_NEXT)
    assert np.abs(fphi - phi).max() <= 2e-4 * np.abs(phi).max()
    assert np.abs(fphiv - phiv).max() <= 1e-3 * np.abs(phiv).max()


def test_host_fmm_converges_with_order():
    m = film()
    p, _, _ = charged(m)
    xyz, qu, qv = p.sources()
    cu, cv = p.corrections()
    _, _, phi, phiv = p.get_state(nv.FG_NEXT)
    tgt = m.arrays()["pos"]
    exact = (phi * 4 * np.pi - cu, phiv * 4 * np.pi - cv)
    errs = []
    for order in (4, 6, 8):
        pu, pv, L = fmm_ref(xyz, qu, qv, tgt, order, 64)
        assert L >= 3
        e = max(np.abs(pu - exact[0]).max() / np.abs(exact[0]).max(),
                np.abs(pv - exact[1]).max() / np.abs(exact[1]).max())
        errs.append(e)
    assert errs[0] < 1e-2 and errs[1] < 3e-3 and errs[2] < 6e-4
    assert errs[2] < errs[1] < errs[0]


def test_host_fmm_far_degree():
    """Source cells three widths away evaluated at the lower far degree: the error stays within
    1.15x of the one-degree FMM at every order, and other far degrees are refused."""
    m = film()
    p, _, _ = charged(m)
    xyz, qu, qv = p.sources()
    cu, cv = p.corrections()
    _, _, phi, phiv = p.get_state(nv.FG_NEXT)
    tgt = m.arrays()["pos"]
    exact = phi * 4 * np.pi - cu
    for order in (6, 8, 10):
        one, _, _ = fmm_ref(xyz, qu, qv, tgt, order, 64, far=order)
        low, _, _ = fmm_ref(xyz, qu, qv, tgt, order, 64)
        low2, _, _ = fmm_ref(xyz, qu, qv, tgt, order, 64, far={6: 4, 8: 5, 10: 6}[order])
        assert np.array_equal(low, low2)
        e1 = np.abs(one - exact).max() / np.abs(exact).max()
        e2 = np.abs(low - exact).max() / np.abs(exact).max()
        assert e2 <= 1.15 * e1, (order, e1, e2)
        assert not np.array_equal(one, low)
    with pytest.raises(AssertionError):
        fmm_ref(xyz, qu, qv, tgt, 8, 64, far=6)


def test_host_fmm_point_charges():
    """Random point charges in a cube: FMM vs brute force at orders 4..10 (exact tail bound)."""
    rng = np.random.default_rng(5)
    src = rng.uniform(-1, 1, size=(4000, 3))
    tgt = rng.uniform(-1, 1, size=(300, 3))
    qu = rng.normal(size=4000)
    qv = rng.uniform(0, 1, size=4000)
    r = np.linalg.norm(tgt[:, None, :] - src[None, :, :], axis=2)
    ex_u = (qu / r).sum(axis=1)
    ex_v = (qv / r).sum(axis=1)
    prev = 1.0
    for order in (4, 6, 8, 10):
        pu, pv, L = fmm_ref(src, qu, qv, tgt, order, 40)
        e = max(np.abs(pu - ex_u).max() / np.abs(ex_u).max(), np.abs(pv - ex_v).max() / np.abs(ex_v).max())
        assert e < prev
        prev = e
    assert prev < 1e-5


@pytest.mark.gpu
def test_device_fmm_matches_host_fmm_and_direct():
    from feellgood_amd.engine import Engine
    m = film()
    p, u, v = charged(m)
    xyz, qu, qv = p.sources()
    cu, cv = p.corrections()
    _, _, phi0, phiv0 = p.get_state(nv.FG_NEXT)
    tgt = m.arrays()["pos"]
    # FP64 default: one degree; the lower far degree on request (both loops of the far field)
    for order, leaf, far in ((8, 64, 0), (8, 64, 5), (6, 128, 0), (10, 32, 0), (10, 32, 6)):
        g = Engine(m.mesh, demag="fmm", fmm_order=order, fmm_leaf=leaf, fmm_far_order=far)
        g.set_state(nv.FG_BOTH, u=u, v=v)
        g.calc_demag()
        _, _, phi1, phiv1 = g.get_state(nv.FG_NEXT)
        pu, pv, _ = fmm_ref(xyz, qu, qv, tgt, order, leaf, far=far or order)
        ref_u = (pu + cu) / (4 * np.pi)
        ref_v = (pv + cv) / (4 * np.pi)
        assert np.abs(phi1 - ref_u).max() <= 1e-10 * np.abs(ref_u).max(), (order, leaf, far)
        assert np.abs(phiv1 - ref_v).max() <= 1e-10 * np.abs(ref_v).max(), (order, leaf, far)
        err = np.abs(phi1 - phi0).max() / np.abs(phi0).max()
        assert err < {6: 3e-3, 8: 6e-4, 10: 2e-4}[order], (order, err)
        g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("order", [6, 8, 10])
def test_device_fmm_with_many_parents_per_level(order):
    """A 48x24x3-cell film with leaves of at most 8 sources on average: 57,350 leaves under 8,192
    level-6 parents, so the upward pass takes the all-row-tile M2M (levels of 2,048 parents or
    more; in FP32 with its input reals in pair order). FP64 device FMM against the host FMM within
    1e-10 of the largest potential; the mixed-precision pass (FP32 translations, lower far and v
    degrees) within 1e-5 (phi) and 3e-3 (phi_v) of the FP64 one at order 8, 1e-4 and 1e-2 at
    orders 6 and 10 (far entries to degrees 4 and 6; measured 1.7e-5 / 1.9e-3 and 5.6e-7 /
    1.8e-3)."""
    from feellgood_amd.engine import Engine
    m = film(48, 24, 3)
    p = oracle.Problem(m.mesh)
    p.set_demag(4, 64)  # the charges do not depend on how the potential is summed
    rng = np.random.default_rng(7)
    u = rng.normal(size=(m.n_nodes, 3)) * 0.3 + [1, 0.5, 0]
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    v = rng.normal(size=(m.n_nodes, 3)) * 1e9
    p.set_state(nv.FG_BOTH, u=u, v=v)
    p.calc_demag()
    xyz, qu, qv = p.sources()
    cu, cv = p.corrections()
    tgt = m.arrays()["pos"]
    pots = {}
    for prec in ("fp64", "mixed"):
        g = Engine(m.mesh, demag="fmm", fmm_order=order, fmm_leaf=8, precision=prec)
        info = g.fmm_info()
        assert info["leaf_level"] == 7 and info["m2m_cells"] >= 8192, info
        g.set_state(nv.FG_BOTH, u=u, v=v)
        g.calc_demag()
        pots[prec] = g.get_state(nv.FG_NEXT)[2:]
        g.close()
    pu, pv, L = fmm_ref(xyz, qu, qv, tgt, order, 8, far=order)
    assert L == 7
    for k, ref in enumerate(((pu + cu) / (4 * np.pi), (pv + cv) / (4 * np.pi))):
        assert np.abs(pots["fp64"][k] - ref).max() <= 1e-10 * np.abs(ref).max(), k
    err = [np.abs(pots["mixed"][k] - pots["fp64"][k]).max() / np.abs(pots["fp64"][k]).max() for k in range(2)]
    print(f"order {order}: mixed against FP64 FMM: phi {err[0]:.2e}, phi_v {err[1]:.2e} of the largest", flush=True)
    bound = {8: (1e-5, 3e-3), 6: (1e-4, 1e-2), 10: (1e-4, 1e-2)}[order]
    assert err[0] <= bound[0] and err[1] <= bound[1], err


@pytest.mark.gpu
@pytest.mark.parametrize("order", [8, 10])
def test_step_with_fmm_against_exact_oracle(order):
    """One LLG step from the same state, FMM demag vs the oracle's exact demag: |dm| <= 1e-6."""
    from feellgood_amd.engine import Engine
    m = film(24, 12, 3)
    _, u, _ = charged(m)
    v = np.zeros_like(u)
    H = np.array([-0.025, 0.0044, 0.0]) / nv.MU0
    dt = 5e-14
    pref = 1 + 0.01 * abs(np.log(0.01))
    p = oracle.Problem(m.mesh, tol=1e-10, max_iter=1000)
    g = Engine(m.mesh, tol=1e-10, max_iter=1000, demag="fmm", fmm_order=order, fmm_leaf=64)
    for o in (p, g):
        o.set_state(nv.FG_BOTH, u=u, v=v)
        o.calc_demag()
        o.evolution()
        o.base_projection(0.3)
        o.prepare_elements(H, dt, pref)
        o.solve(dt)
    dm = np.abs(g.get_state(nv.FG_NEXT)[0] - p.get_state(nv.FG_NEXT)[0]).max()
    assert dm <= 1e-6, dm


@pytest.mark.gpu
def test_mixed_precision_fmm_step():
    """The mixed-precision FMM (FP32 expansions, translations and near-field coefficients; the far
    field of the charges of v to degree 4): phi within 1e-5 relative of the FP64 FMM, phi_v within
    3e-3 (it enters the step only as theta dt H_d(v)), and one LLG step from a state with a nonzero
    previous v within |dm| <= 1e-6 of the exact FP64 oracle and within 1e-7 of the FP64 FMM step."""
    from feellgood_amd.engine import Engine
    m = film(24, 12, 3)
    p, u, v = charged(m)
    pots = {}
    for prec in ("fp64", "mixed"):
        g = Engine(m.mesh, demag="fmm", fmm_order=8, fmm_leaf=64, precision=prec)
        g.set_state(nv.FG_BOTH, u=u, v=v)
        g.calc_demag()
        pots[prec] = g.get_state(nv.FG_NEXT)[2:]
        g.close()
    err = [np.abs(pots["mixed"][k] - pots["fp64"][k]).max() / np.abs(pots["fp64"][k]).max() for k in range(2)]
    print(f"mixed against FP64 FMM: phi {err[0]:.2e}, phi_v {err[1]:.2e} of the largest", flush=True)
    assert err[0] <= 1e-5 and err[1] <= 3e-3, err
    v0 = v  # random, so its charges are as large as those of the film's smooth 1e11 rad/s velocities
    H = np.array([-0.025, 0.0044, 0.0]) / nv.MU0
    dt = 5e-14
    pref = 1 + 0.01 * abs(np.log(0.01))
    o = oracle.Problem(m.mesh, tol=1e-10, max_iter=1000)
    f = Engine(m.mesh, tol=1e-10, max_iter=1000, demag="fmm", fmm_order=8, fmm_leaf=64)
    g = Engine(m.mesh, tol=1e-10, max_iter=1000, demag="fmm", fmm_order=8, fmm_leaf=64, precision="mixed")
    for e in (o, f, g):
        e.set_state(nv.FG_BOTH, u=u, v=v0)
        e.calc_demag()
        e.evolution()
        e.base_projection(0.3)
        e.prepare_elements(H, dt, pref)
        e.solve(dt)
    un = {k: e.get_state(nv.FG_NEXT)[0] for k, e in (("o", o), ("f", f), ("g", g))}
    dm, dmf = np.abs(un["g"] - un["o"]).max(), np.abs(un["g"] - un["f"]).max()
    print(f"step |dm|: mixed against exact {dm:.2e}, against FP64 FMM {dmf:.2e}", flush=True)
    assert dm <= 1e-6 and dmf <= 1e-7, (dm, dmf)
