"""GPU parity: the gfx950 path (libfgmi.so through the C-ABI) against the FP64 CPU oracle on the
same meshes, states and basis angles.

Tolerances (FP64 both sides, differing only in summation order):
  * basis vectors, charges, potentials, element-gathered matrix entries: 1e-12 relative to the
    largest entry of the array compared;
  * BiCGSTAB solutions: both solved to tol 1e-12 -> velocities agree to 1e-8 relative;
  * magnetisation after a step: |dm| <= 1e-6, in practice ~1e-13;
  * energies: 1e-8 relative.
"""
import os

import numpy as np
import pytest

import oracle
from feellgood_amd import native as nv
from feellgood_amd.mesh import HostMesh

pytestmark = pytest.mark.gpu

MU0 = nv.MU0


def engine(mesh, **kw):
    from feellgood_amd.engine import Engine
    return Engine(mesh, **kw)


def cube(n=5, L=25.0, jitter=0.15, seed=1, **prm):
    m = HostMesh.cuboid((L, L, L), (n, n, n), jitter=jitter, seed=seed)
    p = dict(alpha=0.3, Ae=1.3e-11, Ms=8e5, K=4e4, uk=(1, 2, 0.5), K3=-1e4)
    p.update(prm)
    m.prepare({"volume": p})
    return m


def two_regions():
    """A cuboid whose right half is a non-magnetic region: masking and interface charges."""
    import tempfile
    m = HostMesh.cuboid((30.0, 15.0, 15.0), (6, 3, 3), jitter=0.1, seed=4)
    d = tempfile.mkdtemp()
    path = os.path.join(d, "two.msh")
    m.write_msh(path)
    lines = open(path).read().splitlines()
    nodes = {}
    i_n = lines.index("$Nodes")
    for k in range(int(lines[i_n + 1])):
        p = lines[i_n + 2 + k].split()
        nodes[int(p[0])] = float(p[1])
    i_e = lines.index("$Elements")
    for k in range(int(lines[i_e + 1])):
        p = lines[i_e + 2 + k].split()
        if p[1] == "4" and sum(nodes[int(v)] for v in p[5:9]) / 4 > 20:
            p[3] = "301"
        lines[i_e + 2 + k] = " ".join(p)
    i_p = lines.index("$PhysicalNames")
    lines[i_p + 1] = str(int(lines[i_p + 1]) + 1)
    lines.insert(i_p + 2, '3 301 "copper"')
    open(path, "w").write("\n".join(lines) + "\n")
    mm = HostMesh.read(path)
    mm.prepare({"volume": dict(alpha=0.2, Ae=1.1e-11, Ms=9e5), "copper": dict(Ms=0.0)})
    return mm


def state(m, seed=2, mag=None):
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(m.n_nodes, 3)) * 0.3 + np.array([1.0, 1.0, 0.2])
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    v = rng.normal(size=(m.n_nodes, 3)) * 1e9
    v -= np.sum(v * u, axis=1, keepdims=True) * u
    if mag is not None:
        u[~mag] = np.nan
        v[~mag] = 0
    return u, v


def magnetic_nodes(m):
    a = m.arrays()
    vol = m.region_names
    mag = np.zeros(m.n_nodes, bool)
    for t, r in zip(a["tets"], a["tet_region"]):
        if vol[r] != "copper":
            mag[t] = True
    return mag


def pair(m, u, v, tol=1e-12, max_iter=2000):
    """Oracle and engine with the same CURRENT/NEXT state (demag done, evolution done)."""
    p = oracle.Problem(m.mesh, tol=tol, max_iter=max_iter)
    g = engine(m.mesh, tol=tol, max_iter=max_iter)
    for o in (p, g):
        o.set_state(nv.FG_BOTH, u=u, v=v)
        o.calc_demag()
        o.evolution()
    return p, g


def rel(a, b):
    a = np.nan_to_num(np.asarray(a, dtype=np.float64))
    b = np.nan_to_num(np.asarray(b, dtype=np.float64))
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def small():
    return cube()


def test_basis(small):
    u, v = state(small)
    p, g = pair(small, u, v)
    for r in (0.0, 0.37, 0.61):
        p.base_projection(r)
        g.base_projection(r)
        ep0, eq0 = p.basis()
        ep1, eq1 = g.basis()
        assert rel(ep1, ep0) < 1e-14 and rel(eq1, eq0) < 1e-14


def test_demag_direct(small):
    u, v = state(small)
    p, g = pair(small, u, v)
    _, _, ph0, phv0 = p.get_state(nv.FG_NEXT)
    _, _, ph1, phv1 = g.get_state(nv.FG_NEXT)
    assert rel(ph1, ph0) < 1e-12 and rel(phv1, phv0) < 1e-12


def test_system_and_rhs(small):
    u, v = state(small)
    p, g = pair(small, u, v)
    H = np.array([3e4, -2e4, 1e4])
    dt, pref = 2e-14, 1.046052
    for o in (p, g):
        o.base_projection(0.42)
        o.prepare_elements(H, dt, pref)
        o.solve(dt)
    rp0, c0, v0, b0 = p.system()
    rp1, c1, v1, b1 = g.system()
    assert np.array_equal(rp0, rp1) and np.array_equal(c0, c1)
    assert rel(v1, v0) < 1e-12
    assert rel(b1, b0) < 1e-12


def test_solve_and_update(small):
    u, v = state(small)
    p, g = pair(small, u, v)
    H = np.array([3e4, -2e4, 1e4])
    dt, pref = 2e-14, 1.046052
    out = []
    for o in (p, g):
        o.base_projection(0.13)
        o.prepare_elements(H, dt, pref)
        failed = o.solve(dt)
        out.append(failed)
    (rc0, i0), (f1, i1) = out
    assert rc0 == 0 and not f1
    assert i1.status == 0 and abs(i1.iterations - i0.iterations) <= 1
    assert i1.rhs_norm == pytest.approx(i0.rhs_norm, rel=1e-12, abs=0)
    un0, vn0, _, _ = p.get_state(nv.FG_NEXT)
    un1, vn1, _, _ = g.get_state(nv.FG_NEXT)
    assert rel(vn1, vn0) < 1e-8
    assert np.abs(un1 - un0).max() <= 1e-6  # per-step |dm| bound
    assert np.abs(un1 - un0).max() <= 1e-12
    assert i1.v_max == pytest.approx(i0.v_max, rel=1e-8, abs=0)


def test_space_field_step_and_energy(small):
    """The space-dependent applied field (R4toR3, src/linear_algebra.cpp:54-75, src/energy.cpp:11-16):
    a field varying over the Gauss points of fg_gauss_points, scaled by an amplitude, in the
    element pass, one solve and the Zeeman energy, engine against oracle (the copper region
    checks the mesh-to-magnetic-tet mapping of fg_set_space_field)."""
    import ctypes as C
    u, v = state(small)
    p, g = pair(small, u, v)
    T = small.n_tets
    pg = np.zeros((T, 5, 3))
    assert nv.hip_lib().fg_gauss_points(C.byref(small.mesh), nv.dptr(pg)) == nv.FG_OK
    a = small.arrays()
    cen = a["pos"][a["tets"]].mean(axis=1)
    assert np.allclose((pg * np.array([-2 / 15] + [3 / 40] * 4)[None, :, None]).sum(axis=1) / (1 / 6), cen,
                       rtol=0, atol=1e-20)  # the 5-point rule integrates x exactly
    L = np.ptp(a["pos"], axis=0).max()
    h = np.stack([3e4 * np.sin(pg[..., 0] / L * 3), -2e4 * np.cos(pg[..., 1] / L * 2),
                  1e4 + 5e3 * pg[..., 2] / L], axis=-1)
    dt, pref = 2e-14, 1.046052
    for o in (p, g):
        o.set_space_field(h)
        o.base_projection(0.42)
        o.prepare_elements_amp(0.7, dt, pref)
        o.solve(dt)
    rp0, c0, v0, b0 = p.system()
    rp1, c1, v1, b1 = g.system()
    assert rel(v1, v0) < 1e-12 and rel(b1, b0) < 1e-12
    un0, vn0, _, _ = p.get_state(nv.FG_NEXT)
    un1, vn1, _, _ = g.get_state(nv.FG_NEXT)
    assert rel(vn1, vn0) < 1e-8 and np.abs(un1 - un0).max() <= 1e-12
    for o in (p, g):
        o.evolution()
    E0, E1 = p.energy_amp(-1.3), g.energy_amp(-1.3)
    assert np.allclose(E1, E0, rtol=1e-8, atol=1e-8 * np.abs(E0).max()) and E0[3] != 0
    g.set_space_field(None)
    with pytest.raises(RuntimeError):
        g.prepare_elements_amp(0.7, dt, pref)


def test_spin_field_step(small):
    """The spin-accumulation field Hst of the element pass (src/tetra.cpp:93-106,248-255,296-302)
    from a non-uniform node spin accumulation, with v != 0 so that its second-order term is active:
    element pass and one solve, engine against oracle; removing it restores the plain system."""
    u, v = state(small)
    p, g = pair(small, u, v)
    rng = np.random.default_rng(11)
    sacc = rng.normal(size=(small.n_nodes, 3)) * 2e5 + [1e5, 0.0, -3e5]
    vols = [nv.SpinVolume(5e6, 1e47, 0.4, 4e-9, 10e-9)] * small.mesh.n_vol_regions
    H = np.array([1e4, 5e3, -2e4])
    dt, pref = 2e-14, 1.046052
    for o in (p, g):
        o.set_spin_accumulation(sacc, vols)
        o.base_projection(0.42)
        o.prepare_elements(H, dt, pref)
        o.solve(dt)
    rp0, c0, v0, b0 = p.system()
    rp1, c1, v1, b1 = g.system()
    assert rel(v1, v0) < 1e-12 and rel(b1, b0) < 1e-12
    un0, vn0, _, _ = p.get_state(nv.FG_NEXT)
    un1, vn1, _, _ = g.get_state(nv.FG_NEXT)
    assert rel(vn1, vn0) < 1e-8 and np.abs(un1 - un0).max() <= 1e-12
    g.set_spin_accumulation(None)
    g.prepare_elements(H, dt, pref)
    g.solve(dt)
    assert rel(g.system()[3], b0) > 1e-9  # the field changed the right-hand side


def test_surface_anisotropy_step_and_energy():
    """Surface (Neel) anisotropy Ks != 0 on the boundary triangles (Tri::integrales and
    Tri::anisotropyEnergy, src/triangle.cpp:6-43) together with uniaxial K and cubic K3 in the
    volume (src/tetra.cpp:171-208): element pass, assembled system and right-hand side, one solve
    and the energies, engine against oracle. Ks changes the right-hand side and the anisotropy
    energy, so the comparison covers the surface term."""
    m = HostMesh.cuboid((25.0, 20.0, 15.0), (5, 4, 3), jitter=0.15, seed=6)
    prm = dict(alpha=0.3, Ae=1.3e-11, Ms=8e5, K=4e4, uk=(1, 2, 0.5), K3=-1e4)
    m.prepare({"volume": prm}, {"surface": dict(Ks=2e-3, uk=(0.3, -0.2, 1.0))})
    u, v = state(m, seed=8)
    p, g = pair(m, u, v)
    H = np.array([2e4, -1e4, 3e4])
    dt, pref = 2e-14, 1.046052
    for o in (p, g):
        o.base_projection(0.29)
        o.prepare_elements(H, dt, pref)
        o.solve(dt)
    rp0, c0, v0, b0 = p.system()
    rp1, c1, v1, b1 = g.system()
    assert np.array_equal(rp0, rp1) and np.array_equal(c0, c1)
    assert rel(v1, v0) < 1e-12 and rel(b1, b0) < 1e-12
    un0, vn0, _, _ = p.get_state(nv.FG_NEXT)
    un1, vn1, _, _ = g.get_state(nv.FG_NEXT)
    assert rel(vn1, vn0) < 1e-8 and np.abs(un1 - un0).max() <= 1e-12
    for o in (p, g):
        o.calc_demag()
    E0, E1 = p.energy(H), g.energy(H)
    assert np.allclose(E1, E0, rtol=1e-8, atol=1e-8 * np.abs(E0).max())
    m0 = HostMesh.cuboid((25.0, 20.0, 15.0), (5, 4, 3), jitter=0.15, seed=6)
    m0.prepare({"volume": prm})
    p0, _ = pair(m0, u, v)
    p0.base_projection(0.29)
    p0.prepare_elements(H, dt, pref)
    p0.solve(dt)
    assert rel(p0.system()[3], b0) > 1e-9  # the surface term is 1e5 times the 1e-12 tolerance above
    p0.calc_demag()
    assert abs(p0.energy(H)[1] - E0[1]) > 1e-3 * abs(E0[1])


def test_energy_average_max_angle(small):
    u, v = state(small)
    p, g = pair(small, u, v)
    H = np.array([1e4, 5e3, -2e4])
    # NEXT == CURRENT after the evolution of pair(); the energies read NEXT
    E0 = p.energy(H)
    E1 = g.energy(H)
    assert np.allclose(E1, E0, rtol=1e-8, atol=1e-8 * np.abs(E0).max())
    assert np.allclose(g.average(0), p.average(0), rtol=1e-12, atol=1e-14)
    assert np.allclose(g.average(1), p.average(1), rtol=1e-12, atol=1e-14 * np.abs(v).max())
    assert g.max_angle() == pytest.approx(p.max_angle(), rel=1e-12, abs=0)


def test_trajectory_20_steps():
    """Lock-step run at fixed dt: |dm| stays below 1e-6 at every step, energies within 1e-8."""
    m = cube(n=4, alpha=0.5, K=0.0, K3=0.0)
    u, v = state(m, seed=5)
    v[:] = 0
    p, g = pair(m, u, v, tol=1e-10)
    H = np.array([0.0, 0.0, 2e4])
    dt = 1e-14
    pref = 1 + (dt / (100 * dt)) * abs(np.log(dt / (100 * dt)))
    angles = oracle.basis_angles(1, 20)
    for k in range(20):
        res = []
        for o in (p, g):
            o.base_projection(angles[k])
            o.prepare_elements(H, dt, pref)
            res.append(o.solve(dt))
            o.calc_demag()
        E0 = p.energy(H)
        E1 = g.energy(H)
        for o in (p, g):
            o.evolution()
        u0 = p.get_state(nv.FG_CURRENT)[0]
        u1 = g.get_state(nv.FG_CURRENT)[0]
        assert np.abs(u1 - u0).max() <= 1e-6, k
        assert np.allclose(E1, E0, rtol=1e-8, atol=1e-8 * np.abs(E0).max()), k


def test_nonmagnetic_region_masking():
    m = two_regions()
    mag = magnetic_nodes(m)
    assert 0 < mag.sum() < m.n_nodes
    u, v = state(m, mag=mag)
    p, g = pair(m, u, v)
    H = np.array([1e4, 0, 0])
    dt, pref = 1e-14, 1.046052
    for o in (p, g):
        o.base_projection(0.2)
        o.prepare_elements(H, dt, pref)
        o.solve(dt)
    rp0, c0, v0, b0 = p.system()
    rp1, c1, v1, b1 = g.system()
    assert np.array_equal(c0, c1) and rel(v1, v0) < 1e-12 and rel(b1, b0) < 1e-12
    un0 = p.get_state(nv.FG_NEXT)[0][mag]
    un1 = g.get_state(nv.FG_NEXT)[0][mag]
    assert np.abs(un1 - un0).max() <= 1e-12
    for o in (p, g):
        o.calc_demag()
    ph0 = p.get_state(nv.FG_NEXT)[2]
    ph1 = g.get_state(nv.FG_NEXT)[2]
    assert rel(ph1[mag], ph0[mag]) < 1e-12


def test_solver_failure_reports_overflow(small):
    """max(iter) = 1 makes every solve fail with ITER_OVERFLOW (src/algebra/iter.h:119-125);
    the NEXT state is left untouched (src/solver.cpp:62-69)."""
    u, v = state(small)
    g = engine(small.mesh, tol=1e-14, max_iter=1)
    g.set_state(nv.FG_BOTH, u=u, v=v)
    before = g.get_state(nv.FG_NEXT)[0]
    g.base_projection(0.3)
    g.prepare_elements(np.array([1e4, 0, 0]), 1e-14, 1.0)
    failed, info = g.solve(1e-14)
    assert failed and info.status == 1 and info.iterations == 1
    assert np.array_equal(g.get_state(nv.FG_NEXT)[0], before)


def test_bitwise_deterministic(small):
    u, v = state(small)
    outs = []
    for _ in range(2):
        g = engine(small.mesh, tol=1e-10)
        g.set_state(nv.FG_BOTH, u=u, v=v)
        g.calc_demag()
        g.evolution()
        g.base_projection(0.5)
        g.prepare_elements(np.array([1e4, 2e4, 0]), 1e-14, 1.04)
        g.solve(1e-14)
        outs.append((g.solution(), g.get_state(nv.FG_NEXT)[0]))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("demag", ["direct", "fmm"])
def test_step_entry_point_matches_the_separate_calls(small, demag):
    """fg_step, the loop body in one call, against base_projection, prepare_elements, solve,
    calc_demag, energy, evolution and max_angle called one by one on a second engine: the same
    states to the bit, the same energies, max angle and iteration counts over three steps. A
    dumax below dt v_max rejects the step: nothing evolves (CURRENT unchanged, NEXT holds the
    solution), as in the reference's loop (src/time_integration.cpp:215-222)."""
    u, v = state(small)
    H = np.array([1e4, 2e4, 0.0])
    dt, pref = 1e-14, 1.04
    kw = dict(tol=1e-10, demag=demag, fmm_order=8, fmm_leaf=16)
    a, b = engine(small.mesh, **kw), engine(small.mesh, **kw)
    for g in (a, b):
        g.set_state(nv.FG_BOTH, u=u, v=v)
        g.calc_demag()
        g.evolution()
    for r in (0.5, 0.2, 0.9):
        rc, info, E, ang = a.step(r, H, dt, pref)
        assert rc == nv.FG_OK
        b.base_projection(r)
        b.prepare_elements(H, dt, pref)
        failed, info_b = b.solve(dt)
        assert not failed and info.iterations == info_b.iterations and info.v_max == info_b.v_max
        b.calc_demag()
        E_b = b.energy(H)
        b.evolution()
        ang_b = b.max_angle()
        assert np.array_equal(E, E_b) and ang == ang_b
        for which in (nv.FG_CURRENT, nv.FG_NEXT):
            for x, y in zip(a.get_state(which), b.get_state(which)):
                assert np.array_equal(x, y)
    cur = a.get_state(nv.FG_CURRENT)[0]
    rc, info, E, ang = a.step(0.3, H, dt, pref, dumax=0.5 * dt * info.v_max)
    assert rc == nv.FG_STEP_REJECTED and E is None and info.v_max > 0
    assert np.array_equal(a.get_state(nv.FG_CURRENT)[0], cur)
    assert not np.array_equal(a.get_state(nv.FG_NEXT)[0], cur)
    a.close()
    b.close()
    # a solve that cannot converge in its iterations: FG_SOLVE_FAILED with the reference's
    # ITER_OVERFLOW status, CURRENT and NEXT u untouched by the gated update and nothing evolved
    f = engine(small.mesh, tol=1e-30, max_iter=2, demag=demag, fmm_order=8, fmm_leaf=16)
    f.set_state(nv.FG_BOTH, u=u, v=v)
    f.calc_demag()
    f.evolution()
    before = [x.copy() for x in f.get_state(nv.FG_CURRENT)]
    nxt_u = f.get_state(nv.FG_NEXT)[0]
    rc, info, E, ang = f.step(0.5, H, dt, pref)
    assert rc == nv.FG_SOLVE_FAILED and E is None and info.status == nv.FG_ITER_OVERFLOW, (rc, info.status)
    for x, y in zip(f.get_state(nv.FG_CURRENT), before):
        assert np.array_equal(x, y)
    assert np.array_equal(f.get_state(nv.FG_NEXT)[0], nxt_u)
    f.close()


def test_ellipsoid_fixture_step(golden_dir):
    """One step of the full_test.py settings on examples/ellipsoid.msh, engine vs oracle."""
    m = HostMesh.read(os.path.join(golden_dir, "ellipsoid.msh"))
    m.prepare({"ellipsoid_volume": dict(K=3e5, uk=(0, 1, 0))})
    u = np.tile([0.0, 0.0, 1.0], (m.n_nodes, 1))
    p, g = pair(m, u, np.zeros_like(u), tol=1e-6, max_iter=700)
    H = np.array([1.0, 0.0, -1.0]) / MU0
    dt = np.sqrt(5e-14 * 1e-12)
    pref = 1 + (dt / 1e-10) * abs(np.log(dt / 1e-10))
    for o in (p, g):
        o.base_projection(0.25)
        o.prepare_elements(H, dt, pref)
        o.solve(dt)
    un0 = p.get_state(nv.FG_NEXT)[0]
    un1 = g.get_state(nv.FG_NEXT)[0]
    assert np.abs(un1 - un0).max() <= 1e-6


def test_mixed_krylov_solve_and_trajectory():
    """FG_KRYLOV_MIXED (FP32 bicg_dir inside FP64 refinement) on the 5-cell cube: one solve against
    the oracle's FP64 bicg_dir on the same system (both to tol 1e-12: v within 1e-8, u within 1e-12),
    the accepted solution's FP64 residual recomputed from the assembled system at most tol ||b||,
    the refinement counts reported, and a 20-step lock-step run with |dm| <= 1e-6 and energies
    within 1e-8 of the oracle."""
    from feellgood_amd.engine import Engine
    m = cube()
    u, v = state(m)
    p = oracle.Problem(m.mesh, tol=1e-12, max_iter=2000)
    g = Engine(m.mesh, tol=1e-12, max_iter=2000, krylov="mixed")
    for o in (p, g):
        o.set_state(nv.FG_BOTH, u=u, v=v)
        o.calc_demag()
        o.evolution()
    H = np.array([3e4, -2e4, 1e4])
    dt, pref = 2e-14, 1.046052
    for o in (p, g):
        o.base_projection(0.13)
        o.prepare_elements(H, dt, pref)
    rc0, i0 = p.solve(dt)
    f1, i1 = g.solve(dt)
    assert rc0 == 0 and not f1 and i1.status == nv.FG_ITER_CONVERGED
    assert i1.refinements >= 1 and i1.inner_iterations == i1.iterations > 0
    print(f"mixed Krylov: {i1.refinements} corrections, {i1.inner_iterations} FP32 iterations "
          f"(FP64 bicg_dir: {i0.iterations})", flush=True)
    rp, col, val, b = g.system()
    x = g.solution()
    Ax = np.array([val[rp[i]:rp[i + 1]] @ x[col[rp[i]:rp[i + 1]]] for i in range(len(b))])
    assert np.linalg.norm(b - Ax) <= 1e-12 * np.linalg.norm(b) * 1.01
    un0, vn0, _, _ = p.get_state(nv.FG_NEXT)
    un1, vn1, _, _ = g.get_state(nv.FG_NEXT)
    assert rel(vn1, vn0) < 1e-8 and np.abs(un1 - un0).max() <= 1e-12
    g.close()

    m = cube(n=4, alpha=0.5, K=0.0, K3=0.0)
    u, v = state(m, seed=5)
    v[:] = 0
    p = oracle.Problem(m.mesh, tol=1e-10, max_iter=2000)
    g = Engine(m.mesh, tol=1e-10, max_iter=2000, krylov="mixed")
    for o in (p, g):
        o.set_state(nv.FG_BOTH, u=u, v=v)
        o.calc_demag()
        o.evolution()
    H = np.array([0.0, 0.0, 2e4])
    dt = 1e-14
    pref = 1 + 0.01 * abs(np.log(0.01))
    angles = oracle.basis_angles(1, 20)
    for k in range(20):
        for o in (p, g):
            o.base_projection(angles[k])
            o.prepare_elements(H, dt, pref)
            o.solve(dt)
            o.calc_demag()
        E0 = p.energy(H)
        E1 = g.energy(H)
        for o in (p, g):
            o.evolution()
        u0 = p.get_state(nv.FG_CURRENT)[0]
        u1 = g.get_state(nv.FG_CURRENT)[0]
        assert np.abs(u1 - u0).max() <= 1e-6, k
        assert np.allclose(E1, E0, rtol=1e-8, atol=1e-8 * np.abs(E0).max()), k
    g.close()
