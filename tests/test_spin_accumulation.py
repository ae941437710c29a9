"""The spin accumulation of the spin-diffusion problem (spinAcc, src/spinAccumulationSolver.cpp:
41-287, algebra::bicg_dir with Dirichlet values, src/algebra/bicg.h:82-158).

The oracle is pinned by closed forms on a bar (x from x0 to x1, Dirichlet values on both ends):
the diffusion alone (s = [s0 sinh((x1 - x) / lsf) + s1 sinh((x - x0) / lsf)] / sinh(L / lsf),
with s1 = -jn mu_B / e uP from the current-density surface); the s-d exchange in a magnetic bar
with u along z, where w = s_x + i s_y obeys w'' = (1 / lsf^2 - i / lsd^2) w; and the source of a
uniform potential gradient entering at the interface of a non-magnetic and a magnetic half, where
the flux D0 s_z' jumps by cst0 dV/dx. P1 elements with the lumped spin-flip term converge to these
at second order, so the tests check the error and its decay under refinement. The device path is
checked against the oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle
from feellgood_amd import native as nv
from test_electrostatic import bar

E = 1.602176634e-19
HBAR = 6.62607015e-34 / (2 * np.pi)
ME = 9.1093837139e-31
MUB_E = HBAR / (2 * ME)  # mu_B / e (src/config.h.in:33-37)
NAN = float("nan")


def vol(sigma=5e6, N0=1e47, P=0.4, lsd=5e-9, lsf=10e-9):
    return nv.SpinVolume(sigma, N0, P, lsd, lsf)


def surf(jn=NAN, uP=(NAN, NAN, NAN), s=(NAN, NAN, NAN)):
    return nv.SpinSurface(jn, (C.c_double * 3)(*uP), (C.c_double * 3)(*s))


S0 = (2e6, -1e6, 5e5)
JN, UP = 3e10, (0.0, 0.6, 0.8)
SURF = [surf(s=S0), surf(jn=JN, uP=UP), surf()]


def run(nx, ms=(0.0, 0.0), vols=(vol(), vol()), V=None, u=None, surfs=SURF, jitter=0.0):
    m, own, pos = bar(jitter, cells=(nx, 2, 2), ms=ms)
    if u is None:
        u = np.tile([0.0, 0.0, 1.0], (m.n_nodes, 1))
    if V is None:
        V = np.zeros(m.n_nodes)
    elif callable(V):
        V = V(pos)
    s, info = oracle.spin_accumulation(m, list(vols), list(surfs), V, u, tol=1e-12, max_iter=5000)
    assert info.status == nv.FG_ITER_CONVERGED
    return s, pos


def test_diffusion_closed_form():
    errs = []
    for nx in (16, 32):
        s, pos = run(nx)
        x = pos[:, 0] - pos[:, 0].min()
        L, lsf = np.ptp(pos[:, 0]), 10e-9
        s1 = -JN * MUB_E * np.array(UP)
        ref = (np.outer(np.sinh((L - x) / lsf), S0) + np.outer(np.sinh(x / lsf), s1)) / np.sinh(L / lsf)
        errs.append(np.abs(s - ref).max() / np.abs(ref).max())
    assert errs[1] < 2e-3 and errs[0] / errs[1] > 3.0, errs


def test_sd_exchange_closed_form():
    errs = []
    for nx in (16, 32):
        s, pos = run(nx, ms=(8e5, 8e5))
        x = pos[:, 0] - pos[:, 0].min()
        L, lsf, lsd = np.ptp(pos[:, 0]), 10e-9, 5e-9
        s1 = -JN * MUB_E * np.array(UP)
        k = np.sqrt(1 / lsf**2 - 1j / lsd**2)
        w0, w1 = S0[0] + 1j * S0[1], s1[0] + 1j * s1[1]
        w = (w0 * np.sinh(k * (L - x)) + w1 * np.sinh(k * x)) / np.sinh(k * L)
        z = (S0[2] * np.sinh((L - x) / lsf) + s1[2] * np.sinh(x / lsf)) / np.sinh(L / lsf)
        ref = np.stack([w.real, w.imag, z], axis=1)
        errs.append(np.abs(s - ref).max() / np.abs(ref).max())
    assert errs[1] < 5e-3 and errs[0] / errs[1] > 3.0, errs


def test_potential_source_closed_form():
    """Non-magnetic left half (D0a, lsf_a), magnetic right half (D0b, lsf_b, P) with u along z, s = 0
    on both ends, V = g (x - x0): s_z(x_m) = -cst0 g / (D0a coth(xm/la)/la + D0b coth((L-xm)/lb)/lb)."""
    va, vb = vol(sigma=4e6, N0=2e47, lsf=8e-9), vol(sigma=6e6, N0=1e47, P=0.5, lsf=12e-9)
    zero = [surf(s=(0.0, 0.0, 0.0)), surf(jn=0.0, uP=(0.0, 0.0, 1.0)), surf()]
    g = -3e6
    errs = []
    for nx in (16, 32):
        s, pos = run(nx, ms=(0.0, 8e5), vols=(va, vb), V=lambda p: g * (p[:, 0] - p[:, 0].min()), surfs=zero)
        x = pos[:, 0] - pos[:, 0].min()
        L = np.ptp(pos[:, 0])
        xm = L / 2
        d0a, d0b = 2 * va.sigma / (E**2 * va.N0), 2 * vb.sigma / (E**2 * vb.N0)
        la, lb = va.lsf, vb.lsf
        cst0 = MUB_E * E * vb.P * vb.sigma / E
        sm = -cst0 * g / (d0a / np.tanh(xm / la) / la + d0b / np.tanh((L - xm) / lb) / lb)
        ref = np.where(x <= xm, sm * np.sinh(x / la) / np.sinh(xm / la), sm * np.sinh((L - x) / lb) / np.sinh((L - xm) / lb))
        errs.append(np.abs(s[:, 2] - ref).max() / np.abs(ref).max())
        assert np.abs(s[:, :2]).max() <= 1e-9 * np.abs(ref).max()
    assert errs[1] < 5e-3 and errs[0] / errs[1] > 3.0, errs


def test_boundary_condition_errors():
    m, _own, _ = bar()
    u = np.tile([0.0, 0.0, 1.0], (m.n_nodes, 1))
    V = np.zeros(m.n_nodes)
    out = np.zeros(3 * m.n_nodes)
    info = nv.SolveInfo()
    for vols, surfs in (((vol(), vol(N0=0.0)), SURF), ((vol(), vol(P=1.0)), SURF),
                        ((vol(), vol()), [surf(s=S0), surf(s=S0), surf()]),
                        ((vol(), vol()), [surf(jn=1.0), surf(jn=JN, uP=UP), surf()])):
        with pytest.raises(ValueError):
            oracle.spin_accumulation(m, list(vols), list(surfs), V, u)
        va = (nv.SpinVolume * 2)(*vols)
        sa = (nv.SpinSurface * 3)(*surfs)
        rc = nv.hip_lib().fg_spin_accumulation(C.byref(m), va, sa, nv.dptr(V), nv.dptr(u), None, 1e-8, 1000, 0,
                                               nv.dptr(out), C.byref(info))
        assert rc == nv.FG_ERR_ARG


def device_spin(m, vols, surfs, V, u, tol=1e-8, max_iter=1000, u_src=None):
    va = (nv.SpinVolume * len(vols))(*vols)
    sa = (nv.SpinSurface * len(surfs))(*surfs)
    out = np.zeros(3 * m.n_nodes)
    info = nv.SolveInfo()
    rc = nv.hip_lib().fg_spin_accumulation(C.byref(m), va, sa, nv.dptr(np.ascontiguousarray(V, dtype=np.float64)),
                                           nv.dptr(np.ascontiguousarray(u, dtype=np.float64).reshape(-1)),
                                           None if u_src is None else nv.dptr(np.ascontiguousarray(u_src).reshape(-1)), tol,
                                           max_iter, 0, nv.dptr(out), C.byref(info))
    assert rc == nv.FG_OK
    return out.reshape(-1, 3), info


@pytest.mark.gpu
def test_device_spin_accumulation_matches_oracle():
    """Device assembly and bicg_dir against the oracle at the reference's tolerance 1e-8 and 1000
    iterations (src/main.cpp:252) on a jittered bar: a non-magnetic half and a magnetic half with a
    non-uniform magnetization and potential, both source and s-d exchange active."""
    m, _own, pos = bar(0.3, cells=(12, 4, 5), ms=(0.0, 8e5))
    rng = np.random.default_rng(5)
    u = rng.normal(size=(m.n_nodes, 3)) + [0.0, 0.0, 2.0]
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    x = (pos[:, 0] - pos[:, 0].min()) / np.ptp(pos[:, 0])
    V = -0.02 * x + 0.004 * x**2 + 1e-3 * np.sin(pos[:, 1] * 3e8)
    vols = [vol(sigma=4e6, N0=2e47, lsf=8e-9), vol(sigma=6e6, N0=1e47, P=0.5, lsf=12e-9, lsd=4e-9)]
    u2 = u + 0.2 * rng.normal(size=u.shape)
    u2 /= np.linalg.norm(u2, axis=1, keepdims=True)
    s0, i0 = oracle.spin_accumulation(m, vols, SURF, V, u, tol=1e-8, max_iter=1000, u_src=u2)
    s1, i1 = device_spin(m, vols, SURF, V, u, u_src=u2)
    assert i0.status == i1.status == nv.FG_ITER_CONVERGED
    assert abs(i1.iterations - i0.iterations) <= 2, (i0.iterations, i1.iterations)
    assert i1.rhs_norm == pytest.approx(i0.rhs_norm, rel=1e-10, abs=0)
    assert np.abs(s1 - s0).max() <= 1e-6 * np.abs(s0).max()


@pytest.mark.gpu
def test_device_spin_accumulation_closed_form():
    """The s-d exchange closed form (w'' = (1 / lsf^2 - i / lsd^2) w) on the device, and an
    iteration limit that is reached reported as such."""
    m, _own, pos = bar(0.0, cells=(32, 2, 2), ms=(8e5, 8e5))
    u = np.tile([0.0, 0.0, 1.0], (m.n_nodes, 1))
    s, info = device_spin(m, [vol(), vol()], SURF, np.zeros(m.n_nodes), u, tol=1e-12, max_iter=5000)
    assert info.status == nv.FG_ITER_CONVERGED
    ref, _ = oracle.spin_accumulation(m, [vol(), vol()], SURF, np.zeros(m.n_nodes), u, tol=1e-12, max_iter=5000)
    x = pos[:, 0] - pos[:, 0].min()
    L, lsf, lsd = np.ptp(pos[:, 0]), 10e-9, 5e-9
    s1 = -JN * MUB_E * np.array(UP)
    k = np.sqrt(1 / lsf**2 - 1j / lsd**2)
    w = ((S0[0] + 1j * S0[1]) * np.sinh(k * (L - x)) + (s1[0] + 1j * s1[1]) * np.sinh(k * x)) / np.sinh(k * L)
    assert np.abs(s[:, 0] - w.real).max() < 5e-3 * np.abs(w).max()
    assert np.abs(s - ref).max() <= 1e-9 * np.abs(ref).max()
    _, info2 = device_spin(m, [vol(), vol()], SURF, np.zeros(m.n_nodes), u, tol=1e-12, max_iter=2)
    assert info2.status == nv.FG_ITER_OVERFLOW and info2.iterations == 2


def spin_case(seed=5, cells=(10, 3, 4)):
    """A jittered bar, non-magnetic half then magnetic half, with a potential gradient and the
    surface conditions of SURF; the regions' spin parameters and three magnetization pairs."""
    m, own, pos = bar(0.25, cells=cells, ms=(0.0, 8e5))
    x = (pos[:, 0] - pos[:, 0].min()) / np.ptp(pos[:, 0])
    V = -0.02 * x + 0.003 * x**2
    vols = [vol(sigma=4e6, N0=2e47, lsf=8e-9), vol(sigma=6e6, N0=1e47, P=0.5, lsf=12e-9, lsd=4e-9)]
    rng = np.random.default_rng(seed)
    us = []
    for _ in range(6):
        u = rng.normal(size=(m.n_nodes, 3)) * 0.4 + [0.3, 0.0, 1.0]
        us.append(u / np.linalg.norm(u, axis=1, keepdims=True))
    return m, own, pos, V, vols, us[0::2], us[1::2]


def test_oracle_accumulated_solve_is_the_solve_at_the_mean_state():
    """The reference's spin solver adds every call's assembly into K and L_rhs (src/solver.h:
    110-142) and never clears them. The element matrix is affine in the exchange's u and the
    right-hand side affine in the source's u (the surface terms constant), so the k-th solve is the
    single solve at the mean magnetizations: k assemblies at one state give that state's solve,
    and three different states give the solve at their means. A dropped or doubled assembly, or a
    surface term added once instead of k times, fails one of these."""
    m, _own, _pos, V, vols, ux, usrc = spin_case()
    kw = dict(tol=1e-13, max_iter=5000)
    s1, _ = oracle.spin_accumulation(m, vols, SURF, V, ux[0], u_src=usrc[0], **kw)
    s3, i3 = oracle.spin_accumulation_hist(m, vols, SURF, V, [ux[0]] * 3, [usrc[0]] * 3, **kw)
    assert i3.status == nv.FG_ITER_CONVERGED
    assert np.abs(s3 - s1).max() <= 1e-9 * np.abs(s1).max()
    sh, ih = oracle.spin_accumulation_hist(m, vols, SURF, V, ux, usrc, **kw)
    assert ih.status == nv.FG_ITER_CONVERGED
    sm, _ = oracle.spin_accumulation(m, vols, SURF, V, np.mean(ux, axis=0), u_src=np.mean(usrc, axis=0), **kw)
    assert np.abs(sh - sm).max() <= 1e-9 * np.abs(sm).max()
    # and the accumulated solve is not the fresh one at the last state
    s_last, _ = oracle.spin_accumulation(m, vols, SURF, V, ux[-1], u_src=usrc[-1], **kw)
    assert np.abs(sh - s_last).max() > 1e-3 * np.abs(s_last).max()


@pytest.mark.gpu
def test_device_spin_solver_accumulates_like_the_reference():
    """The persistent solver (fg_spin_create / fg_spin_solve) over three calls with accumulate set
    against the oracle's summed system after each call, the same solver with accumulate cleared
    against the oracle's single solve, and the mean-state identity on the device."""
    from feellgood_amd.engine import SpinSolver
    m, _own, _pos, V, vols, ux, usrc = spin_case()
    sp = SpinSolver(m, vols, SURF, V, tol=1e-12, max_iter=5000)
    for k in range(1, 4):
        s_dev, info = sp.solve(ux[k - 1], usrc[k - 1], accumulate=True)
        s_orc, io = oracle.spin_accumulation_hist(m, vols, SURF, V, ux[:k], usrc[:k], tol=1e-12, max_iter=5000)
        assert info.status == io.status == nv.FG_ITER_CONVERGED, k
        assert abs(info.iterations - io.iterations) <= 2, (k, info.iterations, io.iterations)
        assert np.abs(s_dev - s_orc).max() <= 1e-8 * np.abs(s_orc).max(), k
    sm, _ = SpinSolver(m, vols, SURF, V, tol=1e-12, max_iter=5000).solve(np.mean(ux, axis=0), np.mean(usrc, axis=0))
    assert np.abs(s_dev - sm).max() <= 1e-8 * np.abs(sm).max()
    s_new, _ = sp.solve(ux[1], usrc[1], accumulate=False)
    s_one, _ = oracle.spin_accumulation(m, vols, SURF, V, ux[1], u_src=usrc[1], tol=1e-12, max_iter=5000)
    assert np.abs(s_new - s_one).max() <= 1e-8 * np.abs(s_one).max()
    sp.close()


@pytest.mark.gpu
def test_spin_run_in_lock_step_with_the_oracle():
    """Five steps of a spin-transfer run as the reference's loop takes them (src/time_integration.cpp:
    193-231, compute_all src/fem.h:203-224): after each solve the demag, then the spin accumulation
    from the state (s-d exchange from CURRENT, source from NEXT) through the persistent solver with
    accumulation, installed as the field of the next element pass (fg_spin_solve_ctx), then the
    evolution. The oracle re-solves the summed system of all calls so far from its own states. At
    every step s agrees to 1e-7 of its largest entry and u to 1e-9, while the accumulated s of the
    last step differs from a fresh solve at that state by more than 1e-5 (about 4e-5 here).
    Currents, spin densities and potentials are those of a run (s ~ 200 A/m), with dt = 1e-14 s:
    the explicit coupling of the magnetization to its spin accumulation grows without bound at a
    few times that step."""
    from feellgood_amd.engine import Engine, SpinSolver
    m, own, pos, V, _, _, _ = spin_case(cells=(8, 3, 3))
    V = V * 1e-4
    vols = [vol(sigma=4e6, N0=2e47, lsf=8e-9), vol(sigma=6e6, N0=1e47, P=0.5, lsf=12e-9, lsd=2e-8)]
    surfs = [surf(s=(200.0, -100.0, 50.0)), surf(jn=3e6, uP=UP), surf()]
    own[7][1].alpha, own[7][1].Ae = 0.5, 1.3e-11
    rng = np.random.default_rng(8)
    u0 = rng.normal(size=(m.n_nodes, 3)) * 0.1 + [0.3, 0.0, 1.0]
    u0 /= np.linalg.norm(u0, axis=1, keepdims=True)
    u0[pos[:, 0] < pos[:, 0].min() + 0.5 * np.ptp(pos[:, 0]) - 1e-9] = [0.0, 0.0, 1.0]
    tol = 1e-10
    p = oracle.Problem(m, tol=tol, max_iter=2000)
    g = Engine(m, tol=tol, max_iter=2000)
    sp = SpinSolver(m, vols, surfs, V, tol=1e-10, max_iter=2000)
    H = np.array([2e4, 0.0, 5e4])
    dt = 1e-14
    pref = 1 + 0.01 * abs(np.log(0.01))
    hu, hn = [], []

    def compute_all():
        for o in (p, g):
            o.calc_demag()
        hu.append(p.get_state(nv.FG_CURRENT)[0])
        hn.append(p.get_state(nv.FG_NEXT)[0])
        s_orc, io = oracle.spin_accumulation_hist(m, vols, surfs, V, hu, hn, tol=1e-10, max_iter=2000)
        assert io.status == nv.FG_ITER_CONVERGED
        p.set_spin_accumulation(s_orc, vols)
        s_dev, info = g.spin_solve(sp, accumulate=True)
        assert info.status == nv.FG_ITER_CONVERGED
        assert np.abs(s_dev - s_orc).max() <= 1e-7 * np.abs(s_orc).max(), len(hu)
        for o in (p, g):
            o.evolution()
        return s_dev

    for o in (p, g):
        o.set_state(nv.FG_BOTH, u=u0)
    compute_all()
    angles = oracle.basis_angles(3, 5)
    for k in range(5):
        for o in (p, g):
            o.base_projection(angles[k])
            o.prepare_elements(H, dt, pref)
            failed, _ = o.solve(dt)
            assert not failed
        s_dev = compute_all()
        u_p = p.get_state(nv.FG_CURRENT)[0]
        u_g = g.get_state(nv.FG_CURRENT)[0]
        assert np.nanmax(np.abs(u_g - u_p)) <= 1e-9, k
    s_fresh, _ = oracle.spin_accumulation(m, vols, surfs, V, hu[-1], u_src=hn[-1], tol=1e-10, max_iter=2000)
    assert np.abs(s_dev - s_fresh).max() > 1e-5 * np.abs(s_fresh).max()
    sp.close()
    g.close()
