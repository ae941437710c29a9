"""Pins the FP64 CPU oracle (CPU only).

Each check is independent of the oracle's own formula: known answers from FeeLLGood's unit tests
(legacy "ref code" formulations, inputs regenerated with std::mt19937(5489) exactly as the
reference's deterministic unit tests draw them), closed forms, invariants, brute force, and the
reference's end-to-end golden (ci-tests/full_test.py:45-50).
"""
import math
import os

import numpy as np
import pytest

import oracle
from feellgood_amd.mesh import HostMesh, volume_region

SEED = 5489  # std::mt19937_64::default_seed (unit-tests/ut_config.h.in:8-21)
MU0 = 1.25663706127e-6
GAMMA0 = 1.76085962784e11 * MU0
UNIT_TET = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)
A_TET = np.array([[1 - a - b - c for a, b, c in zip([.25, 1/6, 1/6, 1/6, .5], [.25, 1/6, 1/6, .5, 1/6], [.25, 1/6, .5, 1/6, 1/6])],
                  [.25, 1/6, 1/6, 1/6, .5], [.25, 1/6, 1/6, .5, 1/6], [.25, 1/6, .5, 1/6, 1/6]])
PDS_TET = np.array([-2/15, 3/40, 3/40, 3/40, 3/40])


def rand_vec3d(theta, phi):  # unit-tests/ut_tools.h:30-35
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])


class Draw:
    """Sequential draws of uniform_real_distribution<>(0,1) from mt19937(seed)."""

    def __init__(self, seed=SEED, n=4096):
        self.v = oracle.mt_uniform(seed, n)
        self.i = 0

    def __call__(self):
        x = self.v[self.i]
        self.i += 1
        return x


def geometry(pos, ind=(0, 1, 2, 3)):
    ind = np.array(ind, dtype=np.int32)
    da = np.zeros(12)
    w = np.zeros(5)
    rc = oracle.lib().orc_tet_geometry(oracle.d(np.ascontiguousarray(pos[ind], dtype=np.float64).reshape(-1)),
                                      ind.ctypes.data_as(oracle._i32), oracle.d(da), oracle.d(w))
    return rc, ind, da.reshape(4, 3), w


# ---------------------------------------------------------------- element geometry
def test_tet_inner_tables_unit_tet():
    """ut_tetra.cpp:21-120 (Tet_inner_tables, Tet_calc_vol): da = dadu*inv(J), w = detJ*pds."""
    rc, ind, da, w = geometry(UNIT_TET)
    assert rc == 0 and list(ind) == [0, 1, 2, 3]
    dadu = np.array([[-1., -1., -1.], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    assert np.array_equal(da, dadu)
    assert np.array_equal(w, PDS_TET)
    assert w.sum() == pytest.approx(1 / 6, rel=5e-16, abs=0)


def test_tet_geometry_random_tets_against_numpy():
    """Closed form: grad of the barycentric coordinates is inv([p1-p0, p2-p0, p3-p0]) rows;
    volume = |mixed product|/6; negative orientation swaps nodes 2 and 3 (src/tetra.cpp:410-424)."""
    rng = np.random.default_rng(7)
    for _ in range(50):
        P = rng.normal(size=(4, 3))
        rc, ind, da, w = geometry(P)
        assert rc == 0
        Q = P[ind]
        J = np.stack([Q[1] - Q[0], Q[2] - Q[0], Q[3] - Q[0]], axis=1)
        assert np.linalg.det(J) > 0
        Ji = np.linalg.inv(J)
        expect = np.vstack([-Ji.sum(axis=0), Ji])
        assert np.allclose(da, expect, rtol=1e-12, atol=1e-12 * np.abs(expect).max())
        vol = abs(np.dot(P[1] - P[0], np.cross(P[2] - P[0], P[3] - P[0]))) / 6
        assert w.sum() == pytest.approx(vol, rel=1e-13, abs=0)
        # sum of shape-function gradients vanishes (partition of unity)
        assert np.allclose(da.sum(axis=0), 0, atol=1e-12 * np.abs(da).max())
    assert geometry(np.zeros((4, 3)))[0] == -1  # singular


# ---------------------------------------------------------------- alpha_eff, lumping
def test_calc_alpha_eff_ut_time_int():
    """ut_time_int.cpp:45-96: inputs drawn alpha, X; dt = sqrt(1e-14*1e-9)."""
    dr = Draw()
    alpha, X = dr(), dr()
    dt = math.sqrt(1e-14 * 1e-9)
    rdt = GAMMA0 * dt
    M = 2 * alpha * 0.1 / rdt
    out = np.zeros(1)
    for h in (X, -X, 10 * M, -10 * M, 0.0, M, -M):
        oracle.lib().orc_alpha_eff(dt, alpha, oracle.d(np.array([h])), 1, oracle.d(out))
        if h > 0:
            ref = alpha + rdt / 2 * (M if h > M else h)
        else:
            ref = alpha / (1 + rdt / (2 * alpha) * M) if h < -M else alpha / (1 - rdt / (2 * alpha) * h)
        assert out[0] == pytest.approx(ref, rel=5e-16, abs=0)


def test_tet_lumping_legacy_loops():
    """ut_tet_lumping.cpp:94-197: the legacy per-Gauss-point loops (MuMag_integrales.cc) give the
    same 12x12 AE as Tet::lumping, for u drawn on the unit tet."""
    dr = Draw()
    u = np.array([rand_vec3d(math.pi * dr(), 2 * math.pi * dr()) for _ in range(4)])
    rc, ind, da, w = geometry(UNIT_TET)
    a, b = dr(), dr()
    dtmin, dtmax = min(a, b), max(a, b)
    dt = math.sqrt(dtmin * dtmax)
    TAUR = 100 * dtmax
    s_dt = 0.5 * dt
    alpha_llg = 0.5
    uHeff = np.full(5, dr())
    alfa = np.zeros(5)
    oracle.lib().orc_alpha_eff(dt, alpha_llg, oracle.d(uHeff), 5, oracle.d(alfa))
    A = dr()
    Js = 0.5 + dr()
    R = dt / TAUR * abs(math.log(dt / TAUR))
    AE = np.zeros((12, 12))
    for g in range(5):
        wg = w[g]
        for ie in range(4):
            ai = A_TET[ie, g]
            for b3 in range(3):
                AE[4 * b3 + ie, 4 * b3 + ie] += alfa[g] * ai * wg
            AE[ie, 8 + ie] += u[ie, 1] * ai * wg
            AE[ie, 4 + ie] += -u[ie, 2] * ai * wg
            AE[4 + ie, ie] += u[ie, 2] * ai * wg
            AE[4 + ie, 8 + ie] += -u[ie, 0] * ai * wg
            AE[8 + ie, 4 + ie] += u[ie, 0] * ai * wg
            AE[8 + ie, ie] += -u[ie, 1] * ai * wg
            for je in range(4):
                dd = np.dot(da[ie], da[je])
                for b3 in range(3):
                    AE[4 * b3 + ie, 4 * b3 + je] += s_dt * (1 + R) * 2 * A / Js * dd * wg
    got = np.zeros(144)
    prefactor = 1 + R
    oracle.lib().orc_tet_lumping(oracle.d(da.reshape(-1).copy()), oracle.d(w), oracle.d(u.reshape(-1).copy()),
                                 oracle.d(alfa), prefactor * s_dt * (2 * A / Js), oracle.d(got))
    assert np.allclose(got.reshape(12, 12), AE, rtol=1e-14, atol=1e-15 * np.abs(AE).max())


# ---------------------------------------------------------------- element matrices
def _integrales(da, w, u, v, phi, phiv, ep, eq, prm, H, dt, pref):
    Kp = np.zeros(64)
    Lp = np.zeros(8)
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    oracle.lib().orc_tet_integrales(oracle.d(c(da)), oracle.d(c(w)), oracle.d(c(u)), oracle.d(c(v)),
                                    oracle.d(c(phi)), oracle.d(c(phiv)), oracle.d(c(ep)), oracle.d(c(eq)),
                                    oracle.C.byref(prm), oracle.d(c(H)), dt, pref, oracle.d(Kp), oracle.d(Lp))
    return Kp.reshape(8, 8), Lp


def _basis(u, r):
    ep = np.zeros(3)
    eq = np.zeros(3)
    oracle.lib().orc_set_basis(oracle.d(np.ascontiguousarray(u, dtype=np.float64)), r, oracle.d(ep), oracle.d(eq))
    return ep, eq


def test_element_matrix_structure():
    """Block structure implied by AE (src/tetra.cpp:108-140): off-diagonal node blocks are the
    geometric stiffness times basis dot products, diagonal blocks carry the lumped terms.
    Kp[(di,a),(dj,b)] = e_r(di,a).(c G_ab I + delta_ab (D_a I + m_a [u_a]x)).e_s(dj,b),
    rows r = (q, p), columns s = (p, q)."""
    rng = np.random.default_rng(3)
    P = rng.normal(size=(4, 3)) * 3e-9
    rc, ind, da, w = geometry(P)
    u = np.array([rand_vec3d(*rng.uniform(0, 3, 2)) for _ in range(4)])
    v = rng.normal(size=(4, 3)) * 1e9
    phi = rng.normal(size=4) * 1e-3
    phiv = rng.normal(size=4) * 1e6
    eps, eqs = zip(*[_basis(u[i], rng.uniform(0, 1)) for i in range(4)])
    ep, eq = np.array(eps), np.array(eqs)
    prm = volume_region(alpha=0.1, Ae=1.3e-11, Ms=8e5, K=5e4, uk=(1, 2, 3), K3=-2e4)
    H = np.array([1e4, -2e4, 3e4])
    dt, pref = 1e-14, 1.046052
    Kp, Lp = _integrales(da, w, u, v, phi, phiv, ep, eq, prm, H, dt, pref)
    Abis = 2 * prm.Ae / (MU0 * prm.Ms)
    c = pref * 0.5 * dt * GAMMA0 * Abis
    G = w.sum() * da @ da.T
    m = A_TET @ w
    rows = lambda a: [eq[a], ep[a]]
    cols = lambda b: [ep[b], eq[b]]
    scale = np.abs(Kp).max()
    for a in range(4):
        for b in range(4):
            blk = np.array([[rr @ (c * G[a, b] * cc) for cc in cols(b)] for rr in rows(a)])
            if a == b:
                ux = np.array([[0, -u[a, 2], u[a, 1]], [u[a, 2], 0, -u[a, 0]], [-u[a, 1], u[a, 0], 0]])
                extra = np.array([[rr @ (m[a] * ux @ cc) for cc in cols(b)] for rr in rows(a)])
                got = np.array([[Kp[di * 4 + a, dj * 4 + b] for dj in range(2)] for di in range(2)])
                # remaining diagonal part is D_a * B(a,a): solve for D_a from the (0,1) entry
                rest = got - blk - extra
                Baa = np.array([[rr @ cc for cc in cols(b)] for rr in rows(a)])
                D = rest[0, 1] / Baa[0, 1]
                assert np.allclose(rest, D * Baa, rtol=1e-12, atol=1e-13 * scale)
                assert 0 < D < 2 * m[a] * 1.0  # alpha_eff-weighted mass, alpha_eff ~ alpha
            else:
                got = np.array([[Kp[di * 4 + a, dj * 4 + b] for dj in range(2)] for di in range(2)])
                assert np.allclose(got, blk, rtol=1e-12, atol=1e-13 * scale)
    # RHS: uniform m and no anisotropy field contribution at u: exchange part vanishes
    uu = np.tile(u[0], (4, 1))
    prm0 = volume_region(alpha=0.1, Ae=1.3e-11, Ms=8e5)
    Kp0, Lp0 = _integrales(da, w, uu, np.zeros((4, 3)), np.zeros(4), np.zeros(4), ep, eq, prm0,
                           np.zeros(3), dt, pref)
    assert np.allclose(Lp0, 0, atol=1e-30)
    # RHS: uniform field, uniform m, no exchange -> BE_i = m_i H (lumped), Lp = (eq.BE, ep.BE)
    Kp1, Lp1 = _integrales(da, w, uu, np.zeros((4, 3)), np.zeros(4), np.zeros(4), ep, eq, prm0, H, dt, pref)
    for a in range(4):
        assert Lp1[a] == pytest.approx(m[a] * eq[a] @ H, rel=1e-12, abs=1e-12 * np.abs(Lp1).max())
        assert Lp1[4 + a] == pytest.approx(m[a] * ep[a] @ H, rel=1e-12, abs=1e-12 * np.abs(Lp1).max())


# ---------------------------------------------------------------- node basis
def test_node_basis_ut_node():
    """ut_node.cpp:58-147: e_p starts on the axis furthest from u, (u, ep, eq) orthonormal,
    rotated by r; make_evol v = vp ep + vq eq."""
    dr = Draw()
    for _ in range(40):
        th = math.pi * (2 * dr() - 1)
        ph = 2 * dr() - 1
        u = rand_vec3d(th, ph)
        r = math.pi * (2 * dr() - 1)
        ep, eq = _basis(u, r)
        M = np.stack([u, ep, eq])
        assert np.allclose(M @ M.T, np.eye(3), atol=1e-15)
        assert np.allclose(np.cross(ep, eq), u, atol=1e-15)  # right-handed: eq = u x ep rotated
        ep0, eq0 = _basis(u, 0.0)
        k = int(np.argmin(np.abs(u)))
        e = np.zeros(3); e[k] = 1
        e = e - (e @ u) * u
        assert np.allclose(ep0, e / np.linalg.norm(e), atol=1e-15)
        assert np.allclose(ep, math.cos(r) * ep0 - math.sin(r) * eq0, atol=1e-15)


# ---------------------------------------------------------------- triangle potential
def _potential(P, s, dMs, i):
    return oracle.lib().orc_tri_potential(oracle.d(np.ascontiguousarray(P, dtype=np.float64).reshape(-1)),
                                          oracle.d(np.ascontiguousarray(s, dtype=np.float64)), dMs, i)


def _legacy_potential(P, s, dMs, i):
    """ut_triangle.cpp:129-241 (MuMag_potential.cc closed form): returns dMs*pot."""
    ii, iii = (i + 1) % 3, (i + 2) % 3
    p1, p2, p3 = P[i], P[ii], P[iii]
    s1, s2, s3 = s[i], s[ii], s[iii]
    surf = 0.5 * np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0]))
    b = np.linalg.norm(p2 - p1)
    t = np.dot(p2 - p1, p3 - p1) / b
    h = 2 * surf / b
    a = t / h
    c = (t - b) / h
    l = s1
    j = (s2 - s1) / b
    k = t / b / h * (s1 - s2) + (s3 - s1) / h
    cc1 = c * c + 1
    r = math.sqrt(h * h + (c * h + b) ** 2)
    ll = math.log((cc1 * h + c * b + math.sqrt(cc1) * r) / (b * (c + math.sqrt(cc1))))
    pot1 = (b * b / cc1 ** 1.5 * ll + c * b * r / cc1 + h * r - c * b * b / cc1 - math.sqrt(a * a + 1) * h * h) * j / 2
    pot2 = (-c * b * b / cc1 ** 1.5 * ll + b * r / cc1 - h * h / 2 + h * h * math.log(c * h + b + r) - b * b / cc1) * k / 2
    pot3 = (h * math.log(c * h + b + r) - h + b / math.sqrt(cc1) * ll) * l
    pot = pot1 + pot2 + pot3 + h * (k * h / 2 + l) * (1 - math.log(h * (a + math.sqrt(a * a + 1)))) - k * h * h / 4
    return dMs * pot


def _quad_potential(P, s, i, n=400):
    """Brute force: integral over the triangle of the linear density sum_j s_j lambda_j at vertex i
    divided by the distance, by a Duffy transform that removes the 1/r singularity."""
    x = np.polynomial.legendre.leggauss(n)
    g, wg = (x[0] + 1) / 2, x[1] / 2
    ii, iii = (i + 1) % 3, (i + 2) % 3
    Pi, Pj, Pk = P[i], P[ii], P[iii]
    area2 = np.linalg.norm(np.cross(Pj - Pi, Pk - Pi))
    tot = 0.0
    for a_, wa in zip(g, wg):
        y = Pi + a_ * ((Pj - Pi) + g[:, None] * (Pk - Pj))  # (n,3)
        lam_j = a_ * (1 - g)
        lam_k = a_ * g
        dens = s[i] * (1 - lam_j - lam_k) + s[ii] * lam_j + s[iii] * lam_k
        r = np.linalg.norm(y - Pi, axis=1)
        tot += wa * np.sum(wg * dens * a_ * area2 / r)
    return tot


def test_tri_potential_legacy_closed_form_and_quadrature():
    """Tri::potential (src/triangle.cpp:87-125) equals the legacy MuMag closed form
    (ut_triangle.cpp:129-241, result_ref = dMs*pot) and dMs times the brute-force integral of the
    linear density sum_j s_j lambda_j over |x - x_i| at vertex i."""
    dr = Draw()
    for case in range(6):
        P = np.array([[1 + (dr() - .5) / 10, (dr() - .5) / 10, (dr() - .5) / 10] for _ in range(3)])
        s = np.array([2 * dr() - 1 for _ in range(3)])
        dMs = dr()
        for i in range(3):
            got = _potential(P, s, dMs, i)
            assert got == pytest.approx(_legacy_potential(P, s, dMs, i), rel=1e-12, abs=0)
            assert got == pytest.approx(dMs * _quad_potential(P, s, i), rel=1e-9, abs=0)


# ---------------------------------------------------------------- Krylov
def _csr(n, coeffs):
    rows = [dict() for _ in range(n)]
    for r, c, v in coeffs:
        rows[r][c] = rows[r].get(c, 0.0) + v
    rp = [0]
    col, val = [], []
    for r in rows:
        for c in sorted(r):
            col.append(c)
            val.append(r[c])
        rp.append(len(col))
    return np.array(rp), np.array(col), np.array(val)


def test_bicg_dir_small_exact_system():
    """ut_algebra.cpp:245-274 matrix: the converged x reproduces b."""
    rp, col, val = _csr(4, [(1, 1, 3.14), (0, 0, 1.0), (2, 2, 5.0), (3, 3, 42.0), (1, 3, -10.0),
                            (1, 3, 10.0), (0, 3, 0.5)])
    b = np.ones(4)
    x, info = oracle.bicg_dir(rp, col, val, b, np.zeros(4), tol=1e-12)
    A = np.zeros((4, 4))
    for r in range(4):
        A[r, col[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    assert info.status == 0
    assert np.allclose(A @ x, b, atol=1e-12)


def test_bicg_dir_identity_zero_iterations():
    """ut_algebra_bicg.cpp:19-44: identity with x0 = b converges before the first iteration."""
    n = 1000
    rp, col, val = _csr(n, [(i, i, 1.0) for i in range(n)])
    b = np.ones(n)
    x, info = oracle.bicg_dir(rp, col, val, b, b.copy(), tol=1e-8)
    assert info.iterations == 0 and info.status == 0 and np.array_equal(x, b)


def test_bicg_dir_laplacian_dirichlet_linear_profile():
    """ut_algebra.cpp:277-327: 1-D Laplacian, Dirichlet 0 and 1 at the ends -> linear profile.
    The 5-argument bicg_dir masks the Dirichlet rows; the boundary values enter through b."""
    n = 201
    co = [(0, 0, 1.0), (0, 1, -1.0), (n - 1, n - 2, -1.0), (n - 1, n - 1, 1.0)]
    for k in range(1, n - 1):
        co += [(k, k - 1, -1.0), (k, k, 2.0), (k, k + 1, -1.0)]
    rp, col, val = _csr(n, co)
    xd = np.zeros(n)
    xd[-1] = 1.0
    A = np.zeros((n, n))
    for r in range(n):
        A[r, col[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    b = -A @ xd
    x, info = oracle.bicg_dir(rp, col, val, b, np.zeros(n), mask=[0, n - 1], tol=1e-12, max_iter=5000)
    assert info.status == 0
    assert np.allclose(x + xd, np.arange(n) / (n - 1), atol=1e-9)


def test_bicg_iteration_overflow_status():
    """iteration::operator++ sets ITER_OVERFLOW at max_iter (src/algebra/iter.h:119-125)."""
    n = 201
    co = []
    for k in range(n):
        co += [(k, k, 2.0)] + ([(k, k - 1, -1.0)] if k else []) + ([(k, k + 1, -1.0)] if k < n - 1 else [])
    rp, col, val = _csr(n, co)
    x, info = oracle.bicg_dir(rp, col, val, np.ones(n), np.zeros(n), tol=1e-14, max_iter=3)
    assert info.status == 1 and info.iterations == 3


# ---------------------------------------------------------------- whole-problem invariants
def small_cube(n=4, L=20.0, jitter=0.15, seed=1):
    m = HostMesh.cuboid((L, L, L), (n, n, n), jitter=jitter, seed=seed)
    m.prepare({"volume": dict(alpha=0.5, Ae=1.3e-11, Ms=8e5, K=2e4, uk=(1, 1, 0))})
    return m


def random_state(N, seed=2):
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(N, 3)) + np.array([1.0, 1.0, 0.0])
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    return u


def test_solve_invariants_and_basis_independence():
    """v is tangent, |u_next| = 1, the residual recomputed independently meets the tolerance, and
    the 3-D velocity does not depend on the random tangent basis (within the Krylov tolerance)."""
    m = small_cube()
    N = m.n_nodes
    u = random_state(N)
    H = np.array([2e4, -1e4, 5e3])
    dt, pref = 1e-13, 1.046052
    vs = []
    for r in (0.1, 0.55):
        p = oracle.Problem(m.mesh, tol=1e-12, max_iter=2000)
        p.set_state(2, u=u)
        p.calc_demag()
        p.evolution()
        p.base_projection(r)
        p.prepare_elements(H, dt, pref)
        rc, info = p.solve(dt)
        assert rc == 0 and info.status == 0
        rp, col, val, rhs = p.system()
        x = p.solution()
        Ax = np.zeros_like(x)
        for i in range(len(x)):
            Ax[i] = val[rp[i]:rp[i + 1]] @ x[col[rp[i]:rp[i + 1]]]
        assert np.linalg.norm(rhs - Ax) <= 1e-11 * np.linalg.norm(rhs)
        un, vn, _, _ = p.get_state(1)
        assert np.allclose(np.linalg.norm(un, axis=1), 1, atol=1e-15)
        assert np.abs(np.sum(vn * u, axis=1)).max() <= 1e-12 * np.abs(vn).max()
        vs.append(vn)
    assert np.allclose(vs[0], vs[1], rtol=0, atol=1e-9 * np.abs(vs[0]).max())


def test_space_field_reduces_to_the_uniform_field():
    """The R4toR3 path (src/linear_algebra.cpp:54-75, src/energy.cpp:11-16): a space field equal to
    the uniform field H at every Gauss point with amplitude 1 gives the uniform path's system,
    solution and energies exactly; amplitude and field scale together (amp * h = (2 amp) * h / 2)."""
    m = small_cube()
    u = random_state(m.n_nodes)
    H = np.array([2e4, -1e4, 5e3])
    dt, pref = 1e-13, 1.046052
    T = m.n_tets

    def run(setup):
        p = oracle.Problem(m.mesh, tol=1e-12, max_iter=2000)
        p.set_state(2, u=u)
        p.calc_demag()
        p.evolution()
        p.base_projection(0.3)
        E = setup(p)
        rc, _ = p.solve(dt)
        assert rc == 0
        return p.system()[2:], p.solution(), E

    def uniform(p):
        p.prepare_elements(H, dt, pref)
        return p.energy(H)

    def space(amp, scale):
        def f(p):
            p.set_space_field(np.tile(H * scale, (T, 5, 1)))
            p.prepare_elements_amp(amp, dt, pref)
            return p.energy_amp(amp)
        return f
    (v0, b0), x0, E0 = run(uniform)
    for amp, scale in ((1.0, 1.0), (0.5, 2.0)):
        (v1, b1), x1, E1 = run(space(amp, scale))
        assert np.array_equal(v1, v0) and np.array_equal(b1, b0) and np.array_equal(x1, x0)
        assert np.array_equal(E1, E0)
    # a field that varies in space changes the right-hand side only through the Zeeman terms
    p = oracle.Problem(m.mesh, tol=1e-12, max_iter=2000)
    p.set_state(2, u=u)
    p.calc_demag()
    p.evolution()
    p.base_projection(0.3)
    h = np.random.default_rng(4).normal(size=(T, 5, 3)) * 1e4
    p.set_space_field(h)
    p.prepare_elements_amp(0.0, dt, pref)
    E = p.energy_amp(0.0)
    assert E[3] == 0.0


def test_charges_sum_to_zero_divergence_theorem():
    """For a P1 field the volume charges and the surface charges cancel exactly (divergence
    theorem per element, with the 4-point triangle rule exact for linear u.n)."""
    m = small_cube()
    p = oracle.Problem(m.mesh)
    p.set_state(2, u=random_state(m.n_nodes), v=random_state(m.n_nodes, 5))
    p.calc_demag()
    xyz, qu, qv = p.sources()
    assert abs(qu.sum()) <= 1e-12 * np.abs(qu).sum()
    assert abs(qv.sum()) <= 1e-12 * np.abs(qv).sum()


def test_direct_demag_against_numpy():
    """phi_i = (sum_s q_s/|x_i-y_s| + corr_i)/(4 pi) (src/fmm_demag.h:211-223), recomputed."""
    m = small_cube(n=3)
    p = oracle.Problem(m.mesh)
    p.set_state(2, u=random_state(m.n_nodes))
    p.calc_demag()
    xyz, qu, qv = p.sources()
    cu, cv = p.corrections()
    X = m.arrays()["pos"]
    r = np.linalg.norm(X[:, None, :] - xyz[None, :, :], axis=2)
    phi = ((qu[None, :] / r).sum(axis=1) + cu) / (4 * np.pi)
    _, _, got, _ = p.get_state(1)
    assert np.allclose(got, phi, rtol=1e-12, atol=1e-12 * np.abs(phi).max())


def test_uniform_sphere_demag_field():
    """Textbook: a uniformly magnetised sphere has H_d = -M/3 inside; the P1 potential's gradient
    averaged over the volume matches to mesh accuracy."""
    m = HostMesh.ellipsoid((0, 0, 0), (10.0, 10.0, 10.0), 1.25)
    m.prepare({"volume": dict(Ms=8e5)})
    p = oracle.Problem(m.mesh)
    N = m.n_nodes
    p.set_state(2, u=np.tile([0.0, 0.0, 1.0], (N, 1)))
    p.calc_demag()
    _, _, phi, _ = p.get_state(1)
    a = m.arrays()
    pos, tets = a["pos"], a["tets"]
    Hz, vol = 0.0, 0.0
    for t in tets:
        P = pos[t]
        J = np.stack([P[1] - P[0], P[2] - P[0], P[3] - P[0]], axis=1)
        Ji = np.linalg.inv(J)
        da = np.vstack([-Ji.sum(axis=0), Ji])
        V = abs(np.linalg.det(J)) / 6
        Hz += -(phi[t] @ da[:, 2]) * V
        vol += V
    assert Hz / vol == pytest.approx(-8e5 / 3, rel=0.03, abs=0)


def test_energy_closed_forms():
    """Zeeman energy of uniform m = -mu0 Ms V m.H; exchange energy of a linear field is exact for
    P1 elements: A |grad u|^2 V; uniform m has zero exchange energy."""
    m = small_cube(n=3, jitter=0.2)
    N = m.n_nodes
    pos = m.arrays()["pos"]
    p = oracle.Problem(m.mesh)
    mz = np.tile([0.6, 0.0, 0.8], (N, 1))
    p.set_state(2, u=mz)
    H = np.array([1e4, 2e4, -3e4])
    E = p.energy(H)
    V = (20e-9) ** 3
    assert E[3] == pytest.approx(-MU0 * 8e5 * V * (mz[0] @ H), rel=1e-12, abs=0)
    assert abs(E[0]) < 1e-30
    lin = np.zeros((N, 3))
    lin[:, 0] = pos[:, 1] * 1e7  # u_x = 1e7 * y  -> |grad u|^2 = 1e14
    p.set_state(2, u=lin)
    E = p.energy(np.zeros(3))
    assert E[0] == pytest.approx(1.3e-11 * 1e14 * V, rel=1e-12, abs=0)


def test_demag_energy_volume_equals_surface_form():
    """ut_energy.cpp:78-168: -1/2 mu0 Ms int u.Hd over a tet equals the volume charge term plus
    the surface-charge term of its four faces (divergence theorem). Checked on a whole cube: the
    oracle's E_demag (volume + surface form) equals -1/2 mu0 Ms int u . (-grad phi)."""
    m = small_cube(n=3, jitter=0.2)
    N = m.n_nodes
    p = oracle.Problem(m.mesh)
    u = random_state(N, 9)
    rng = np.random.default_rng(4)
    phi = rng.normal(size=N)
    p.set_state(2, u=u, phi=phi)
    E = p.energy(np.zeros(3))
    a = m.arrays()
    tot = 0.0
    for t in a["tets"]:
        P = a["pos"][t]
        J = np.stack([P[1] - P[0], P[2] - P[0], P[3] - P[0]], axis=1)
        Ji = np.linalg.inv(J)
        da = np.vstack([-Ji.sum(axis=0), Ji])
        V = abs(np.linalg.det(J)) / 6
        Hd = -(phi[t] @ da)
        ubar = u[t].mean(axis=0)  # P1 field integrates to the vertex mean
        tot += -0.5 * MU0 * 8e5 * V * (ubar @ Hd)
    assert E[2] == pytest.approx(tot, rel=1e-10, abs=0)


# ---------------------------------------------------------------- the reference's golden
@pytest.mark.slow
def test_full_test_golden(golden_dir):
    """ci-tests/full_test.py:21-50 with --seed 2 on examples/ellipsoid.msh (NPI = 5 build):
    last .evol row <m> = (0.307542, 0.475877, -0.823989), E_tot = -4.445605e-19, both within 1e-5."""
    m = HostMesh.read(os.path.join(golden_dir, "ellipsoid.msh"))
    m.prepare({"ellipsoid_volume": dict(K=3e5, uk=(0, 1, 0))})
    assert (m.n_nodes, m.n_tets, m.n_tris) == (167, 499, 274)
    st, evol = oracle.run(m.mesh, np.tile([0, 0, 1.0], (m.n_nodes, 1)), final_time=2e-11,
                          evol_time_step=1e-12, dt_min=5e-14, dt_max=1e-12, du_max=0.1, Bext=(1, 0, -1), seed=2)
    assert st == 0 and len(evol) == 21
    last = evol[-1]
    assert np.linalg.norm(last[1:4] - np.array([0.307542, 0.475877, -0.823989])) < 1e-5
    assert abs((last[8] - (-4.445605e-19)) / -4.445605e-19) < 1e-5


def test_basis_angles_reproduce_rand_mt19937():
    """base_projection draws r = M_2_PI * U(0,1) from std::mt19937(rand()) after srand(seed)."""
    a = oracle.basis_angles(2, 4)
    assert np.all((a >= 0) & (a < 2 / math.pi))
    assert np.array_equal(a, oracle.basis_angles(2, 4))
    assert not np.array_equal(a, oracle.basis_angles(3, 4))


def test_spin_field_reduces_to_an_applied_field():
    """The spin-accumulation field of the element pass (src/tetra.cpp:93-106,248-255,296-302):
    with a uniform spin accumulation c, Hst = D0 / (lsd^2 gamma0 Ms) c at every Gauss point, and
    with v = 0 the system is the one of the applied field Hext + Hst (the second-order term
    vanishes); with v != 0 the right-hand side differs from it by a term linear in v."""
    from feellgood_amd import native as nv
    m = small_cube()
    u = random_state(m.n_nodes)
    H = np.array([2e4, -1e4, 5e3])
    c = np.array([3e5, 1e5, -2e5])
    sv = nv.SpinVolume(5e6, 1e47, 0.4, 5e-9, 10e-9)
    vols = [sv] * m.mesh.n_vol_regions
    E = 1.602176634e-19
    gamma0 = 1.76085962784e11 * 1.25663706127e-6
    hst = 2 * sv.sigma / (E**2 * sv.N0) / (sv.lsd**2 * gamma0 * m.mesh.vol[0].Ms) * c
    dt, pref = 1e-13, 1.046052
    rng = np.random.default_rng(2)
    vrand = np.cross(u, rng.normal(size=u.shape)) * 1e9

    def system(v, spin, Hext):
        p = oracle.Problem(m.mesh, tol=1e-12, max_iter=2000)
        p.set_state(2, u=u, v=v)
        p.calc_demag()
        p.evolution()
        p.base_projection(0.3)
        if spin:
            p.set_spin_accumulation(np.tile(c, (m.n_nodes, 1)), vols)
        p.prepare_elements(Hext, dt, pref)
        p.solve(dt)  # assembles the system
        return p.system()[2:]

    v0 = np.zeros_like(u)
    (a1, b1), (a0, b0) = system(v0, True, H), system(v0, False, H + hst)
    assert np.allclose(a1, a0, rtol=1e-12, atol=1e-12 * np.abs(a0).max())
    assert np.allclose(b1, b0, rtol=1e-12, atol=1e-12 * np.abs(b0).max())
    d = [system(k * vrand, True, H)[1] - system(k * vrand, False, H + hst)[1] for k in (1.0, 2.0)]
    assert np.abs(d[0]).max() > 1e-6 * np.abs(b0).max()
    assert np.allclose(d[1], 2 * d[0], rtol=1e-6, atol=1e-9 * np.abs(d[0]).max())
