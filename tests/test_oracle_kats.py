"""Pins the oracle's anisotropy, charge, energy and surface-anisotropy functions (CPU only).

Known answers are the legacy "ref code" formulations FeeLLGood's own unit tests hold, evaluated on
inputs drawn from std::mt19937(5489) in the order the tests draw them:

* unit-tests/ut_anisotropy.cpp:22-120 (uniaxial field and uHeff term), :121-252 (cubic, including
  the component-wise theta term of :228-231);
* unit-tests/ut_energy.cpp:18-76 (uniaxial energy), :170-196 (Zeeman with a field varying over the
  Gauss points);
* unit-tests/ut_tet_charges.cpp:18-58, unit-tests/ut_tri_charges.cpp:18-53.

What those leave open is pinned by the mathematics: the P1 mass matrices (the Gauss rules are exact
for quadratics), so the surface-anisotropy right-hand side (src/triangle.cpp:6-36) and the surface
anisotropy and demag energies of a triangle have closed forms for any nodal values; the volume
energies of a tetrahedron likewise; the anisotropy fields are minus the gradient of their energy
densities (finite differences on the tangent plane); the cubic energy is invariant under a common
rotation of the axes and of m.
"""
import math

import numpy as np
import pytest

import oracle
from feellgood_amd import native as nv
from tests.test_oracle import A_TET, Draw, UNIT_TET, geometry, rand_vec3d

MU0 = 1.25663706127e-6
THETA = 0.5
UT_TOL = 5e-16  # unit-tests/ut_config.h.in:24
TRI3 = np.array([[1.0, 0, 0], [0, 1.0, 0], [1.0, 1.0, 0]])  # dummyNodes<3> (unit-tests/ut_tools.h:20-27)
PDS_TRI = np.array([-27 / 96, 25 / 96, 25 / 96, 25 / 96])
A_TRI = np.array([[1 - a - b for a, b in zip([1 / 3, 1 / 5, 3 / 5, 1 / 5], [1 / 3, 1 / 5, 1 / 5, 3 / 5])],
                  [1 / 3, 1 / 5, 3 / 5, 1 / 5], [1 / 3, 1 / 5, 1 / 5, 3 / 5]])

c64 = lambda a: np.ascontiguousarray(a, dtype=np.float64).reshape(-1)


def aniso_uniax(U, V, uk, Kbis, s_dt):
    H, ctb = np.zeros(15), np.zeros(5)
    oracle.lib().orc_aniso_uniax(oracle.d(c64(U)), oracle.d(c64(V)), oracle.d(c64(uk)), Kbis, s_dt,
                                 oracle.d(H), oracle.d(ctb))
    return H.reshape(5, 3), ctb


def aniso_cub(U, V, ex, ey, ez, K3bis, s_dt):
    H, ctb = np.zeros(15), np.zeros(5)
    oracle.lib().orc_aniso_cub(oracle.d(c64(U)), oracle.d(c64(V)), oracle.d(c64(ex)), oracle.d(c64(ey)),
                               oracle.d(c64(ez)), K3bis, s_dt, oracle.d(H), oracle.d(ctb))
    return H.reshape(5, 3), ctb


def region(**kw):
    r = nv.VolumeRegion()
    r.alpha, r.Ae, r.Ms = kw.get("alpha", 0.5), kw.get("Ae", 0.0), kw.get("Ms", 1.0)
    r.K, r.K3 = kw.get("K", 0.0), kw.get("K3", 0.0)
    for name, default in (("uk", (0, 0, 1)), ("ex", (1, 0, 0)), ("ey", (0, 1, 0)), ("ez", (0, 0, 1))):
        getattr(r, name)[:] = kw.get(name, default)
    return r


def tet_energy(pos, u, phi, prm, Hg):
    rc, ind, da, w = geometry(pos)
    assert rc == 0 and list(ind) == [0, 1, 2, 3]
    e = np.zeros(5)
    oracle.lib().orc_tet_energy(oracle.d(c64(da)), oracle.d(w), oracle.d(c64(u)), oracle.d(c64(phi)),
                                oracle.C.byref(prm), oracle.d(c64(Hg)), oracle.d(e))
    return e, w


def tri_energy(pos, u, phi, dMs, Ks, uk):
    e = np.zeros(2)
    oracle.lib().orc_tri_energy(oracle.d(c64(pos)), oracle.d(c64(u)), oracle.d(c64(phi)), dMs, Ks,
                                oracle.d(c64(uk)), oracle.d(e))
    return e


def tri_mass(pos):
    """P1 consistent mass matrix of a triangle: int lambda_i lambda_j dA = A/12 (1 + delta_ij)."""
    A = 0.5 * np.linalg.norm(np.cross(pos[1] - pos[0], pos[2] - pos[0]))
    return A / 12 * (np.ones((3, 3)) + np.eye(3)), A


def tet_mass(pos):
    """P1 consistent mass matrix of a tetrahedron: int lambda_i lambda_j dV = V/20 (1 + delta_ij)."""
    V = abs(np.dot(pos[1] - pos[0], np.cross(pos[2] - pos[0], pos[3] - pos[0]))) / 6
    return V / 20 * (np.ones((4, 4)) + np.eye(4)), V


# ---------------------------------------------------------------- anisotropy fields
def test_uniaxial_field_ut_anisotropy():
    """ut_anisotropy.cpp:22-120: the legacy MuMag_integrales.cc field H + theta dt Ht and uHau."""
    dr = Draw()
    u, v = np.zeros((4, 3)), np.zeros((4, 3))
    for i in range(4):
        u[i] = rand_vec3d(math.pi * dr(), 2 * math.pi * dr())
        v[i] = rand_vec3d(math.pi * dr(), 2 * math.pi * dr())
    dt = dr()
    uk = rand_vec3d(math.pi * dr(), 2 * math.pi * dr())
    K = dr()
    Js = 0.5 + dr()
    U, V = A_TET.T @ u, A_TET.T @ v  # tiny::mult(u_nod, Tetra::a)
    H, ctb = aniso_uniax(U, V, uk, 2.0 * K / Js, THETA * dt)
    for g in range(5):
        uk0_u = uk[0] * U[g, 0] + uk[1] * U[g, 1] + uk[2] * U[g, 2]
        uk0_v = uk[0] * V[g, 0] + uk[1] * V[g, 1] + uk[2] * V[g, 2]
        uHau = 2 * K / Js * uk0_u * uk0_u
        Ht = 2 * K / Js * uk0_v * uk
        Hl = 2 * K / Js * uk0_u * uk
        ref = Hl + THETA * dt * Ht
        assert np.linalg.norm(ref - H[g]) <= 10 * UT_TOL * np.linalg.norm(ref)
        assert ctb[g] == pytest.approx(uHau, rel=10 * UT_TOL, abs=0)


def cubic_inputs(dr):
    u, v = np.zeros((4, 3)), np.zeros((4, 3))
    for i in range(4):
        u[i] = rand_vec3d(math.pi * dr(), 2 * math.pi * dr())
        v[i] = rand_vec3d(math.pi * dr(), 2 * math.pi * dr())
    dt = dr()
    rand_vect = rand_vec3d(math.pi * dr(), 2 * math.pi * dr())
    ex = rand_vec3d(math.pi * dr(), 2 * math.pi * dr())
    ey = np.cross(ex, rand_vect)
    ez = np.cross(ex, ey)
    ey /= np.linalg.norm(ey)
    ez /= np.linalg.norm(ez)
    K3 = dr()
    Js = 0.5 + dr()
    return u, v, dt, ex, ey, ez, K3, Js


def test_cubic_field_ut_anisotropy():
    """ut_anisotropy.cpp:121-252: the legacy cubic field, whose theta term takes the components of
    ex only (Ht[k] = -2 K3/Js (e_k . V)(1 - 3 (e_k . U)^2) ex[k], :228-231), and uHa3u."""
    u, v, dt, ex, ey, ez, K3, Js = cubic_inputs(Draw())
    U, V = A_TET.T @ u, A_TET.T @ v
    H, ctb = aniso_cub(U, V, ex, ey, ez, 2.0 * K3 / Js, THETA * dt)
    E = np.stack([ex, ey, ez])  # rows: the axes
    for g in range(5):
        b = E @ U[g]
        bv = E @ V[g]
        uHa3u = -2 * K3 / Js * sum(b[a] * (1 - b[a] * b[a]) * b[a] for a in range(3))
        Ht = np.array([-2 * K3 / Js * bv[k] * (1 - 3 * b[k] * b[k]) * ex[k] for k in range(3)])
        Hl = sum(-2 * K3 / Js * b[a] * (1 - b[a] * b[a]) * E[a] for a in range(3))
        ref = Hl + THETA * dt * Ht
        assert np.linalg.norm(ref - H[g]) <= 10 * UT_TOL * np.linalg.norm(ref) + 1e-300
        assert ctb[g] == pytest.approx(uHa3u, rel=1e-14, abs=1e-15)


def _energy_density(u, kind, par):
    """Energy density of a uniform unit m over the unit tet, from the oracle's volume energy."""
    prm = region(**par)
    e, w = tet_energy(UNIT_TET, np.tile(u, (4, 1)), np.zeros(4), prm, np.zeros((5, 3)))
    return (e[1] if kind == "uniax" else e[2]) / w.sum()


@pytest.mark.parametrize("kind", ["uniax", "cubic"])
def test_anisotropy_field_is_minus_the_energy_gradient(kind):
    """H_an = -(1/(mu0 Ms)) grad_m e(m) on the tangent plane, and uHeff's anisotropy term is
    m . H_an: central differences of the oracle's energy density along two tangent directions
    against the oracle's field (V = 0, the first-order part)."""
    rng = np.random.default_rng(11)
    Ms = 7.3e5
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    for _ in range(5):
        m = rng.normal(size=3)
        m /= np.linalg.norm(m)
        if kind == "uniax":
            uk = rng.normal(size=3)
            uk /= np.linalg.norm(uk)
            par = dict(K=3.1e4, uk=tuple(uk))
            H, ctb = aniso_uniax(np.tile(m, (5, 1)), np.zeros((5, 3)), uk, 2 * par["K"] / (MU0 * Ms), 0.0)
        else:
            par = dict(K3=-2.7e4, ex=tuple(Q[0]), ey=tuple(Q[1]), ez=tuple(Q[2]))
            H, ctb = aniso_cub(np.tile(m, (5, 1)), np.zeros((5, 3)), Q[0], Q[1], Q[2],
                               2 * par["K3"] / (MU0 * Ms), 0.0)
        t1 = np.cross(m, rng.normal(size=3))
        t1 /= np.linalg.norm(t1)
        t2 = np.cross(m, t1)
        h = 1e-5
        for t in (t1, t2):
            ep = _energy_density((m + h * t) / np.linalg.norm(m + h * t), kind, par)
            em = _energy_density((m - h * t) / np.linalg.norm(m - h * t), kind, par)
            grad_t = (ep - em) / (2 * h)
            assert -grad_t / (MU0 * Ms) == pytest.approx(np.dot(H[0], t), rel=1e-6, abs=1e-9 * np.abs(H).max())
        assert np.allclose(ctb, np.dot(H[0], m), rtol=1e-13, atol=1e-13 * abs(ctb).max())


def test_theta_term_of_the_uniaxial_field_is_its_derivative_along_v():
    """The second-order uniaxial term is s_dt d/de H_an(U + e V): H(U, V) = H(U + s_dt V, 0)."""
    rng = np.random.default_rng(12)
    U, V = rng.normal(size=(5, 3)), rng.normal(size=(5, 3))
    uk = rng.normal(size=3)
    uk /= np.linalg.norm(uk)
    H1, _ = aniso_uniax(U, V, uk, 1.7, 0.31)
    H2, _ = aniso_uniax(U + 0.31 * V, np.zeros((5, 3)), uk, 1.7, 0.0)
    assert np.allclose(H1, H2, rtol=1e-14, atol=1e-14)


def test_cubic_energy_closed_form_and_rotation_invariance():
    """Uniform m: E = K3 V (b0^2 b1^2 + b1^2 b2^2 + b2^2 b0^2); rotating the axes and m together
    leaves E unchanged; along an axis E = 0, along (1,1,1)/sqrt3 E = K3 V / 3."""
    rng = np.random.default_rng(13)
    P = UNIT_TET * 2e-9 + np.array([1e-9, 0, 0])
    Vol = abs(np.linalg.det(P[1:] - P[0])) / 6
    for _ in range(4):
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        m = rng.normal(size=3)
        m /= np.linalg.norm(m)
        prm = region(K3=4.2e4, ex=tuple(Q[0]), ey=tuple(Q[1]), ez=tuple(Q[2]))
        e, _ = tet_energy(P, np.tile(m, (4, 1)), np.zeros(4), prm, np.zeros((5, 3)))
        b = Q @ m
        ref = 4.2e4 * Vol * (b[0] ** 2 * b[1] ** 2 + b[1] ** 2 * b[2] ** 2 + b[2] ** 2 * b[0] ** 2)
        assert e[2] == pytest.approx(ref, rel=1e-13, abs=0)
        R, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        prm2 = region(K3=4.2e4, ex=tuple(R @ Q[0]), ey=tuple(R @ Q[1]), ez=tuple(R @ Q[2]))
        e2, _ = tet_energy(P, np.tile(R @ m, (4, 1)), np.zeros(4), prm2, np.zeros((5, 3)))
        assert e2[2] == pytest.approx(e[2], rel=1e-12, abs=0)
    prm = region(K3=4.2e4)
    assert tet_energy(P, np.tile([0, 1.0, 0], (4, 1)), np.zeros(4), prm, np.zeros((5, 3)))[0][2] == 0.0
    e, _ = tet_energy(P, np.tile(np.ones(3) / math.sqrt(3), (4, 1)), np.zeros(4), prm, np.zeros((5, 3)))
    assert e[2] == pytest.approx(4.2e4 * Vol / 3, rel=1e-14, abs=0)


# ---------------------------------------------------------------- energies
def test_uniaxial_energy_ut_energy():
    """ut_energy.cpp:18-76: the legacy MuMag_energy.cc loop -K sum_g w_g (uk . U_g)^2."""
    dr = Draw()
    u = np.array([rand_vec3d(math.pi * dr(), 2 * math.pi * dr()) for _ in range(4)])
    K = dr()
    uk = rand_vec3d(math.pi * dr(), 2 * math.pi * dr())
    e, w = tet_energy(UNIT_TET, u, np.zeros(4), region(K=K, uk=tuple(uk)), np.zeros((5, 3)))
    U = A_TET.T @ u
    ref = 0.0
    for g in range(5):
        al0 = uk[0] * U[g, 0] + uk[1] * U[g, 1] + uk[2] * U[g, 2]
        ref += w[g] * (-K * al0 * al0)
    assert e[1] == pytest.approx(ref, rel=UT_TOL * 4, abs=0)


def test_volume_energies_closed_forms():
    """Exact P1 integrals on a random tetrahedron with random nodal u, phi: uniaxial
    -K sum_ij s_i s_j M_ij (s = uk . u_i), demag -1/2 mu0 Ms div(u) sum_ij M_ij phi_j... (the
    5-point rule integrates quadratics exactly), exchange A V |grad u|^2, Zeeman of a linear field
    H(x) = G x: -mu0 Ms sum_ij u_i . G x_j M_ij."""
    rng = np.random.default_rng(21)
    for _ in range(5):
        P = rng.normal(size=(4, 3)) * 3e-9
        if np.linalg.det(P[1:] - P[0]) < 0:
            P[[2, 3]] = P[[3, 2]]
        M, Vol = tet_mass(P)
        u = rng.normal(size=(4, 3))
        phi = rng.normal(size=4)
        uk = rng.normal(size=3)
        uk /= np.linalg.norm(uk)
        G = rng.normal(size=(3, 3)) * 1e13
        rc, ind, da, w = geometry(P)
        Xg = A_TET.T @ P  # Gauss points
        Hg = Xg @ G.T
        prm = region(K=2.3e4, uk=tuple(uk), Ae=1.3e-11, Ms=8e5)
        e, _ = tet_energy(P, u, phi, prm, Hg)
        s = u @ uk
        assert e[1] == pytest.approx(-2.3e4 * s @ M @ s, rel=1e-12, abs=0)
        grad = da.T @ u  # grad[k] = sum_i da[i, k] u_i  (rows: d/dx_k)
        assert e[0] == pytest.approx(1.3e-11 * Vol * np.sum(grad ** 2), rel=1e-12, abs=0)
        divu = np.trace(grad)
        assert e[3] == pytest.approx(-0.5 * MU0 * 8e5 * divu * np.sum(M @ phi), rel=1e-12, abs=0)
        Hn = P @ G.T  # the linear field at the nodes
        assert e[4] == pytest.approx(-MU0 * 8e5 * np.sum(u * (M @ Hn)), rel=1e-12, abs=0)


def test_zeeman_varying_field_ut_energy():
    """ut_energy.cpp:170-196: the density of the varying-field Zeeman term is the column-wise dot
    product of u and Hext at each Gauss point; with u uniform the energy is -mu0 Ms sum_g w_g u . H_g
    for the field drawn per Gauss point in the test's order."""
    dr = Draw()
    Hext = np.zeros((3, 5))
    uu = np.zeros((3, 5))
    for i in range(3):
        for j in range(5):
            Hext[i, j] = dr()
            uu[i, j] = dr()
    m = uu[:, 0]
    e, w = tet_energy(UNIT_TET, np.tile(m, (4, 1)), np.zeros(4), region(Ms=2.0), Hext.T)
    ref = -MU0 * 2.0 * sum(w[g] * np.dot(Hext[:, g], m) for g in range(5))
    assert e[4] == pytest.approx(ref, rel=1e-15, abs=0)


def test_triangle_energies_closed_forms():
    """Surface anisotropy -Ks sum_ij s_i s_j M_ij (s = uk . u_i) and surface demag
    1/2 mu0 dMs sum_ij (u_i . n) M_ij phi_j: the 4-point rule integrates quadratics exactly."""
    rng = np.random.default_rng(22)
    for _ in range(5):
        P = rng.normal(size=(3, 3)) * 5e-9
        M, A = tri_mass(P)
        n = np.cross(P[1] - P[0], P[2] - P[0])
        n /= np.linalg.norm(n)
        u = rng.normal(size=(3, 3))
        phi = rng.normal(size=3)
        uk = rng.normal(size=3)
        uk /= np.linalg.norm(uk)
        e = tri_energy(P, u, phi, 6.1e5, 1.7e-3, uk)
        s = u @ uk
        assert e[0] == pytest.approx(-1.7e-3 * s @ M @ s, rel=1e-12, abs=0)
        assert e[1] == pytest.approx(0.5 * MU0 * 6.1e5 * (u @ n) @ M @ phi, rel=1e-12, abs=0)


# ---------------------------------------------------------------- charges
def test_tet_charges_ut_tet_charges():
    """ut_tet_charges.cpp:18-58: -Ms w_j (dudx_x + dudy_y + dudz_z) on the unit tet."""
    dr = Draw()
    u = np.array([rand_vec3d(math.pi * dr(), 2 * math.pi * dr()) for _ in range(4)])
    Ms = dr()
    rc, ind, da, w = geometry(UNIT_TET)
    q = np.zeros(5)
    oracle.lib().orc_tet_charges(oracle.d(c64(da)), oracle.d(w), oracle.d(c64(u)), Ms, oracle.d(q))
    dudx = u.T @ da[:, 0]
    dudy = u.T @ da[:, 1]
    dudz = u.T @ da[:, 2]
    ref = -Ms * w * (dudx[0] + dudy[1] + dudz[2])
    assert np.allclose(q, ref, rtol=UT_TOL * 4, atol=0)


def test_tri_charges_ut_tri_charges():
    """ut_tri_charges.cpp:18-53: dMs w_g (U_g . n) on dummyNodes<3>, n = (0, 0, -1)."""
    dr = Draw()
    u = np.array([rand_vec3d(math.pi * dr(), 2 * math.pi * dr()) for _ in range(3)])
    dMs = dr()
    q = np.zeros(4)
    oracle.lib().orc_tri_charges(oracle.d(c64(TRI3)), oracle.d(c64(u)), dMs, oracle.d(q))
    n = np.array([0.0, 0.0, -1.0])
    w = 2 * 0.5 * PDS_TRI
    Ug = A_TRI.T @ u
    ref = dMs * w * (Ug @ n)
    assert np.allclose(q, ref, rtol=UT_TOL * 4, atol=1e-17)
    q0 = np.zeros(4)
    oracle.lib().orc_tri_charges(oracle.d(c64(TRI3)), oracle.d(c64(u)), 0.0, oracle.d(q0))
    assert not q0.any()


# ---------------------------------------------------------------- surface anisotropy RHS
def test_tri_integrales_closed_form():
    """Tri::integrales (src/triangle.cpp:6-36) against the exact P1 integral
    BE_i = Kbis uk sum_j M_ij (uk . u_j), Kbis = 2 Ks / dMs, and Lp = [eq_i . BE_i ; ep_i . BE_i]
    (the reference's Perm puts the e_q-tested rows first), for random nodal u and bases."""
    rng = np.random.default_rng(23)
    for _ in range(5):
        P = rng.normal(size=(3, 3)) * 4e-9
        M, A = tri_mass(P)
        u = rng.normal(size=(3, 3))
        ep = rng.normal(size=(3, 3))
        eq = rng.normal(size=(3, 3))
        uk = rng.normal(size=3)
        uk /= np.linalg.norm(uk)
        Ks, dMs = 2.5e-3, -8e5
        Lp = np.zeros(6)
        oracle.lib().orc_tri_integrales(oracle.d(c64(P)), oracle.d(c64(u)), oracle.d(c64(ep)), oracle.d(c64(eq)),
                                        dMs, Ks, oracle.d(c64(uk)), oracle.d(Lp))
        Kbis = 2 * Ks / dMs
        BE = Kbis * np.outer(M @ (u @ uk), uk)
        assert np.allclose(Lp[:3], np.sum(eq * BE, axis=1), rtol=1e-12, atol=1e-12 * np.abs(BE).max())
        assert np.allclose(Lp[3:], np.sum(ep * BE, axis=1), rtol=1e-12, atol=1e-12 * np.abs(BE).max())
        # uniform u: the lumped form Kbis (uk . u) uk A / 3
        m = rng.normal(size=3)
        oracle.lib().orc_tri_integrales(oracle.d(c64(P)), oracle.d(c64(np.tile(m, (3, 1)))), oracle.d(c64(ep)),
                                        oracle.d(c64(eq)), dMs, Ks, oracle.d(c64(uk)), oracle.d(Lp))
        be = Kbis * np.dot(uk, m) * uk * A / 3
        assert np.allclose(Lp[3:], ep @ be, rtol=1e-12, atol=1e-14 * np.abs(be).max())
