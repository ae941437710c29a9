"""The synthetic workloads of the benchmark and the full-size tests: meshes, materials, initial
states, applied fields and time steps, all seeded.

* cube  -- 100 nm cube, 20 x 20 x 20 cells x 6 tetrahedra (48,000 tets, 9,261 nodes), Ms 8e5 A/m,
           A 1.3e-11 J/m, alpha 0.5, no anisotropy, no field, m = normalize((1,1,0)/sqrt2 +
           U(-0.05, 0.05)^3) (seed 1), dt 1e-14 s, 200 steps, tolerance 1e-10, FMM order 8.
* film  -- 1000 x 500 x 20 nm, 333 x 167 x 6 cells x 6 tets (2,001,996 tets, 392,784 nodes),
           interior nodes jittered by +/-10 % of the cell (seed 2), alpha 0.02, S-state
           m = (cos p, sin p, 0), p = 0.3 sin(pi x / L), B = 25 mT at 170 deg in the plane,
           dt 5e-14 s, 2,000 steps.
* guess -- the film under seeded applied fields of 0-50 mT (seed 3): training directions and
           held-out directions for the learned initial guess.
* dots  -- 256 ellipsoidal nanodots, semi-axes U[10,60] x U[10,60] x U[3,20] nm (seed 4), per dot
           Ms U[6e5, 1.4e6], uniaxial K U[0, 5e4] J/m^3 along a random axis, alpha 0.1, random unit
           m, dt 1e-14 s, 500 steps; about 1.2 M tets in total.
* wire  -- cylinder r 50 nm, L 4 um (about 10 M tets), +z with a reversed 200 nm end domain
           (seed 5), alpha 0.05, field along -z ramped 0 -> 60 mT over 5,000 steps at dt 2e-14 s.

The meshes come from the host library's generators (the 6-tetrahedron split of the cuboid, the
ellipsoid and cylinder meshers).
"""
import math

import numpy as np

MU0 = 1.25663706127e-6


def cube_workload():
    return dict(name="cube", lengths=(100.0, 100.0, 100.0), cells=(20, 20, 20), jitter=0.0, seed=1,
                Ms=8e5, Ae=1.3e-11, alpha=0.5, B=0.0, angle=0.0, dt=1e-14, tol=1e-10, steps=200)


def film_workload():
    return dict(name="film", lengths=(1000.0, 500.0, 20.0), cells=(333, 167, 6), jitter=0.1, seed=2,
                Ms=8e5, Ae=1.3e-11, alpha=0.02, B=0.025, angle=170.0, dt=5e-14, tol=1e-10, steps=2000)


def wire_workload():
    return dict(name="wire", radius=50.0, length=4000.0, h=2.66, seed=5, Ms=8e5, Ae=1.3e-11, alpha=0.05,
                dt=2e-14, tol=1e-10, reversed=200.0, ramp_T_per_s=-0.06 / 1e-10, steps=5000)


def dots_workload():
    return dict(name="dots", count=256, seed=4, h=4.5, dt=1e-14, tol=1e-10, alpha=0.1, B=0.0, angle=0.0,
                steps=500)


def build_mesh(w):
    from feellgood_amd.mesh import HostMesh
    if w["name"] == "wire":
        m = HostMesh.cylinder((0.0, 0.0), 0.0, w["radius"], w["length"], w["h"])
        m.prepare({"volume": dict(Ms=w["Ms"], Ae=w["Ae"], alpha=w["alpha"])})
        return m
    m = HostMesh.cuboid(w["lengths"], w["cells"], jitter=w["jitter"], seed=w["seed"])
    m.prepare({"volume": dict(Ms=w["Ms"], Ae=w["Ae"], alpha=w["alpha"])})
    return m


def build_dots(w, count=None):
    """Ellipsoidal nanodots: semi-axes U[10,60] x U[10,60] x U[3,20] nm (seed 4), a grid of
    spacing h split into 6 tetrahedra per cell with the boundary projected on the ellipsoid; per dot
    Ms U[6e5, 1.4e6] A/m, A = 1.3e-11 J/m, uniaxial K U[0, 5e4] J/m^3 along a random axis,
    alpha = 0.1, random unit initial m. count < w['count'] builds the first dots of the same
    sequence. Returns (meshes, u0 of the union in instance order)."""
    from feellgood_amd.mesh import HostMesh
    rng = np.random.default_rng(w["seed"])
    meshes, us = [], []
    for k in range(w["count"] if count is None else count):
        a, b, c = rng.uniform(10, 60), rng.uniform(10, 60), rng.uniform(3, 20)
        m = HostMesh.ellipsoid((0.0, 0.0, 0.0), (a, b, c), w["h"])
        uk = rng.normal(size=3)
        m.prepare({"volume": dict(Ms=rng.uniform(6e5, 1.4e6), Ae=1.3e-11, alpha=w["alpha"], K=rng.uniform(0, 5e4),
                                  uk=tuple(uk / np.linalg.norm(uk)))})
        u = rng.normal(size=(m.n_nodes, 3))
        meshes.append(m)
        us.append(u / np.linalg.norm(u, axis=1, keepdims=True))
    return meshes, np.concatenate(us)


def initial_state(m, w):
    pos = m.arrays()["pos"]
    if w["name"] == "wire":
        # +z, a reversed domain over the first `reversed` nm, a seeded +/-0.02 perturbation
        rng = np.random.default_rng(w["seed"])
        u = np.zeros((m.n_nodes, 3))
        u[:, 2] = np.where(pos[:, 2] < w["reversed"] * 1e-9, -1.0, 1.0)
        u += rng.uniform(-0.02, 0.02, size=u.shape)
        return u / np.linalg.norm(u, axis=1, keepdims=True)
    if w["name"] == "film":
        L = w["lengths"][0] * 1e-9
        ph = 0.3 * np.sin(np.pi * pos[:, 0] / L)
        return np.stack([np.cos(ph), np.sin(ph), np.zeros_like(ph)], axis=1)
    rng = np.random.default_rng(w["seed"])
    u = np.array([1.0, 1.0, 0.0]) / math.sqrt(2) + rng.uniform(-0.05, 0.05, size=(m.n_nodes, 3))
    return u / np.linalg.norm(u, axis=1, keepdims=True)


def field(w, t=0.0):
    """Applied field H = B / mu0 [A/m] at time t."""
    if "ramp_T_per_s" in w:
        return np.array([0.0, 0.0, w["ramp_T_per_s"] * t]) / MU0
    a = math.radians(w["angle"])
    return np.array([math.cos(a), math.sin(a), 0.0]) * w["B"] / MU0


def prefactor(dt):
    """timing::set_dt with dt_min = dt_max = dt (src/time_integration.h:37-42)."""
    tt = dt / (100.0 * dt)
    return 1.0 + tt * abs(math.log(tt))


def guess_field(rng, bmax=0.05):
    """Seeded applied field of the learned-guess workload: in-plane angle U(0, 2 pi), out-of-plane
    tilt U(-0.2, 0.2), |B| U(0, bmax) T; H = B / mu0."""
    th = rng.uniform(0, 2 * math.pi)
    tilt = rng.uniform(-0.2, 0.2)
    b = rng.uniform(0.0, bmax)
    d = np.array([math.cos(th), math.sin(th), tilt])
    return d / np.linalg.norm(d) * b / MU0


def guess_fields(train=32, held_out=8, seed=3):
    """The 32 training and 8 held-out applied fields of the learned-guess workload."""
    rng = np.random.default_rng(seed)
    tr = [guess_field(rng) for _ in range(train)]
    te = [guess_field(rng) for _ in range(held_out)]
    return tr, te


def workload_text(w, m):
    if w["name"] == "wire":
        return (f"wire r {w['radius']:g} nm, L {w['length']:g} nm, h {w['h']:g} nm ({m.n_tets} tets, {m.n_nodes} "
                f"nodes, {m.n_tris} surface triangles), +z with a reversed {w['reversed']:g} nm end, field ramp "
                f"{w['ramp_T_per_s'] * 1e-10:g} T per 0.1 ns along z, dt {w['dt']:g} s, tol {w['tol']:g}")
    return (f"{w['name']} {w['lengths'][0]:g}x{w['lengths'][1]:g}x{w['lengths'][2]:g} nm, "
            f"{w['cells'][0]}x{w['cells'][1]}x{w['cells'][2]} cells x 6 tets "
            f"({m.n_tets} tets, {m.n_nodes} nodes, {m.n_tris} surface triangles), "
            f"jitter {w['jitter']}, dt {w['dt']:g} s, tol {w['tol']:g}")


def guess_weights_path(name):
    """The committed parameters of the learned initial guess trained for a workload."""
    import os
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", f"{name}_guess.npy")
