"""Autotuning of the FMM tiling: the leaf occupancy target (sources per leaf) decides the octree
depth, hence the balance between the near-field operator (bytes grow with the leaf size) and the
multipole evaluations (work grows with the depth). The best value depends on the geometry, so it is
measured: each candidate builds a context and times the demag pass of the actual meshes on the GPU.
"""
import time

import numpy as np


def tune_fmm_leaf(make_engine, u0, candidates=(32, 64, 128, 256), calls=3):
    """make_engine(fmm_leaf) -> Engine (single or batch) with the state space of u0.
    Returns (best leaf, {leaf: ms per demag call})."""
    from feellgood_amd import native as nv
    times = {}
    for leaf in candidates:
        g = make_engine(leaf)
        g.set_state(nv.FG_BOTH, u=u0)
        g.calc_demag()
        g.synchronize()
        t0 = time.perf_counter()
        for _ in range(calls):
            g.calc_demag()
        g.synchronize()
        times[leaf] = (time.perf_counter() - t0) * 1e3 / calls
        g.close()
    best = min(times, key=times.get)
    return best, times


def random_unit(rng, n):
    u = rng.normal(size=(n, 3))
    return u / np.linalg.norm(u, axis=1, keepdims=True)
