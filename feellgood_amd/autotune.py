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


def tune_batch_block(make_engine, u0, H, dt, prefactor, candidates=(256, 512, 1024), calls=3):
    """make_engine(batch_block) -> batch Engine with the persistent solver. Times fg_solve (one
    launch solving every instance) from the same assembled system for each block size; returns
    (best block, {block: ms per solve})."""
    from feellgood_amd import native as nv
    times = {}
    for blk in candidates:
        g = make_engine(blk)
        g.set_state(nv.FG_BOTH, u=u0)
        g.calc_demag()
        g.evolution()
        g.base_projection(0.3)
        g.prepare_elements(H, dt, prefactor)
        g.solve(dt)
        g.synchronize()
        t0 = time.perf_counter()
        for _ in range(calls):
            failed, _ = g.solve(dt)
            if failed:
                raise RuntimeError("solver failure while tuning the batch block")
        g.synchronize()
        times[blk] = (time.perf_counter() - t0) * 1e3 / calls
        g.close()
    best = min(times, key=times.get)
    return best, times
