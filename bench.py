#!/usr/bin/env python3
"""LLG steps/s of the gfx950 engine on the 2M-tetrahedron film (one MI355X).

Workload (the configuration the headline metric is quoted on): thin film 1000 x 500 x 20 nm,
333 x 167 x 6 cells split into 6 tetrahedra each (2,001,996 tets, 392,784 nodes), interior node
coordinates jittered by +/-10% of the cell size (seed 2); Ms = 8e5 A/m, A = 1.3e-11 J/m,
alpha = 0.02; S-state m = (cos p, sin p, 0), p = 0.3 sin(pi x / L); B = 25 mT at 170 deg in the
plane; fixed dt = 5e-14 s; BiCGSTAB tolerance 1e-10.

One step = the reference's hot loop body (src/time_integration.cpp:193-244): base_projection,
prepareElements, solve (assembly + BiCGSTAB + update), compute_all (demag u and v, energies,
evolution) and max_angle. Inputs are resident in HBM before the timed region.

The measured configuration (--mode, default learned+mixed: the trained initial-guess network and
the mixed-precision FMM pass) is timed beside the plain FP64 engine on the same steps; the JSON
line carries both rates, their ratio and the per-step |dm| of the fast mode against plain FP64.

Prints one JSON line (see the keys below). Options: --gpus N (only 1 is supported), --steps K,
--warmup W, --mode, --demag {fmm,direct}, --guess weights.npy, --no-cpu, --small (48k-tet cube).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from feellgood_amd.workloads import (MU0, build_dots, build_mesh, cube_workload, dots_workload, field,  # noqa: E402
                                     film_workload, initial_state, prefactor, wire_workload, workload_text)

METRIC = "LLG steps/s at 2M tets: learned+mixed vs plain FP64 GPU at |dm|<=1e-6"
HBM_PEAK_GBS = 8000.0
FP64_VALU_TFLOPS = 78.6
FP32_VALU_TFLOPS = 157.3


def iters_text(its):
    """The per-step BiCGSTAB iteration counts, listed when few, else as counts of each value."""
    if len(its) <= 20:
        return str(list(its))
    vals, cnt = np.unique(np.asarray(its), return_counts=True)
    return "{" + ", ".join(f"{v}: {c} steps" for v, c in zip(vals, cnt)) + "}"


def bench_dots(args):
    """Batched nanodots in one context: dot-steps/s and tet-steps/s, mixed precision FMM demag with
    the leaf size picked by the autotuner."""
    from feellgood_amd import native as nv
    from feellgood_amd.autotune import tune_fmm_leaf
    from feellgood_amd.engine import Engine, options
    w = dots_workload()
    t0 = time.perf_counter()
    meshes, u0 = build_dots(w)
    T = sum(m.n_tets for m in meshes)
    print(f"[bench] {len(meshes)} dots, {T} tets, {u0.shape[0]} nodes built in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    prec = "fp64" if args.mode == "plain" else "mixed"

    block = 0

    def make(leaf, profile=False):
        return Engine.batch([m.mesh for m in meshes], options(tol=w["tol"], max_iter=10000, demag="fmm",
                                                              fmm_order=args.fmm_order, fmm_leaf=leaf,
                                                              fmm_far_order=args.fmm_far_order,
                                                              precision=prec, profile=profile,
                                                              batch_solver=args.batch_solver, batch_block=block))
    leaf, tune = tune_fmm_leaf(make, u0)
    print(f"[bench] autotuned fmm_leaf {leaf}: {tune}", file=sys.stderr, flush=True)
    H = np.zeros(3)
    dt, pref = w["dt"], prefactor(w["dt"])
    tune_block = None
    if args.batch_solver == "persistent":
        from feellgood_amd.autotune import tune_batch_block

        def make_blk(b):
            nonlocal block
            block = b
            return make(leaf)
        block, tune_block = tune_batch_block(make_blk, u0, H, dt, pref)
        print(f"[bench] autotuned batch block {block}: {tune_block}", file=sys.stderr, flush=True)
    g = make(leaf)
    rng = np.random.default_rng(w["seed"])
    g.set_state(nv.FG_BOTH, u=u0)
    g.calc_demag()
    g.energy(H)
    g.evolution()
    iters = []

    def step():
        rc, info, _, _ = g.step((2 / math.pi) * rng.uniform(), H, dt, pref)
        if rc != 0:
            raise RuntimeError("solver failure")
        iters.append(info.iterations)
    for _ in range(args.warmup):
        step()
    iters.clear()
    g.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step()
    g.synchronize()
    ms = (time.perf_counter() - t1) * 1e3 / args.steps
    g.close()
    out = {"metric": "batched nanodot steps: dot-steps/s", "value": round(len(meshes) * 1e3 / ms, 2),
           "unit": "dot-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64" if prec == "fp64" else "f32+f64", "data": "synthetic",
           "config": {"workload": f"{len(meshes)} ellipsoidal dots, {T} tets, h {w['h']} nm, dt {w['dt']:g} s",
                      "mode": prec, "fmm_leaf": leaf, "autotune_ms_per_demag": tune,
                      "batch_solver": args.batch_solver, "batch_block": block if tune_block else None,
                      "autotune_ms_per_solve": tune_block},
           "tet_steps_per_s": round(T * 1e3 / ms, 1),
           "bicgstab_iterations_per_step": round(float(np.mean(iters)), 2)}
    print(json.dumps(out), flush=True)


# ------------------------------------------------------------------ byte / flop models
def m2p_flops(p):
    """FP64 (or FP32) operations of one multipole-to-target evaluation of two expansions of degree
    p, as m2p_eval2_r performs them: setup (r^2, refined 1/r^2 and 1/r), the real m = 0 column, the
    complex columns m >= 1 (diagonal step, first off-diagonal, three-term recurrence) with four
    fused multiply-adds per term, and the caller's offset and scaling."""
    f = 25                                              # setup
    f += 2 + (5 if p >= 1 else 0) + 10 * max(0, p - 1)  # m = 0 column
    for m in range(1, p + 1):
        f += 18 + (15 if m < p else 0) + 17 * max(0, p - m - 1)
    return f + 14


V_DEGREE = 4  # degree of the far field of the charges of v in the mixed-precision mode


def m2p_flops_v(p, pv):
    """m2p_flops with the second expansion (charges of v) taken to degree pv <= p only: its fused
    multiply-adds of degrees above pv (one per real m = 0 term, two per complex term) drop out."""
    return m2p_flops(p) - sum(2 + 4 * n for n in range(pv + 1, p + 1))


FAR_DEGREE = {4: 2, 6: 4, 8: 5, 10: 6}  # lower degree of the far list entries, per order


def far_degree(p, far_order, precision):
    """Degree the far list entries are evaluated to (fg_options.fmm_far_order semantics)."""
    if far_order:
        return far_order
    return FAR_DEGREE[p] if precision == "mixed" else p


def m2l_flops(p):
    """FP64 matrix-core operations of one M2L translation of both charge sets: the dense real
    operator of the offset (2 fcount(p) rows and columns) times each set's real coefficients."""
    nr = (p + 1) * (p + 2)
    return 2 * 2 * nr * nr


def xlate_ksteps(p, kind):
    """Sum over the 16-row tiles of the k-steps the M2M (kind 0) or L2L (kind 1) operator tiles run
    (the degree coupling of each tile, as k_fmm_xlate takes it)."""
    nco = (p + 1) * (p + 2) // 2
    nr, nks = 2 * nco, (2 * nco + 3) // 4
    fidx = lambda n, m: n * (n + 1) // 2 + m

    def deg(j):
        n = 0
        while fidx(n + 1, 0) <= (j >> 1):
            n += 1
        return n
    tot = 0
    for r in range((nr + 15) // 16):
        if kind == 0:
            tot += (2 * ((deg(min(16 * r + 15, nr - 1)) + 1) * (deg(min(16 * r + 15, nr - 1)) + 2) // 2) + 3) // 4
        else:
            tot += nks - (2 * fidx(deg(16 * r), 0)) // 4
    return tot


def m2m_flops(p):
    """FP64 matrix-core operations of one parent's M2M, both charge sets: for each of 8 child octants,
    every 16-row operator tile times the parent's 16-column tile over its k-steps (16x16x4 MFMAs,
    2 x 1024 flops each, per 16 parents)."""
    return 2 * 8 * xlate_ksteps(p, 0) * 2 * 16 * 16 * 4 // 16


def p2m_flops(p):
    """Operations of accumulating one source into both expansions: its charges from the element
    values, the regular recurrence once, and four fused multiply-adds per complex coefficient."""
    rec = 5 + 7 * p + 2 * p + 9 * sum(max(0, p - m - 1) for m in range(p + 1))
    return 4 + rec + 8 * (p + 1) * (p + 2) // 2


def kernel_model(name, n):
    """(bound, algorithmic bytes or flops per launch) of the engine's kernels; n holds the sizes."""
    N, T, nnz, F, S = n["N"], n["T"], n["nnz"], n["F"], n["S"]
    models = {
        "spmv": ("hbm", 12 * nnz + 180 * N),
        "spmv_r0": ("hbm", 12 * nnz + 180 * N),
        "iter_p": ("hbm", 192 * N),
        "iter_s": ("hbm", 176 * N),
        "iter_x": ("hbm", 128 * N),
        "node_assemble": ("hbm", 332 * N + 144 * T),
        "element": ("hbm", 284 * T + 64 * N),
        "update": ("hbm", 164 * N),
        "evolution": ("hbm", 128 * N),
        "charges_tet": ("hbm", 16 * T + 96 * T + 4 * T + 80 * T + 64 * N),
        "charges_tri": ("hbm", 16 * F + 128 * F + 64 * F + 48 * F + 96 * N),
        "corr_gather": ("hbm", 8 * N + 4 * 3 * F + 16 * 3 * F + 16 * N),
        "energy": ("hbm", 16 * T + 136 * T + 4 * T + 64 * N),
        "max_angle": ("hbm", 4 * nnz + 64 * N),
        "p2p_direct": ("valu64", 2 * 10.0 * N * S),
    }
    f = n.get("fmm")
    if f:
        p, Tg = f["order"], f["targets"]
        mixed = n.get("precision") == "mixed"
        pf = far_degree(p, n.get("far_order"), n.get("precision"))
        pv = V_DEGREE if mixed else p
        models.update({
            # coefficient rows (FP32 in the mixed mode) + shared columns + element values read once
            # + per-target I/O
            "fmm_near": ("hbm", (4 if mixed else 8) * f["near_coefs"] + 4 * f["near_columns"] + 16 * (T + 3 * F)
                         + (8 + 16 + 16 + 4) * Tg + 12 * f["target_leaves"]),
            "fmm_m2p": ("valu32" if mixed else "valu64",
                        m2p_flops_v(p, min(p, pv)) * (f["m2p_pairs"] - f["m2p_far_pairs"])
                        + m2p_flops_v(pf, min(pf, pv)) * f["m2p_far_pairs"]),
            "fmm_m2l": ("mfma32" if mixed else "mfma64", m2l_flops(p) * f["m2l_pairs"]),
            "fmm_m2m": ("mfma32" if mixed else "mfma64", m2m_flops(p) * f["m2m_cells"]),
            "fmm_p2m": ("valu32" if mixed else "valu64", p2m_flops(p) * f["sources"]),
        })
    return models.get(name)


def make_engine(args, w, m, precision, guess, profile):
    from feellgood_amd.engine import Engine, options
    # the Krylov mode applies to the measured mixed modes; the plain FP64 engine keeps FP64 bicg_dir
    krylov = getattr(args, "krylov", "fp64") if precision == "mixed" else "fp64"
    opts = options(tol=w["tol"], max_iter=10000, demag=args.demag, fmm_order=args.fmm_order,
                   fmm_leaf=args.fmm_leaf, precision=precision, profile=profile, fmm_far_order=args.fmm_far_order,
                   krylov=krylov, krylov_inner_tol=getattr(args, "krylov_inner_tol", 0.0))
    g = Engine(m.mesh, opts)
    if guess is not None:
        g.guess_set(guess)
        g.guess_enable(True)
    return g


class Stepper:
    """The reference loop body at fixed dt on one engine, with its own basis-angle stream."""

    def __init__(self, g, w, u0):
        from feellgood_amd import native as nv
        self.g, self.w, self.H, self.dt, self.t = g, w, field(w), w["dt"], 0.0
        self.pref = prefactor(self.dt)
        self.rng = np.random.default_rng(w["seed"])
        self.iters, self.fails = [], 0
        # compute_all at t0 (src/time_integration.cpp:136)
        g.set_state(nv.FG_BOTH, u=u0)
        g.calc_demag()
        g.energy(self.H)
        g.evolution()

    def angle(self):
        return (2 / math.pi) * self.rng.uniform()

    def step(self, r=None):
        """base_projection, prepare_elements, solve, calc_demag, energy, evolution, max_angle: one
        fg_step call (the same kernels as the separate entry points, one host read at the end)."""
        from feellgood_amd import native as nv
        r = self.angle() if r is None else r
        self.H = field(self.w, self.t)  # time-dependent for the wire's ramp
        rc, info, _, _ = self.g.step(r, self.H, self.dt, self.pref)
        if rc == nv.FG_SOLVE_FAILED:
            self.fails += 1
            return info
        self.t += self.dt
        self.iters.append(info.iterations)
        return info


def run_gpu(args, w, m, u0, precision, guess):
    """Clean timing of K steps after W warmup steps, then a profiled engine for per-kernel times."""
    g = make_engine(args, w, m, precision, guess, False)
    st = Stepper(g, w, u0)
    for _ in range(args.warmup):
        st.step()
    st.iters.clear()
    g.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st.step()
    g.synchronize()
    t1 = time.perf_counter()
    clean_ms = (t1 - t0) * 1e3 / args.steps
    clean_iters, fails = list(st.iters), st.fails
    g.close()
    prof_steps = max(1, min(args.steps, args.prof_steps))
    g = make_engine(args, w, m, precision, guess, True)
    st = Stepper(g, w, u0)
    st.step()
    g.reset_kernel_stats()
    g.synchronize()
    tp0 = time.perf_counter()
    for _ in range(prof_steps):
        st.step()
    g.synchronize()
    tp1 = time.perf_counter()
    stats = g.kernel_stats()
    fmm = g.fmm_info() if args.demag == "fmm" else None
    g.close()
    return dict(ms_per_step=clean_ms, iters=clean_iters, fails=fails, stats=stats, fmm=fmm,
                prof_ms_per_step=(tp1 - tp0) * 1e3 / prof_steps, prof_steps=prof_steps)


def step_accuracy(args, w, m, u0, precision, guess, steps=2):
    """Per-step accuracy, untimed. The exact engine (FP64 direct-sum demag, pinned against the oracle)
    steps from the initial state; before every step the plain FP64 FMM engine and the measured engine
    are put on its u, v (each with the demag potentials of its own method), then all three take the
    step with the same basis angle. Returns the max over `steps` steps of |dm| of each FMM engine
    against the exact one and of the measured engine against the plain one, and the relative errors
    of their demag energies."""
    from feellgood_amd import native as nv
    ex_args = argparse.Namespace(**vars(args))
    ex_args.demag = "direct"
    ex = make_engine(ex_args, w, m, "fp64", None, False)
    ref = make_engine(args, w, m, "fp64", None, False)
    fast = make_engine(args, w, m, precision, guess, False)
    a, b, c = Stepper(ex, w, u0), Stepper(ref, w, u0), Stepper(fast, w, u0)
    out = dict(dm_plain_vs_exact=0.0, dm_vs_exact=0.0, dm_vs_plain=0.0, edemag_plain_vs_exact=0.0,
               edemag_vs_exact=0.0)
    H = field(w)
    for _ in range(steps):
        u, v, _, _ = ex.get_state(nv.FG_CURRENT)
        for g in (ref, fast):
            g.set_state(nv.FG_BOTH, u=u, v=v)
            g.calc_demag()
            g.evolution()
        r = a.angle()
        for st in (a, b, c):
            st.step(r)
        un = {k: g.get_state(nv.FG_CURRENT)[0] for k, g in (("ex", ex), ("ref", ref), ("fast", fast))}
        # fg_energy's third component: the demag energy
        E = {k: g.energy(H)[2] for k, g in (("ex", ex), ("ref", ref), ("fast", fast))}
        out["dm_plain_vs_exact"] = max(out["dm_plain_vs_exact"], float(np.abs(un["ref"] - un["ex"]).max()))
        out["dm_vs_exact"] = max(out["dm_vs_exact"], float(np.abs(un["fast"] - un["ex"]).max()))
        out["dm_vs_plain"] = max(out["dm_vs_plain"], float(np.abs(un["fast"] - un["ref"]).max()))
        out["edemag_plain_vs_exact"] = max(out["edemag_plain_vs_exact"], abs(E["ref"] / E["ex"] - 1))
        out["edemag_vs_exact"] = max(out["edemag_vs_exact"], abs(E["fast"] / E["ex"] - 1))
    for g in (ex, ref, fast):
        g.close()
    return {k: float(f"{v:.3e}") for k, v in out.items()}


def measured_traffic(kernel, args):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE scaled by a copy kernel), when it describes this configuration: the
    FP64 far field for --mode plain, the FP32 far field at the lower far degree for the mixed modes
    (the learned guess does not change that kernel)."""
    name = "film_fmm_fp64_traffic.json" if args.mode == "plain" else "film_learned_mixed_traffic.json"
    if (args.small or args.demag != "fmm" or args.fmm_order != 8 or args.fmm_leaf != 64 or args.fmm_far_order != 0
            or getattr(args, "dots", False) or getattr(args, "wire", False)):
        return None
    try:
        d = json.load(open(os.path.join(ROOT, "profiles", name)))
    except (OSError, ValueError):
        return None
    return d.get("bytes_per_launch") if d.get("kernel") == kernel else None


def roofline(stats, sizes):
    best = None
    for name, (launches, ms) in stats.items():
        if launches == 0:
            continue
        if best is None or ms > best[2]:
            best = (name, launches, ms)
    if best is None:
        return None, {}
    per_kernel = {}
    for name, (launches, ms) in stats.items():
        mdl = kernel_model(name, sizes)
        entry = {"launches": launches, "total_ms": round(ms, 4), "avg_us": round(1e3 * ms / max(launches, 1), 3)}
        if mdl:
            bound, work = mdl
            avg_s = ms / 1e3 / max(launches, 1)
            if bound == "hbm":
                entry["GBps"] = round(work / avg_s / 1e9, 1)
                entry["frac_hbm"] = round(work / avg_s / 1e9 / HBM_PEAK_GBS, 4)
            else:
                pk = FP32_VALU_TFLOPS if bound.endswith("32") else FP64_VALU_TFLOPS
                entry["TFLOPs"] = round(work / avg_s / 1e12, 2)
                entry["frac_valu"] = round(work / avg_s / 1e12 / pk, 4)
        per_kernel[name] = entry
    name, launches, ms = best
    mdl = kernel_model(name, sizes)
    if mdl is None:
        return {"kernel": name, "bound": None}, per_kernel
    bound, work = mdl
    avg_s = ms / 1e3 / launches
    if bound == "hbm":
        achieved = work / avg_s / 1e9
        rl = {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
              "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "bytes_per_launch": int(work),
              "avg_launch_us": round(avg_s * 1e6, 3)}
    else:
        achieved = work / avg_s / 1e12
        pk = FP32_VALU_TFLOPS if bound.endswith("32") else FP64_VALU_TFLOPS
        rl = {"kernel": name, "bound": "mfma" if bound.startswith("mfma") else "valu", "achieved": round(achieved, 2),
              "peak": pk, "unit": "TFLOP/s", "frac": round(achieved / pk, 4), "traffic": None,
              "flops_per_launch": work, "avg_launch_us": round(avg_s * 1e6, 3),
              "precision": "fp32" if bound.endswith("32") else "fp64"}
    return rl, per_kernel


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, m, u0, fmm_order=8, fmm_leaf=64):
    """CPU restatement of FeeLLGood's algorithm (TBB path unavailable): the FP64 oracle with OpenMP on
    the host threads, one full step of the reference loop body on the whole mesh -- element pass,
    assembly, BiCGSTAB, update, the demag of u and v by the oracle's octree FMM (same order and leaf
    size as the GPU's), energies, evolution and max_angle -- after the compute_all of the start."""
    import oracle
    from feellgood_amd import native as nv
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    p = oracle.Problem(m.mesh, tol=w["tol"], max_iter=10000)
    p.set_demag(fmm_order, fmm_leaf)
    H = field(w)
    p.set_state(nv.FG_BOTH, u=u0)
    p.calc_demag()
    p.energy(H)
    p.evolution()
    t_setup = time.perf_counter() - t0
    dt = w["dt"]
    part = {}
    t1 = time.perf_counter()
    p.base_projection(0.3)
    p.prepare_elements(H, dt, prefactor(dt))
    rc, info = p.solve(dt)
    t2 = time.perf_counter()
    p.calc_demag()
    t3 = time.perf_counter()
    p.energy(H)
    p.evolution()
    p.max_angle()
    t4 = time.perf_counter()
    part = dict(elements_and_solve=t2 - t1, demag_fmm=t3 - t2, energy_evolution_max_angle=t4 - t3)
    per_step = t4 - t1
    return {"value": round(1.0 / per_step, 6), "unit": "steps/s", "cores": threads, "kind": "port",
            "label": "CPU restatement of FeeLLGood's algorithm (TBB path unavailable)",
            "cpu": cpu_model(), "nproc": os.cpu_count(),
            "sample": (f"one full step of the FP64 oracle on {threads} OpenMP threads ({cpu_model()}): element "
                       f"pass, assembly, BiCGSTAB ({info.iterations} iterations), update, the demag of u and v "
                       f"by the oracle's octree FMM (order {fmm_order}, leaves of {fmm_leaf} sources), "
                       f"energies, evolution and max_angle on the whole mesh; start-up {t_setup:.1f} s excluded"),
            "seconds_per_step": round(per_step, 3), "seconds": {k: round(v, 3) for k, v in part.items()}}


def default_guess_weights(w):
    """Trained network parameters committed for this workload, or None."""
    path = os.path.join(ROOT, "feellgood_amd", "weights", f"{w['name']}_guess.npy")
    if not os.path.exists(path):
        return None, None
    return np.load(path, allow_pickle=False).astype(np.float32), os.path.relpath(path, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000,
                    help="timed steps (default: the film's whole 2,000-step run, about 10 s of GPU time per mode)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", default="learned+mixed", choices=["learned+mixed", "mixed", "plain"],
                    help="the measured configuration; the plain FP64 engine is timed beside it")
    ap.add_argument("--demag", default="fmm", choices=["fmm", "direct"])
    ap.add_argument("--fmm-order", type=int, default=8)
    ap.add_argument("--fmm-leaf", type=int, default=64)
    ap.add_argument("--fmm-far-order", type=int, default=0,
                    help="degree of the far source cells (0: the order in FP64, the lower far degree in mixed)")
    ap.add_argument("--guess", default=None, help="network parameters (.npy); default: the committed ones")
    ap.add_argument("--prof-steps", type=int, default=3)
    ap.add_argument("--krylov", default="fp64", choices=["fp64", "mixed"],
                    help="the mixed modes' Krylov solve: FP64 bicg_dir, or FP32 bicg_dir inside FP64 refinement")
    ap.add_argument("--krylov-inner-tol", type=float, default=0.0,
                    help="relative target of each FP32 solve of --krylov mixed (0: the library's default)")
    ap.add_argument("--batch-solver", default="persistent", choices=["union", "persistent"],
                    help="--dots: one bicg_dir over the union, or one launch with a block per dot")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the plain FP64 run and the accuracy check (profiling runs of one mode)")
    ap.add_argument("--small", action="store_true", help="the 48k-tet cube instead of the film")
    ap.add_argument("--dots", action="store_true", help="the batched nanodots instead of the film")
    ap.add_argument("--wire", action="store_true",
                    help="the 10M-tet nanowire (FMM leaf size autotuned over 32..256) instead of the film")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding the path across GPUs is designed, not implemented)", file=sys.stderr)
        sys.exit(2)
    if args.dots:
        bench_dots(args)
        return
    w = cube_workload() if args.small else wire_workload() if args.wire else film_workload()
    t0 = time.perf_counter()
    m = build_mesh(w)
    u0 = initial_state(m, w)
    print(f"[bench] mesh {m.n_nodes} nodes {m.n_tets} tets {m.n_tris} tris built in "
          f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    precision = "fp64" if args.mode == "plain" else "mixed"
    tune = None
    if args.wire:
        from feellgood_amd.autotune import tune_fmm_leaf

        def make_tuned(leaf):
            args.fmm_leaf = leaf
            return make_engine(args, w, m, precision, None, False)
        args.fmm_leaf, tune = tune_fmm_leaf(make_tuned, u0)
        print(f"[bench] autotuned fmm_leaf {args.fmm_leaf}: {tune}", file=sys.stderr, flush=True)
    guess, guess_src = None, None
    if args.mode == "learned+mixed":
        if args.guess:
            guess, guess_src = np.load(args.guess, allow_pickle=False).astype(np.float32), args.guess
        else:
            guess, guess_src = default_guess_weights(w)
        if guess is None:
            # no trained weights for this workload: the untrained head, i.e. the extrapolated base
            from feellgood_amd import native as nv
            guess = np.zeros(nv.FG_GUESS_NPARAM, dtype=np.float32)
            guess[-nv.FG_GUESS_NSCALE:] = 1.0
            guess_src = "untrained head: the extrapolation of the last two solutions"
    args.precision = precision
    res = run_gpu(args, w, m, u0, precision, guess)
    print(f"[bench] {args.mode}: {res['ms_per_step']:.2f} ms/step, iterations {iters_text(res['iters'])}",
          file=sys.stderr, flush=True)
    plain = None
    if args.mode != "plain" and not args.no_compare:
        g = make_engine(args, w, m, "fp64", None, False)
        st = Stepper(g, w, u0)
        for _ in range(args.warmup):
            st.step()
        st.iters.clear()
        g.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            st.step()
        g.synchronize()
        plain = dict(ms=(time.perf_counter() - t1) * 1e3 / args.steps, iters=list(st.iters))
        g.close()
        print(f"[bench] plain fp64: {plain['ms']:.2f} ms/step, iterations {iters_text(plain['iters'])}", file=sys.stderr,
              flush=True)
        acc = step_accuracy(args, w, m, u0, precision, guess)
        dm = acc["dm_vs_plain"]
        print(f"[bench] per-step accuracy: {acc}", file=sys.stderr, flush=True)
    sizes = dict(N=m.n_nodes, T=m.n_tets, F=m.n_tris, S=5 * m.n_tets + 4 * m.n_tris)
    # stencil entries N + 2E from the engine's own count
    from feellgood_amd.engine import Engine, options
    import ctypes as C
    g = Engine(m.mesh, options(tol=w["tol"]))
    rows, nz = C.c_int64(), C.c_int64()
    g.lib.fg_system_size(g._h, C.byref(rows), C.byref(nz))
    g.close()
    sizes["nnz"] = nz.value // 4
    sizes["fmm"] = res["fmm"]
    sizes["precision"] = precision
    sizes["far_order"] = args.fmm_far_order
    rl, per_kernel = roofline(res["stats"], sizes)
    if rl and rl.get("bound"):
        rl["traffic"] = measured_traffic(rl["kernel"], args)
    mode_txt = {"plain": "plain FP64", "mixed": "mixed precision (FP32 FMM near field and M2P, FP64 elsewhere)",
                "learned+mixed": "learned initial guess + mixed precision"}[args.mode]
    out = {
        "metric": METRIC,
        "value": round(1e3 / res["ms_per_step"], 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": workload_text(w, m),
                   "mode": mode_txt,
                   "initial_guess": (f"learned ({guess_src})" if guess is not None else
                                     "previous v (reference buildInitGuess)"),
                   "demag": args.demag, "fmm_order": args.fmm_order if args.demag == "fmm" else None,
                   "krylov": args.krylov if precision == "mixed" else "fp64",
                   "krylov_inner_tol": args.krylov_inner_tol if precision == "mixed" and args.krylov == "mixed" else None,
                   "parallelism": "single GPU"},
        "bicgstab_iterations_per_step": (round(float(np.mean(res["iters"])), 2) if res["iters"] else None),
        "solver_failures": res["fails"],
        "roofline": rl,
        "kernels": per_kernel,
        "profiled_ms_per_step": round(res["prof_ms_per_step"], 4),
        "fmm_plan": res["fmm"],
    }
    if tune:
        out["config"]["fmm_leaf"] = args.fmm_leaf
        out["config"]["autotune_ms_per_demag"] = tune
    if plain:
        out["plain_fp64"] = {"steps_per_s": round(1e3 / plain["ms"], 4), "ms_per_step": round(plain["ms"], 4),
                             "bicgstab_iterations_per_step": round(float(np.mean(plain["iters"])), 2)}
        out["speedup_vs_plain_fp64"] = round(plain["ms"] / res["ms_per_step"], 4)
        out["max_step_dm_vs_plain_fp64"] = dm
        out["accuracy_vs_exact_demag"] = acc
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(w, m, u0)
        except Exception as e:  # the baseline must not hide the GPU measurement
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
