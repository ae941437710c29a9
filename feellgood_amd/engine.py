"""Python mirror of the fg_* C-ABI (include/fgmi.h): one Engine per fg_ctx.

The method names follow the reference's LinAlgebra / fmm / Fem calls that they replace
(base_projection, prepare_elements, solve, calc_demag, energy, evolution, max_angle).
"""
import ctypes as C

import numpy as np

from . import native as nv


class EngineError(RuntimeError):
    pass


def options(tol=1e-6, max_iter=700, demag="direct", fmm_order=8, fmm_leaf=64, precision="fp64",
            device=0, profile=False, fmm_far_order=0, krylov="fp64", krylov_inner_tol=0.0, batch_solver="union",
            batch_block=0):
    o = nv.Options()
    nv.hip_lib().fg_default_options(C.byref(o))
    o.tol = tol
    o.max_iter = max_iter
    o.demag_method = {"direct": nv.FG_DEMAG_DIRECT, "fmm": nv.FG_DEMAG_FMM}[demag]
    o.fmm_order = fmm_order
    o.fmm_leaf = fmm_leaf
    o.precision = {"fp64": nv.FG_PREC_FP64, "mixed": nv.FG_PREC_MIXED}[precision]
    o.device = device
    o.profile = 1 if profile else 0
    o.fmm_far_order = fmm_far_order
    o.krylov = {"fp64": nv.FG_KRYLOV_FP64, "mixed": nv.FG_KRYLOV_MIXED}[krylov]
    o.krylov_inner_tol = krylov_inner_tol
    o.batch_solver = {"union": nv.FG_BATCH_UNION, "persistent": nv.FG_BATCH_PERSISTENT}[batch_solver]
    o.batch_block = batch_block
    return o


class Engine:
    def __init__(self, mesh, opts=None, **kw):
        self.lib = nv.hip_lib()
        self.N = mesh.n_nodes
        self._mesh = mesh
        o = opts if opts is not None else options(**kw)
        h = C.c_void_p()
        rc = self.lib.fg_create(C.byref(mesh), C.byref(o), C.byref(h))
        if rc != nv.FG_OK:
            raise EngineError(f"fg_create: {rc}: {self.lib.fg_last_error(None).decode()}")
        self._h = h

    @classmethod
    def batch(cls, meshes, opts=None, **kw):
        """One context over independent meshes side by side (fg_create_batch)."""
        self = cls.__new__(cls)
        self.lib = nv.hip_lib()
        arr = (nv.Mesh * len(meshes))(*[m for m in meshes])
        self._meshes = meshes
        self.N = sum(m.n_nodes for m in meshes)
        o = opts if opts is not None else options(**kw)
        h = C.c_void_p()
        rc = self.lib.fg_create_batch(arr, len(meshes), C.byref(o), C.byref(h))
        if rc != nv.FG_OK:
            raise EngineError(f"fg_create_batch: {rc}: {self.lib.fg_last_error(None).decode()}")
        self._h = h
        return self

    def layout(self):
        """(node offsets, magnetic-tetrahedron offsets) of the instances, each of size k + 1."""
        k = C.c_int()
        self._chk(self.lib.fg_batch_layout(self._h, None, None, C.byref(k)), "fg_batch_layout")
        no = np.zeros(k.value + 1, np.int64)
        to = np.zeros(k.value + 1, np.int64)
        self._chk(self.lib.fg_batch_layout(self._h, nv.lptr(no), nv.lptr(to), C.byref(k)), "fg_batch_layout")
        return no, to

    def close(self):
        if getattr(self, "_h", None):
            self.lib.fg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc < 0:
            raise EngineError(f"{what}: {rc}: {self.lib.fg_last_error(self._h).decode()}")
        return rc

    def set_state(self, which, u=None, v=None, phi=None, phiv=None):
        c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        u, v, phi, phiv = c(u), c(v), c(phi), c(phiv)
        self._chk(self.lib.fg_set_state(self._h, which, nv.dptr(u), nv.dptr(v), nv.dptr(phi), nv.dptr(phiv)),
                  "fg_set_state")

    def get_state(self, which):
        u = np.zeros(3 * self.N); v = np.zeros(3 * self.N)
        phi = np.zeros(self.N); phiv = np.zeros(self.N)
        self._chk(self.lib.fg_get_state(self._h, which, nv.dptr(u), nv.dptr(v), nv.dptr(phi), nv.dptr(phiv)),
                  "fg_get_state")
        return u.reshape(-1, 3), v.reshape(-1, 3), phi, phiv

    def base_projection(self, r):
        self._chk(self.lib.fg_base_projection(self._h, r), "fg_base_projection")

    def prepare_elements(self, Hext, dt, prefactor):
        H = np.ascontiguousarray(Hext, dtype=np.float64)
        self._chk(self.lib.fg_prepare_elements(self._h, nv.dptr(H), dt, prefactor), "fg_prepare_elements")

    def set_space_field(self, h):
        """Space-dependent applied field, [n_tets, 5, 3] A/m at the Gauss points, or None."""
        if h is None:
            self._chk(self.lib.fg_set_space_field(self._h, None), "fg_set_space_field")
            return
        hh = np.ascontiguousarray(h, dtype=np.float64).reshape(-1)
        self._chk(self.lib.fg_set_space_field(self._h, nv.dptr(hh)), "fg_set_space_field")

    def set_spin_accumulation(self, s, vol=None):
        """The spin-accumulation field Hst of the element pass from s [n_nodes, 3] (A/m) and the
        regions' nv.SpinVolume parameters, or None to remove it."""
        if s is None:
            self._chk(self.lib.fg_set_spin_accumulation(self._h, None, None), "fg_set_spin_accumulation")
            return
        ss = np.ascontiguousarray(s, dtype=np.float64).reshape(-1)
        va = (nv.SpinVolume * len(vol))(*vol)
        self._chk(self.lib.fg_set_spin_accumulation(self._h, nv.dptr(ss), va), "fg_set_spin_accumulation")

    def spin_solve(self, spin, accumulate=True, want_s=True):
        """One solve of a SpinSolver at this engine's state (exchange from CURRENT, source from NEXT),
        installed as the spin-accumulation field; returns (s [n_nodes, 3] or None, SolveInfo)."""
        info = nv.SolveInfo()
        s = np.zeros(3 * self.N) if want_s else None
        self._chk(self.lib.fg_spin_solve_ctx(spin._h, self._h, int(accumulate), None if s is None else nv.dptr(s),
                                             C.byref(info)), "fg_spin_solve_ctx")
        return (None if s is None else s.reshape(-1, 3)), info

    def prepare_elements_amp(self, amp, dt, prefactor):
        self._chk(self.lib.fg_prepare_elements_amp(self._h, amp, dt, prefactor), "fg_prepare_elements_amp")

    def energy_amp(self, amp):
        E = np.zeros(4)
        self._chk(self.lib.fg_energy_amp(self._h, amp, nv.dptr(E)), "fg_energy_amp")
        return E

    def solve(self, dt):
        """Returns (failed, SolveInfo) like LinAlgebra::solve's bool (true = failure)."""
        info = nv.SolveInfo()
        rc = self._chk(self.lib.fg_solve(self._h, dt, C.byref(info)), "fg_solve")
        return rc == nv.FG_SOLVE_FAILED, info

    def calc_demag(self):
        self._chk(self.lib.fg_calc_demag(self._h), "fg_calc_demag")

    def energy(self, Hext):
        H = np.ascontiguousarray(Hext, dtype=np.float64)
        E = np.zeros(4)
        self._chk(self.lib.fg_energy(self._h, nv.dptr(H), nv.dptr(E)), "fg_energy")
        return E

    def evolution(self):
        self._chk(self.lib.fg_evolution(self._h), "fg_evolution")

    def max_angle(self):
        a = C.c_double()
        self._chk(self.lib.fg_max_angle(self._h, C.byref(a)), "fg_max_angle")
        return a.value

    def average(self, which=0):
        out = np.zeros(3)
        self._chk(self.lib.fg_average(self._h, which, nv.dptr(out)), "fg_average")
        return out

    def basis(self):
        ep = np.zeros(3 * self.N); eq = np.zeros(3 * self.N)
        self._chk(self.lib.fg_get_basis(self._h, nv.dptr(ep), nv.dptr(eq)), "fg_get_basis")
        return ep.reshape(-1, 3), eq.reshape(-1, 3)

    def system(self):
        rows, nnz = C.c_int64(), C.c_int64()
        self._chk(self.lib.fg_system_size(self._h, C.byref(rows), C.byref(nnz)), "fg_system_size")
        rp = np.zeros(rows.value + 1, np.int64); col = np.zeros(nnz.value, np.int32)
        val = np.zeros(nnz.value); rhs = np.zeros(rows.value)
        self._chk(self.lib.fg_get_system(self._h, nv.lptr(rp), nv.iptr(col), nv.dptr(val), nv.dptr(rhs)),
                  "fg_get_system")
        return rp, col, val, rhs

    def solution(self):
        x = np.zeros(2 * self.N)
        self._chk(self.lib.fg_get_solution(self._h, nv.dptr(x)), "fg_get_solution")
        return x

    def synchronize(self):
        self._chk(self.lib.fg_synchronize(self._h), "fg_synchronize")

    def kernel_stats(self):
        out = {}
        k = 0
        name = C.create_string_buffer(64)
        n = C.c_int64()
        ms = C.c_double()
        while self.lib.fg_kernel_stats(self._h, k, name, 64, C.byref(n), C.byref(ms)) == nv.FG_OK:
            out[name.value.decode()] = (n.value, ms.value)
            k += 1
        return out

    def fmm_info(self):
        """Octree plan statistics of an FMM context (keys as in fg_fmm_info), or None."""
        a = np.zeros(16, dtype=np.int64)
        if self.lib.fg_fmm_info(self._h, nv.lptr(a)) != nv.FG_OK:
            return None
        keys = ["leaf_level", "cells", "leaves", "groups", "list_entries", "m2p_pairs", "m2l_pairs",
                "near_coefs", "near_columns", "targets", "sources", "order", "target_leaves", "m2m_cells",
                "m2p_slots", "m2p_far_pairs"]
        return {k: int(a[i]) for i, k in enumerate(keys)}

    # ---- learned initial guess ----
    def guess_init(self, seed=0):
        self._chk(self.lib.fg_guess_init(self._h, seed), "fg_guess_init")

    def guess_set(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        self._chk(self.lib.fg_guess_set(self._h, nv.fptr(p), p.size), "fg_guess_set")

    def guess_get(self):
        p = np.zeros(nv.FG_GUESS_NPARAM, dtype=np.float32)
        self._chk(self.lib.fg_guess_get(self._h, nv.fptr(p), p.size), "fg_guess_get")
        return p

    def guess_enable(self, on=True):
        self._chk(self.lib.fg_guess_enable(self._h, 1 if on else 0), "fg_guess_enable")

    def guess_train(self, lr):
        a, b = C.c_double(), C.c_double()
        self._chk(self.lib.fg_guess_train(self._h, lr, C.byref(a), C.byref(b)), "fg_guess_train")
        return a.value, b.value

    def guess_gradient(self):
        g = np.zeros(nv.FG_GUESS_NPARAM - nv.FG_GUESS_NSCALE, dtype=np.float32)
        a, b = C.c_double(), C.c_double()
        self._chk(self.lib.fg_guess_gradient(self._h, nv.fptr(g), C.byref(a), C.byref(b)), "fg_guess_gradient")
        return g, a.value, b.value

    def guess_features(self):
        f = np.zeros(nv.FG_GUESS_FEATURE_ROW * self.N, dtype=np.float32)
        self._chk(self.lib.fg_guess_features(self._h, nv.fptr(f)), "fg_guess_features")
        return f.reshape(self.N, nv.FG_GUESS_FEATURE_ROW)

    def reset_kernel_stats(self):
        self.lib.fg_reset_kernel_stats(self._h)

    def step(self, r, Hext, dt, prefactor, dumax=0.0):
        """One step of the reference loop body (src/time_integration.cpp:193-242) in one call
        (fg_step): base_projection(r), prepare_elements, solve, then, unless the solve failed or
        dt * v_max > dumax (dumax <= 0: no limit), compute_all (demag, energy, evolution) and
        max_angle. Returns (status, SolveInfo, E, max_angle): status FG_OK, FG_SOLVE_FAILED or
        FG_STEP_REJECTED; E and max_angle are None unless the step was taken."""
        H = np.ascontiguousarray(Hext, dtype=np.float64)
        info = nv.SolveInfo()
        E = np.zeros(4)
        a = C.c_double()
        rc = self._chk(self.lib.fg_step(self._h, nv.dptr(H), dt, prefactor, r, dumax, C.byref(info), nv.dptr(E),
                                        C.byref(a)), "fg_step")
        if rc != nv.FG_OK:
            return rc, info, None, None
        return rc, info, E, a.value


class SpinSolver:
    """The persistent spin-diffusion solver of a run (fg_spin_*, spinAcc in the reference):
    built once from the mesh view, the regions' nv.SpinVolume / nv.SpinSurface parameters and the
    node potential V; every solve with accumulate=True adds its assembly to the system of the
    earlier ones, as the reference's solver does."""

    def __init__(self, mesh, vol, surf, V, tol=1e-8, max_iter=1000, device=0):
        self.lib = nv.hip_lib()
        self.N = mesh.n_nodes
        va = (nv.SpinVolume * len(vol))(*vol)
        sa = (nv.SpinSurface * len(surf))(*surf)
        VV = np.ascontiguousarray(V, dtype=np.float64)
        h = C.c_void_p()
        rc = self.lib.fg_spin_create(C.byref(mesh), va, sa, nv.dptr(VV), tol, max_iter, device, C.byref(h))
        if rc != nv.FG_OK:
            raise EngineError(f"fg_spin_create: {rc}")
        self._h = h

    def solve(self, u, u_src=None, accumulate=True):
        """s [n_nodes, 3] and SolveInfo at the magnetizations u (s-d exchange) and u_src (source)."""
        uu = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        us = None if u_src is None else np.ascontiguousarray(u_src, dtype=np.float64).reshape(-1)
        s = np.zeros(3 * self.N)
        info = nv.SolveInfo()
        rc = self.lib.fg_spin_solve(self._h, nv.dptr(uu), None if us is None else nv.dptr(us), int(accumulate),
                                    nv.dptr(s), C.byref(info))
        if rc != nv.FG_OK:
            raise EngineError(f"fg_spin_solve: {rc}")
        return s.reshape(-1, 3), info

    def close(self):
        if getattr(self, "_h", None):
            self.lib.fg_spin_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
