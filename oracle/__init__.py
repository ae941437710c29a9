"""Python handle on the FP64 CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg, and there only as the checker. The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from feellgood_amd import native as nv

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)


class SolveInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("residual", C.c_double),
                ("rhs_norm", C.c_double), ("v_max", C.c_double)]


class RunSettings(C.Structure):
    _fields_ = [("final_time", C.c_double), ("evol_time_step", C.c_double), ("dt_min", C.c_double),
                ("dt_max", C.c_double), ("du_max", C.c_double), ("tol", C.c_double),
                ("max_iter", C.c_int32), ("seed", C.c_uint32), ("Bext", C.c_double * 3),
                ("max_rows", C.c_int32), ("verbose", C.c_int32)]


SYMBOLS = {
    "orc_create": (C.c_void_p, [C.POINTER(nv.Mesh), C.c_double, C.c_int]),
    "orc_destroy": (None, [C.c_void_p]),
    "orc_last_error": (C.c_char_p, []),
    "orc_n_edges": (C.c_int, [C.c_void_p]),
    "orc_n_mag_tris": (C.c_int, [C.c_void_p]),
    "orc_set_state": (None, [C.c_void_p, C.c_int, _d, _d, _d, _d]),
    "orc_get_state": (None, [C.c_void_p, C.c_int, _d, _d, _d, _d]),
    "orc_base_projection": (None, [C.c_void_p, C.c_double]),
    "orc_get_basis": (None, [C.c_void_p, _d, _d]),
    "orc_prepare_elements": (None, [C.c_void_p, _d, C.c_double, C.c_double]),
    "orc_solve": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(SolveInfo)]),
    "orc_cg_dir": (C.c_int, [C.c_int64, _i64, _i32, _d, _d, _d, _d, C.c_int, _i32, C.c_double, C.c_int,
                             C.POINTER(SolveInfo)]),
    "orc_electrostatic_potential": (C.c_int, [C.POINTER(nv.Mesh), _d, _d, _d, C.c_double, C.c_int, _d,
                                              C.POINTER(SolveInfo)]),
    "orc_spin_accumulation": (C.c_int, [C.POINTER(nv.Mesh), C.POINTER(nv.SpinVolume), C.POINTER(nv.SpinSurface),
                                        _d, _d, _d, C.c_double, C.c_int, _d, C.POINTER(SolveInfo)]),
    "orc_bicg_dir": (C.c_int, [C.c_int64, _i64, _i32, _d, _d, _d, C.c_int, _i32, C.c_double, C.c_int,
                               C.POINTER(SolveInfo)]),
    "orc_system_size": (None, [C.c_void_p, _i64, _i64]),
    "orc_get_system": (None, [C.c_void_p, _i64, _i32, _d, _d]),
    "orc_get_solution": (None, [C.c_void_p, _d]),
    "orc_get_tet_element": (None, [C.c_void_p, C.c_int, _d, _d]),
    "orc_calc_demag": (None, [C.c_void_p]),
    "orc_set_demag": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "orc_fmm_potentials": (C.c_int, [_d, C.c_int64, _d, _d, _d, C.c_int64, C.c_int, C.c_int, _d, _d]),
    "orc_demag_sample": (None, [C.c_void_p, _i32, C.c_int, _d, _d]),
    "orc_n_sources": (C.c_int64, [C.c_void_p]),
    "orc_get_sources": (None, [C.c_void_p, _d, _d, _d]),
    "orc_get_corrections": (None, [C.c_void_p, _d, _d]),
    "orc_energy": (None, [C.c_void_p, _d, _d]),
    "orc_set_space_field": (None, [C.c_void_p, _d]),
    "orc_set_spin_accumulation": (None, [C.c_void_p, _d, C.POINTER(nv.SpinVolume)]),
    "orc_spin_accumulation_hist": (C.c_int, [C.POINTER(nv.Mesh), C.POINTER(nv.SpinVolume), C.POINTER(nv.SpinSurface),
                                             _d, C.c_int, _d, _d, C.c_double, C.c_int, _d, C.POINTER(SolveInfo)]),
    "orc_prepare_elements_amp": (None, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
    "orc_energy_amp": (None, [C.c_void_p, C.c_double, _d]),
    "orc_evolution": (None, [C.c_void_p]),
    "orc_max_angle": (C.c_double, [C.c_void_p]),
    "orc_average": (None, [C.c_void_p, C.c_int, _d]),
    "orc_tet_geometry": (C.c_int, [_d, _i32, _d, _d]),
    "orc_alpha_eff": (None, [C.c_double, C.c_double, _d, C.c_int, _d]),
    "orc_tet_lumping": (None, [_d, _d, _d, _d, C.c_double, _d]),
    "orc_tet_integrales": (None, [_d, _d, _d, _d, _d, _d, _d, _d, C.POINTER(nv.VolumeRegion), _d,
                                  C.c_double, C.c_double, _d, _d]),
    "orc_tri_potential": (C.c_double, [_d, _d, C.c_double, C.c_int]),
    "orc_aniso_uniax": (None, [_d, _d, _d, C.c_double, C.c_double, _d, _d]),
    "orc_aniso_cub": (None, [_d, _d, _d, _d, _d, C.c_double, C.c_double, _d, _d]),
    "orc_tet_energy": (None, [_d, _d, _d, _d, C.POINTER(nv.VolumeRegion), _d, _d]),
    "orc_tri_energy": (None, [_d, _d, _d, C.c_double, C.c_double, _d, _d]),
    "orc_tet_charges": (None, [_d, _d, _d, C.c_double, _d]),
    "orc_tri_charges": (None, [_d, _d, C.c_double, _d]),
    "orc_tri_integrales": (None, [_d, _d, _d, _d, C.c_double, C.c_double, _d, _d]),
    "orc_gauss_tables": (None, [_d, _d, _d, _d]),
    "orc_set_basis": (None, [_d, C.c_double, _d, _d]),
    "orc_mt_uniform": (None, [C.c_uint32, C.c_int, _d]),
    "orc_basis_angles": (None, [C.c_uint32, C.c_int, _d]),
    "orc_run": (C.c_int, [C.POINTER(nv.Mesh), _d, C.POINTER(RunSettings), _d, _i32]),
    "orc_guess_eval": (None, [C.POINTER(C.c_float), C.c_int, C.c_int, C.POINTER(C.c_float),
                              C.POINTER(C.c_float), C.POINTER(C.c_int64), _i32, _i32, _d, _d, _d, _d, _d]),
}

_lib = None


def build():
    """Compile the oracle with its Makefile (g++, OpenMP)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        l = C.CDLL(LIB)
        for name, (res, args) in SYMBOLS.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def d(a):
    return nv.dptr(a)


def mt_uniform(seed, n):
    out = np.zeros(n)
    lib().orc_mt_uniform(seed, n, d(out))
    return out


def fmm_potentials(xyz, qu, qv, tgt, order=8, leaf=64):
    """The CPU FMM: sum_s q_s / |x_t - y_s| for both charge sets at the targets; (pu, pv, leaf level)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1)
    tgt = np.ascontiguousarray(tgt, dtype=np.float64).reshape(-1)
    qu = np.ascontiguousarray(qu, dtype=np.float64)
    qv = np.ascontiguousarray(qv, dtype=np.float64)
    pu = np.zeros(len(tgt) // 3); pv = np.zeros(len(tgt) // 3)
    L = lib().orc_fmm_potentials(d(xyz), len(qu), d(qu), d(qv), d(tgt), len(pu), order, leaf, d(pu), d(pv))
    return pu, pv, L


def basis_angles(seed, n):
    out = np.zeros(n)
    lib().orc_basis_angles(seed, n, d(out))
    return out


def bicg_dir(rowptr, col, val, rhs, x0, mask=(), tol=1e-6, max_iter=700):
    """algebra::bicg_dir on a CSR matrix; returns (x, SolveInfo)."""
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    x = np.array(x0, dtype=np.float64)
    m = np.ascontiguousarray(mask, dtype=np.int32)
    info = SolveInfo()
    lib().orc_bicg_dir(len(rhs), rowptr.ctypes.data_as(_i64), col.ctypes.data_as(_i32), d(val), d(rhs),
                       d(x), len(m), m.ctypes.data_as(_i32), tol, max_iter, C.byref(info))
    return x, info


def cg_dir(rowptr, col, val, rhs, xd, x0, mask=(), tol=1e-6, max_iter=700):
    """algebra::cg_dir on a CSR matrix with Dirichlet values xd on the masked indices; returns
    (x, SolveInfo)."""
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    xd = np.ascontiguousarray(xd, dtype=np.float64)
    x = np.array(x0, dtype=np.float64)
    m = np.ascontiguousarray(mask, dtype=np.int32)
    info = SolveInfo()
    lib().orc_cg_dir(len(rhs), rowptr.ctypes.data_as(_i64), col.ctypes.data_as(_i32), d(val), d(rhs), d(xd),
                     d(x), len(m), m.ctypes.data_as(_i32), tol, max_iter, C.byref(info))
    return x, info


def electrostatic_potential(mesh, sigma, jn, V, tol=1e-8, max_iter=1000):
    """The electrostatic potential on the nodes of an fg_mesh view (sigma per volume region, jn and
    V per surface region, NaN for none); returns (V, SolveInfo). Raises ValueError on bad boundary
    conditions."""
    sigma = np.ascontiguousarray(sigma, dtype=np.float64)
    jn = np.ascontiguousarray(jn, dtype=np.float64)
    V = np.ascontiguousarray(V, dtype=np.float64)
    out = np.zeros(mesh.n_nodes)
    info = SolveInfo()
    rc = lib().orc_electrostatic_potential(C.byref(mesh), d(sigma), d(jn), d(V), tol, max_iter, d(out),
                                           C.byref(info))
    if rc < 0:
        raise ValueError("incorrect boundary conditions for the potential solver")
    return out, info


def spin_accumulation(mesh, vol, surf, V, u, tol=1e-8, max_iter=1000, u_src=None):
    """The spin accumulation (n_nodes, 3) of an fg_mesh view: vol / surf are sequences of
    nv.SpinVolume / nv.SpinSurface per region, V the node potential, u the node magnetization;
    returns (s, SolveInfo). Raises ValueError on bad boundary conditions."""
    va = (nv.SpinVolume * len(vol))(*vol)
    sa = (nv.SpinSurface * len(surf))(*surf)
    V = np.ascontiguousarray(V, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
    out = np.zeros(3 * mesh.n_nodes)
    info = SolveInfo()
    us = None if u_src is None else np.ascontiguousarray(u_src, dtype=np.float64).reshape(-1)
    rc = lib().orc_spin_accumulation(C.byref(mesh), va, sa, d(V), d(u), None if us is None else d(us), tol, max_iter,
                                     d(out), C.byref(info))
    if rc < 0:
        raise ValueError("incorrect boundary conditions for the spin diffusion solver")
    return out.reshape(-1, 3), info


def spin_accumulation_hist(mesh, vol, surf, V, uh, ush=None, tol=1e-8, max_iter=1000):
    """The k-th solve of a spin-transfer run, k = len(uh): the system summed over the assemblies at
    the magnetizations uh[j] (s-d exchange) and ush[j] (source; None: uh), as the reference's
    solver accumulates K and L_rhs over the calls of spinAcc::solve. Returns (s, SolveInfo)."""
    va = (nv.SpinVolume * len(vol))(*vol)
    sa = (nv.SpinSurface * len(surf))(*surf)
    V = np.ascontiguousarray(V, dtype=np.float64)
    uh = np.ascontiguousarray(uh, dtype=np.float64).reshape(-1)
    k = uh.size // (3 * mesh.n_nodes)
    us = None if ush is None else np.ascontiguousarray(ush, dtype=np.float64).reshape(-1)
    out = np.zeros(3 * mesh.n_nodes)
    info = SolveInfo()
    rc = lib().orc_spin_accumulation_hist(C.byref(mesh), va, sa, d(V), k, d(uh), None if us is None else d(us), tol,
                                          max_iter, d(out), C.byref(info))
    if rc < 0:
        raise ValueError("incorrect boundary conditions for the spin diffusion solver")
    return out.reshape(-1, 3), info


class Problem:
    """The oracle's whole-problem state, mirroring the fg_* calls of include/fgmi.h."""

    def __init__(self, mesh, tol=1e-6, max_iter=700):
        self.N = mesh.n_nodes
        self._mesh = mesh
        h = lib().orc_create(C.byref(mesh), tol, max_iter)
        if not h:
            raise RuntimeError(lib().orc_last_error().decode())
        self._h = C.c_void_p(h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def set_state(self, which, u=None, v=None, phi=None, phiv=None):
        c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        u, v, phi, phiv = c(u), c(v), c(phi), c(phiv)
        lib().orc_set_state(self._h, which, d(u), d(v), d(phi), d(phiv))

    def get_state(self, which):
        u = np.zeros(3 * self.N); v = np.zeros(3 * self.N)
        phi = np.zeros(self.N); phiv = np.zeros(self.N)
        lib().orc_get_state(self._h, which, d(u), d(v), d(phi), d(phiv))
        return u.reshape(-1, 3), v.reshape(-1, 3), phi, phiv

    def base_projection(self, r):
        lib().orc_base_projection(self._h, r)

    def basis(self):
        ep = np.zeros(3 * self.N); eq = np.zeros(3 * self.N)
        lib().orc_get_basis(self._h, d(ep), d(eq))
        return ep.reshape(-1, 3), eq.reshape(-1, 3)

    def prepare_elements(self, Hext, dt, prefactor):
        H = np.ascontiguousarray(Hext, dtype=np.float64)
        lib().orc_prepare_elements(self._h, d(H), dt, prefactor)

    def set_space_field(self, h):
        """Space-dependent applied field, [n_tets, 5, 3] A/m at the Gauss points, or None."""
        if h is None:
            lib().orc_set_space_field(self._h, None)
        else:
            self._hsp = np.ascontiguousarray(h, dtype=np.float64).reshape(-1)
            lib().orc_set_space_field(self._h, d(self._hsp))

    def prepare_elements_amp(self, amp, dt, prefactor):
        lib().orc_prepare_elements_amp(self._h, amp, dt, prefactor)

    def set_spin_accumulation(self, s, vol=None):
        """The spin-accumulation field of the element pass from s [n_nodes, 3] and the regions'
        nv.SpinVolume parameters, or None."""
        if s is None:
            lib().orc_set_spin_accumulation(self._h, None, None)
            return
        ss = np.ascontiguousarray(s, dtype=np.float64).reshape(-1)
        va = (nv.SpinVolume * len(vol))(*vol)
        lib().orc_set_spin_accumulation(self._h, d(ss), va)

    def solve(self, dt):
        info = SolveInfo()
        rc = lib().orc_solve(self._h, dt, C.byref(info))
        return rc, info

    def system(self):
        rows, nnz = C.c_int64(), C.c_int64()
        lib().orc_system_size(self._h, C.byref(rows), C.byref(nnz))
        rp = np.zeros(rows.value + 1, np.int64); col = np.zeros(nnz.value, np.int32)
        val = np.zeros(nnz.value); rhs = np.zeros(rows.value)
        lib().orc_get_system(self._h, rp.ctypes.data_as(_i64), col.ctypes.data_as(_i32), d(val), d(rhs))
        return rp, col, val, rhs

    def solution(self):
        x = np.zeros(2 * self.N)
        lib().orc_get_solution(self._h, d(x))
        return x

    def tet_element(self, k):
        Kp = np.zeros(64); Lp = np.zeros(8)
        lib().orc_get_tet_element(self._h, k, d(Kp), d(Lp))
        return Kp.reshape(8, 8), Lp

    def calc_demag(self):
        lib().orc_calc_demag(self._h)

    def set_demag(self, order=0, leaf=64):
        """Demag method of calc_demag: order 0 = exact direct sums (the default), else the CPU FMM."""
        if lib().orc_set_demag(self._h, order, leaf) != 0:
            raise ValueError("bad FMM order or leaf size")

    def demag_sample(self, targets):
        t = np.ascontiguousarray(targets, dtype=np.int32)
        phi = np.zeros(len(t)); phiv = np.zeros(len(t))
        lib().orc_demag_sample(self._h, t.ctypes.data_as(_i32), len(t), d(phi), d(phiv))
        return phi, phiv

    def sources(self):
        n = lib().orc_n_sources(self._h)
        xyz = np.zeros(3 * n); qu = np.zeros(n); qv = np.zeros(n)
        lib().orc_get_sources(self._h, d(xyz), d(qu), d(qv))
        return xyz.reshape(-1, 3), qu, qv

    def corrections(self):
        cu = np.zeros(self.N); cv = np.zeros(self.N)
        lib().orc_get_corrections(self._h, d(cu), d(cv))
        return cu, cv

    def energy(self, Hext):
        H = np.ascontiguousarray(Hext, dtype=np.float64)
        E = np.zeros(4)
        lib().orc_energy(self._h, d(H), d(E))
        return E

    def energy_amp(self, amp):
        E = np.zeros(4)
        lib().orc_energy_amp(self._h, amp, d(E))
        return E

    def evolution(self):
        lib().orc_evolution(self._h)

    def max_angle(self):
        return lib().orc_max_angle(self._h)

    def average(self, which=0):
        out = np.zeros(3)
        lib().orc_average(self._h, which, d(out))
        return out

    @property
    def n_edges(self):
        return lib().orc_n_edges(self._h)

    @property
    def n_mag_tris(self):
        return lib().orc_n_mag_tris(self._h)


def run(mesh, u0, final_time, evol_time_step, dt_min, dt_max, du_max, Bext, seed, tol=1e-6,
        max_iter=700, max_rows=100000, verbose=False):
    """The reference time loop with direct-sum demag; returns (status, evol rows)."""
    s = RunSettings(final_time, evol_time_step, dt_min, dt_max, du_max, tol, max_iter, seed,
                    (C.c_double * 3)(*Bext), max_rows, 1 if verbose else 0)
    evol = np.zeros(12 * max_rows)
    n = C.c_int32()
    u0 = np.ascontiguousarray(u0, dtype=np.float64).reshape(-1)
    st = lib().orc_run(C.byref(mesh), d(u0), C.byref(s), d(evol), C.byref(n))
    return st, evol[:12 * n.value].reshape(-1, 12)


def guess_eval(params, feat, adj_ptr, adj_idx, mag, v_target=None, round_bf16=True, want_grad=True, v_base=None):
    """The learned-guess network on the CPU: (v_pred [n,3], loss, loss_ref, grad). feat is [n, 12]
    (u, h, previous v, second difference) or the engine's [n, 16] export, whose columns 12..14 are
    the base the network corrects (v_base, else the previous v)."""
    n = len(mag)
    params = np.ascontiguousarray(params, dtype=np.float32)
    feat = np.asarray(feat, dtype=np.float32).reshape(n, -1)
    if feat.shape[1] == nv.FG_GUESS_FEATURE_ROW and v_base is None:
        v_base = feat[:, 12:15]
    feat = np.ascontiguousarray(feat[:, :12]).reshape(-1)
    vb = None if v_base is None else np.ascontiguousarray(v_base, dtype=np.float32).reshape(-1)
    fp = C.POINTER(C.c_float)
    adj_ptr = np.ascontiguousarray(adj_ptr, dtype=np.int64)
    adj_idx = np.ascontiguousarray(adj_idx, dtype=np.int32)
    mag = np.ascontiguousarray(mag, dtype=np.int32)
    vt = None if v_target is None else np.ascontiguousarray(v_target, dtype=np.float64).reshape(-1)
    vp = np.zeros(3 * n)
    grad = np.zeros(len(params) - nv.FG_GUESS_NSCALE) if want_grad else None
    loss, loss_ref = C.c_double(), C.c_double()
    lib().orc_guess_eval(params.ctypes.data_as(C.POINTER(C.c_float)), 1 if round_bf16 else 0, n,
                         feat.ctypes.data_as(fp), None if vb is None else vb.ctypes.data_as(fp),
                         adj_ptr.ctypes.data_as(C.POINTER(C.c_int64)),
                         adj_idx.ctypes.data_as(_i32), mag.ctypes.data_as(_i32), d(vt), d(vp), C.byref(loss),
                         C.byref(loss_ref), d(grad))
    return vp.reshape(n, 3), loss.value, loss_ref.value, grad
