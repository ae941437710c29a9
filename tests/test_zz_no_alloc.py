"""fg_electrostatic_potential_noalloc: the potential solve with every device buffer taken from the
library's static scratch array instead of hipMalloc. Same kernels and order as
fg_electrostatic_potential, so the results are identical, on two calls in a row (the second on a
smaller mesh, which reuses the array's leftover contents) and in either order; a mesh whose set-up
does not fit the array is refused with FG_ERR_RANGE."""
import ctypes as C

import numpy as np
import pytest

from feellgood_amd import native as nv
from tests.test_electrostatic import BC, bar

pytestmark = pytest.mark.gpu


def potential(entry, m, tol=1e-8, max_iter=1000):
    V = np.full(m.n_nodes, np.nan)
    info = nv.SolveInfo()
    rc = getattr(nv.hip_lib(), entry)(C.byref(m), nv.dptr(np.array(BC["sigma"], dtype=np.float64)),
                                      nv.dptr(np.array(BC["jn"], dtype=np.float64)),
                                      nv.dptr(np.array(BC["V"], dtype=np.float64)), tol, max_iter, 0, nv.dptr(V),
                                      C.byref(info))
    return rc, V, info


def test_noalloc_potential_matches_the_allocating_entry_point():
    for cells, first in (((12, 4, 5), "noalloc"), ((6, 3, 3), "alloc")):
        m, _own, _ = bar(0.3, cells=cells)
        order = ["fg_electrostatic_potential_noalloc", "fg_electrostatic_potential"]
        if first == "alloc":
            order.reverse()
        out = {}
        for entry in order:
            rc, V, info = potential(entry, m)
            assert rc == nv.FG_OK, (entry, rc)
            out[entry] = (V, info)
        (V0, i0), (V1, i1) = out["fg_electrostatic_potential"], out["fg_electrostatic_potential_noalloc"]
        print(f"{m.n_nodes} nodes: {i1.iterations} iterations, max |dV| {np.abs(V1 - V0).max():.1e}", flush=True)
        assert i1.status == i0.status == nv.FG_ITER_CONVERGED
        assert i1.iterations == i0.iterations
        assert i1.residual == i0.residual and i1.rhs_norm == i0.rhs_norm
        assert np.array_equal(V1, V0)


def test_noalloc_potential_refuses_a_mesh_beyond_the_scratch():
    m, _own, _ = bar(0.0, cells=(60, 30, 30))  # 58,621 nodes, 324,000 tetrahedra
    rc, V, _ = potential("fg_electrostatic_potential_noalloc", m)
    assert rc == nv.FG_ERR_RANGE
    # the allocating entry point still takes it
    rc, V, info = potential("fg_electrostatic_potential", m, max_iter=5)
    assert rc == nv.FG_OK and info.iterations == 5
