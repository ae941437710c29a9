"""Host-side meshes: gmsh files, synthetic generators and preprocessing, through libfgmi_host.

Region defaults follow FeeLLGood's embedded settings (src/default-settings.yml:101-171).
"""
import ctypes as C

import numpy as np

from . import native as nv

VOLUME_DEFAULTS = dict(alpha=0.5, Ae=1e-11, Ms=795774.7, K=0.0, uk=(0.0, 0.0, 1.0), K3=0.0,
                       ex=(1.0, 0.0, 0.0), ey=(0.0, 1.0, 0.0), ez=(0.0, 0.0, 1.0))
SURFACE_DEFAULTS = dict(Ks=0.0, uk=(0.0, 0.0, 1.0), suppress_charges=False)


def _unit(v):
    v = np.asarray(v, dtype=np.float64)
    n = np.linalg.norm(v)
    return v / n if n > 0 else v


def volume_region(**kw):
    """fg_volume_region from keyword overrides of the reference defaults (axes are normalised,
    as Settings::read does with NORMALIZE, src/settings.cpp:404-409)."""
    p = dict(VOLUME_DEFAULTS)
    unknown = set(kw) - set(p)
    if unknown:
        raise KeyError(f"unknown volume region keys {sorted(unknown)}")
    p.update(kw)
    r = nv.VolumeRegion()
    r.alpha, r.Ae, r.Ms, r.K, r.K3 = p["alpha"], p["Ae"], p["Ms"], p["K"], p["K3"]
    for name in ("uk", "ex", "ey", "ez"):
        getattr(r, name)[:] = list(_unit(p[name]))
    return r


def surface_region(**kw):
    p = dict(SURFACE_DEFAULTS)
    p.update(kw)
    r = nv.SurfaceRegion()
    r.Ks = p["Ks"]
    r.uk[:] = list(_unit(p["uk"]))
    r.suppress_charges = 1 if p["suppress_charges"] else 0
    return r


class HostMesh:
    """A raw mesh (file or generator) that becomes an fg_mesh after :meth:`prepare`."""

    def __init__(self, handle):
        if not handle:
            raise RuntimeError(nv.host_lib().fgh_last_error().decode())
        self._h = C.c_void_p(handle)
        self._view = None
        self._keep = []

    def __del__(self):
        if getattr(self, "_h", None):
            nv.host_lib().fgh_free(self._h)
            self._h = None

    @classmethod
    def read(cls, path):
        err = C.create_string_buffer(512)
        h = nv.host_lib().fgh_read_msh(str(path).encode(), err, 512)
        if not h:
            raise RuntimeError(err.value.decode())
        return cls(h)

    @classmethod
    def cuboid(cls, lengths, cells, origin=(0.0, 0.0, 0.0), jitter=0.0, seed=0):
        lx, ly, lz = lengths
        nx, ny, nz = cells
        x0, y0, z0 = origin
        return cls(nv.host_lib().fgh_gen_cuboid(x0, y0, z0, lx, ly, lz, nx, ny, nz, jitter, seed))

    @classmethod
    def ellipsoid(cls, centre, axes, h):
        return cls(nv.host_lib().fgh_gen_ellipsoid(*centre, *axes, h))

    @classmethod
    def cylinder(cls, axis_xy, z0, radius, length, h):
        return cls(nv.host_lib().fgh_gen_cylinder(axis_xy[0], axis_xy[1], z0, radius, length, h))

    def append(self, other):
        rc = nv.host_lib().fgh_append(self._h, other._h)
        if rc != nv.FG_OK:
            raise RuntimeError("fgh_append failed")

    def raw_counts(self):
        n, t, f = C.c_int64(), C.c_int64(), C.c_int64()
        nv.host_lib().fgh_raw_counts(self._h, C.byref(n), C.byref(t), C.byref(f))
        return n.value, t.value, f.value

    def write_msh(self, path):
        nv.host_lib().fgh_write_msh22(self._h, str(path).encode())

    def prepare(self, volumes, surfaces=None, scale=1e-9):
        """volumes / surfaces: dict region name -> dict of parameter overrides."""
        surfaces = surfaces or {}
        vn = list(volumes)
        sn = list(surfaces)
        vol = (nv.VolumeRegion * len(vn))(*[volume_region(**volumes[k]) for k in vn])
        surf = (nv.SurfaceRegion * max(1, len(sn)))(*[surface_region(**surfaces[k]) for k in sn])
        vnames = (C.c_char_p * len(vn))(*[k.encode() for k in vn])
        snames = (C.c_char_p * max(1, len(sn)))(*[k.encode() for k in sn])
        rc = nv.host_lib().fgh_prepare(self._h, scale, len(vn), vnames, vol, len(sn), snames, surf)
        if rc != nv.FG_OK:
            raise RuntimeError(nv.host_lib().fgh_last_error().decode())
        m = nv.Mesh()
        nv.host_lib().fgh_view(self._h, C.byref(m))
        self._view = m
        self.region_names = vn
        return self

    @property
    def mesh(self):
        """The fg_mesh view (valid while this object lives)."""
        if self._view is None:
            raise RuntimeError("prepare() first")
        return self._view

    def _arr(self, ptr, n, dtype):
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True) if n else np.zeros(0, dtype)

    @property
    def n_nodes(self):
        return self.mesh.n_nodes

    @property
    def n_tets(self):
        return self.mesh.n_tets

    @property
    def n_tris(self):
        return self.mesh.n_tris

    def arrays(self):
        m = self.mesh
        return dict(
            pos=self._arr(m.pos, 3 * m.n_nodes, np.float64).reshape(-1, 3),
            tets=self._arr(m.tet_nodes, 4 * m.n_tets, np.int32).reshape(-1, 4),
            tet_region=self._arr(m.tet_region, m.n_tets, np.int32),
            tris=self._arr(m.tri_nodes, 3 * m.n_tris, np.int32).reshape(-1, 3),
            tri_region=self._arr(m.tri_region, m.n_tris, np.int32),
            dMs=self._arr(m.tri_dMs, m.n_tris, np.float64),
        )

    def node_index(self):
        p = nv.host_lib().fgh_node_index(self._h)
        return np.ctypeslib.as_array(p, shape=(self.n_nodes,)).copy()

    def bbox(self):
        l = np.zeros(3)
        c = np.zeros(3)
        nv.host_lib().fgh_bbox(self._h, nv.dptr(l), nv.dptr(c))
        return l, c

    def read_sol(self, path):
        u = np.zeros(3 * self.n_nodes)
        phi = np.zeros(self.n_nodes)
        t = C.c_double()
        rc = nv.host_lib().fgh_read_sol(self._h, str(path).encode(), nv.dptr(u), nv.dptr(phi), C.byref(t))
        if rc != nv.FG_OK:
            raise RuntimeError(nv.host_lib().fgh_last_error().decode())
        return u.reshape(-1, 3), phi, t.value

    def write_sol(self, path, u, phi, t=0.0, precision=7):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        meta = f"## time: {t:.16g}\n".encode()
        rc = nv.host_lib().fgh_write_sol(self._h, str(path).encode(), meta, precision, nv.dptr(u), nv.dptr(phi))
        if rc != nv.FG_OK:
            raise RuntimeError(nv.host_lib().fgh_last_error().decode())
