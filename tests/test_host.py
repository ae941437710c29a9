"""Host side (CPU): gmsh reading, preprocessing (orientation, controlTriangles, sortNodes),
generators and .sol files."""
import os

import numpy as np
import pytest

from feellgood_amd.mesh import HostMesh


def test_read_ellipsoid_counts(golden_dir):
    """ut_readMesh.cpp:18-75: examples/ellipsoid.msh holds 167 nodes, 499 tets, 274 triangles."""
    m = HostMesh.read(os.path.join(golden_dir, "ellipsoid.msh"))
    assert m.raw_counts() == (167, 499, 274)
    # naming the surface region reads its 274 triangles; none is then created
    m.prepare({"ellipsoid_volume": {}}, {"ellipsoid_surface": {}})
    assert (m.n_nodes, m.n_tets, m.n_tris) == (167, 499, 274)
    assert set(m.arrays()["tri_region"]) == {0}
    # leaving it unnamed: controlTriangles creates the 274 boundary triangles itself
    m2 = HostMesh.read(os.path.join(golden_dir, "ellipsoid.msh"))
    m2.prepare({"ellipsoid_volume": {}})
    assert m2.n_tris == 274 and set(m2.arrays()["tri_region"]) == {0}  # the created region, index n_surf


def _parse_msh41_ascii(path):
    """The sections of an ASCII MSH 4.1 file as Python data (test helper for the binary writers)."""
    lines = open(path).read().split("\n")
    sec, i = {}, 0
    while i < len(lines):
        t = lines[i].strip()
        if t.startswith("$") and not t.startswith("$End"):
            j = lines.index("$End" + t[1:], i)
            sec[t[1:]] = lines[i + 1:j]
            i = j
        i += 1
    toks = " ".join(sec["Entities"]).split()
    pos = 0

    def nxt(f=int):
        nonlocal pos
        pos += 1
        return f(toks[pos - 1])
    counts = [nxt() for _ in range(4)]
    ents = []
    for dim in range(4):
        for _ in range(counts[dim]):
            tag = nxt()
            box = [nxt(float) for _ in range(3 if dim == 0 else 6)]
            ph = [nxt() for _ in range(nxt())]
            bnd = [nxt() for _ in range(nxt())] if dim > 0 else []
            ents.append((dim, tag, box, ph, bnd))
    toks, pos = " ".join(sec["Nodes"]).split(), 0
    nblocks, nnodes, nmin, nmax = (nxt() for _ in range(4))
    nblk = []
    for _ in range(nblocks):
        edim, etag, par, nb = (nxt() for _ in range(4))
        tags = [nxt() for _ in range(nb)]
        xyz = [nxt(float) for _ in range(3 * nb)]
        nblk.append((edim, etag, tags, xyz))
    toks, pos = " ".join(sec["Elements"]).split(), 0
    eblocks, nelem, emin, emax = (nxt() for _ in range(4))
    nn = {15: 1, 1: 2, 2: 3, 4: 4}
    eblk = []
    for _ in range(eblocks):
        edim, etag, typ, nb = (nxt() for _ in range(4))
        els = [[nxt() for _ in range(1 + nn[typ])] for _ in range(nb)]
        eblk.append((edim, etag, typ, els))
    return dict(phys=sec["PhysicalNames"], ents=ents, counts=counts, nodes=(nnodes, nmin, nmax, nblk),
                elems=(nelem, emin, emax, eblk))


def _write_msh41_binary(d, path):
    import struct
    i, q, f = (lambda v: struct.pack("<i", v)), (lambda v: struct.pack("<Q", v)), (lambda v: struct.pack("<d", v))
    out = [b"$MeshFormat\n4.1 1 8\n", i(1), b"\n$EndMeshFormat\n"]
    out += [b"$PhysicalNames\n", ("\n".join(d["phys"]) + "\n").encode(), b"$EndPhysicalNames\n"]
    out += [b"$Entities\n"] + [q(c) for c in d["counts"]]
    for dim, tag, box, ph, bnd in d["ents"]:
        out += [i(tag)] + [f(v) for v in box] + [q(len(ph))] + [i(v) for v in ph]
        if dim > 0:
            out += [q(len(bnd))] + [i(v) for v in bnd]
    out += [b"\n$EndEntities\n"]
    # a section the reader does not use, holding bytes that spell section names
    out += [b"$NodeData\n", b"\x00$EndNodes$Nodes\x07" * 3, b"\n$EndNodeData\n"]
    nnodes, nmin, nmax, nblk = d["nodes"]
    out += [b"$Nodes\n", q(len(nblk)), q(nnodes), q(nmin), q(nmax)]
    for edim, etag, tags, xyz in nblk:
        out += [i(edim), i(etag), i(0), q(len(tags))] + [q(t) for t in tags] + [f(v) for v in xyz]
    out += [b"\n$EndNodes\n"]
    nelem, emin, emax, eblk = d["elems"]
    out += [b"$Elements\n", q(len(eblk)), q(nelem), q(emin), q(emax)]
    for edim, etag, typ, els in eblk:
        out += [i(edim), i(etag), i(typ), q(len(els))] + [q(v) for e in els for v in e]
    out += [b"\n$EndElements\n"]
    open(path, "wb").write(b"".join(out))


def _write_msh22_binary(d, path):
    """Element blocks of one type each, tags (physical, elementary): the MSH 2.2 binary layout."""
    import struct
    i, f = (lambda v: struct.pack("<i", v)), (lambda v: struct.pack("<d", v))
    ph = {(dim, tag): (p[0] if p else 0) for dim, tag, _, p, _ in d["ents"]}
    out = [b"$MeshFormat\n2.2 1 8\n", i(1), b"\n$EndMeshFormat\n"]
    out += [b"$PhysicalNames\n", ("\n".join(d["phys"]) + "\n").encode(), b"$EndPhysicalNames\n"]
    nnodes, _, _, nblk = d["nodes"]
    out += [("$Nodes\n%d\n" % nnodes).encode()]
    for _, _, tags, xyz in nblk:
        for k, t in enumerate(tags):
            out += [i(t)] + [f(v) for v in xyz[3 * k:3 * k + 3]]
    out += [b"\n$EndNodes\n"]
    nelem, _, _, eblk = d["elems"]
    out += [("$Elements\n%d\n" % nelem).encode()]
    for edim, etag, typ, els in eblk:
        out += [i(typ), i(len(els)), i(2)]
        for e in els:
            out += [i(e[0]), i(ph[(edim, etag)]), i(etag)] + [i(v) for v in e[1:]]
    out += [b"\n$EndElements\n"]
    open(path, "wb").write(b"".join(out))


@pytest.mark.parametrize("fmt", ["4.1", "2.2"])
def test_read_binary_msh(golden_dir, tmp_path, fmt):
    """Binary gmsh files (the reference reads any file the gmsh library opens, src/read.cpp:12-60):
    the golden ellipsoid rewritten as binary MSH 4.1 and 2.2 reads as the ASCII original, array
    for array; truncated files are reported, not read short."""
    src = os.path.join(golden_dir, "ellipsoid.msh")
    d = _parse_msh41_ascii(src)
    path = tmp_path / f"ellipsoid_{fmt}.msh"
    (_write_msh41_binary if fmt == "4.1" else _write_msh22_binary)(d, path)
    a, b = HostMesh.read(src), HostMesh.read(path)
    assert b.raw_counts() == a.raw_counts() == (167, 499, 274)
    for m in (a, b):
        m.prepare({"ellipsoid_volume": dict(K=3e5, uk=(0, 1, 0))}, {"ellipsoid_surface": {}})
    x, y = a.arrays(), b.arrays()
    for k in x:
        assert np.array_equal(x[k], y[k]), k
    data = open(path, "rb").read()
    cut = tmp_path / "cut.msh"
    cut.write_bytes(data[:data.index(b"$EndNodes") - 40])
    with pytest.raises(RuntimeError, match="truncated"):
        HostMesh.read(cut)


def _outward_ok(m):
    a = m.arrays()
    pos, tets, tris, dMs = a["pos"], a["tets"], a["tris"], a["dMs"]
    # positive orientation of every tetrahedron
    P = pos[tets]
    mixed = np.einsum("ij,ij->i", P[:, 1] - P[:, 0], np.cross(P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]))
    assert (mixed > 0).all()
    # dMs * n points out of the magnet: closed surface => sum dMs * area * n = 0, and the
    # centroid of the body lies on the inner side of every face
    Q = pos[tris]
    nrm = np.cross(Q[:, 1] - Q[:, 0], Q[:, 2] - Q[:, 0])
    c = pos.mean(axis=0)
    side = np.einsum("ij,ij->i", (Q.mean(axis=1) - c), nrm) * dMs
    return side, (dMs[:, None] * nrm).sum(axis=0), np.abs(dMs[:, None] * nrm).sum()


def test_cuboid_control_triangles_and_sorting():
    m = HostMesh.cuboid((30.0, 20.0, 10.0), (6, 4, 2), jitter=0.1, seed=11)
    assert m.raw_counts()[1] == 6 * 48
    m.prepare({"volume": dict(Ms=8e5)})
    nx, ny, nz = 6, 4, 2
    assert m.n_nodes == (nx + 1) * (ny + 1) * (nz + 1)
    assert m.n_tris == 4 * (nx * ny + ny * nz + nx * nz)  # 2 triangles per boundary quad
    side, closure, scale = _outward_ok(m)
    assert (side > 0).all()
    assert np.abs(closure).max() < 1e-12 * scale
    assert np.all(np.abs(m.arrays()["dMs"]) == 8e5)
    pos = m.arrays()["pos"]
    assert np.all(np.diff(pos[:, 0]) >= 0)  # sorted along the longest axis (x)
    l, c = m.bbox()
    assert np.allclose(l, [30e-9, 20e-9, 10e-9]) and np.allclose(c, [15e-9, 10e-9, 5e-9])
    ni = m.node_index()
    assert sorted(ni) == list(range(m.n_nodes))


def test_interface_triangles_signed_dMs(tmp_path):
    """Two volume regions: the interface gets one triangle per face with dMs = Ms1 - Ms2 in the
    triangle's orientation (src/mesh.cpp:227-242)."""
    path = tmp_path / "two.msh"
    HostMesh.cuboid((20.0, 10.0, 10.0), (4, 2, 2)).write_msh(path)
    # relabel the tets of the right half (centroid x > 10) to a second volume region
    txt = path.read_text().splitlines()
    nodes = {}
    lines = txt
    i_nodes = lines.index("$Nodes")
    nn = int(lines[i_nodes + 1])
    for k in range(nn):
        parts = lines[i_nodes + 2 + k].split()
        nodes[int(parts[0])] = float(parts[1])
    i_el = lines.index("$Elements")
    ne = int(lines[i_el + 1])
    for k in range(ne):
        parts = lines[i_el + 2 + k].split()
        if parts[1] == "4":
            xs = [nodes[int(v)] for v in parts[5:9]]
            if sum(xs) / 4 > 10:
                parts[3] = "301"
        lines[i_el + 2 + k] = " ".join(parts)
    i_pn = lines.index("$PhysicalNames")
    lines[i_pn + 1] = str(int(lines[i_pn + 1]) + 1)
    lines.insert(i_pn + 2, '3 301 "right"')
    path.write_text("\n".join(lines) + "\n")
    m = HostMesh.read(path)
    m.prepare({"volume": dict(Ms=8e5), "right": dict(Ms=5e5)})
    arr = m.arrays()
    pos, tris, dMs = arr["pos"], arr["tris"], arr["dMs"]
    xs = pos[tris][:, :, 0]
    iface = np.all(np.abs(xs - 10e-9) < 1e-15, axis=1)
    assert iface.sum() == 8
    nrm = np.cross(pos[tris[:, 1]] - pos[tris[:, 0]], pos[tris[:, 2]] - pos[tris[:, 0]])
    # dMs * n_x on the interface is (8e5 - 5e5) pointing from the left (Ms 8e5) into the right
    assert np.allclose(np.sign(nrm[iface, 0]) * dMs[iface], 3e5)


@pytest.mark.parametrize("k,msg", [(1, "instances of the same surface triangle"),
                                   (2, "belongs to no tetrahedron"),
                                   (3, "shared by 3 tetrahedra"),
                                   (4, "internal triangle")])
def test_control_triangles_errors(golden_dir, k, msg):
    """ut_readMesh.cpp:77-127 with unit-tests/meshes/bad_cuboid_{1..4}.msh."""
    m = HostMesh.read(os.path.join(golden_dir, f"bad_cuboid_{k}.msh"))
    with pytest.raises(RuntimeError, match=msg):
        m.prepare({"whole_volume": {}}, {"whole_surface": {}})


def test_missing_region_is_an_error(golden_dir):
    m = HostMesh.read(os.path.join(golden_dir, "ellipsoid.msh"))
    with pytest.raises(RuntimeError, match="not found"):
        m.prepare({"no_such_volume": {}})


def test_sol_round_trip(tmp_path):
    m = HostMesh.cuboid((10.0, 10.0, 30.0), (2, 2, 6), jitter=0.1, seed=3)
    m.prepare({"volume": {}})
    rng = np.random.default_rng(0)
    u = rng.normal(size=(m.n_nodes, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    phi = rng.normal(size=m.n_nodes)
    m.write_sol(tmp_path / "a.sol", u, phi, t=1.5e-12, precision=15)
    u2, phi2, t = m.read_sol(tmp_path / "a.sol")
    assert t == 1.5e-12
    assert np.allclose(u2, u, rtol=1e-14) and np.allclose(phi2, phi, rtol=1e-14)
    # file order is the original node numbering
    text = (tmp_path / "a.sol").read_text().splitlines()
    assert text[0].startswith("## time:") and text[1].startswith("0\t")


def test_ellipsoid_generator_valid():
    m = HostMesh.ellipsoid((0, 0, 0), (30.0, 20.0, 8.0), 2.0)
    m.prepare({"volume": dict(Ms=1e6)})
    side, closure, scale = _outward_ok(m)
    assert np.abs(closure).max() < 1e-12 * scale


def test_cylinder_generator_valid():
    """Positive tetrahedra, a closed outward surface, and a volume within the staircase error of
    pi r^2 L (half-way projection of the rim)."""
    r, L, h = 20.0, 60.0, 2.0
    m = HostMesh.cylinder((1.0, -2.0), 5.0, r, L, h)
    m.prepare({"volume": dict(Ms=1e6)})
    side, closure, scale = _outward_ok(m)
    assert np.abs(closure).max() < 1e-12 * scale
    a = m.arrays()
    pos, tn = a["pos"], a["tets"]
    vol = np.linalg.det(pos[tn[:, 1:]] - pos[tn[:, :1]]) / 6
    assert vol.min() > 0
    exact = np.pi * r * r * L * 1e-27
    assert abs(vol.sum() - exact) < 0.03 * exact
    # radial extent: the rim within half a cell diagonal of r (half-way projection), axis at
    # (1, -2), ends at z0 and z0 + L
    rho = np.hypot(pos[:, 0] - 1e-9, pos[:, 1] + 2e-9)
    assert r - h <= rho.max() * 1e9 <= r + h / np.sqrt(2)
    assert pos[:, 2].min() == pytest.approx(5e-9, rel=1e-12, abs=0) and pos[:, 2].max() == pytest.approx(65e-9, rel=1e-12, abs=0)


def _ev(expr, names="t", vals=(0.0,)):
    import ctypes as C
    from feellgood_amd import native as nv
    v = np.ascontiguousarray(vals, dtype=np.float64)
    n = len(v) // max(1, len(names.split(",")))
    out = np.zeros(max(n, 1))
    rc = nv.host_lib().fgh_eval_expression(expr.encode(), names.encode(), n, nv.dptr(v), nv.dptr(out))
    if rc != 0:
        raise ValueError(nv.host_lib().fgh_last_error().decode())
    return out[:n]


def test_settings_expressions():
    """The arithmetic subset of the JavaScript FeeLLGood evaluates for Bext(t) and
    initial_magnetization(x, y, z) (src/expression_parser.cpp:9-33: Math members usable without the
    prefix, cosh/sinh/tanh helpers): values against Python's math, JavaScript's operator semantics,
    and errors that name the problem."""
    import math
    t = np.array([0.0, 1e-10, 3.3e-10, 1e-9])
    A, w = 0.01, 2 * math.pi * 5e9
    got = _ev(f"{A}*cos({w}*t)", "t", t)
    assert np.allclose(got, A * np.cos(w * t), rtol=1e-15, atol=0)
    assert np.allclose(_ev("Math.sin(2*Math.PI*1e9*t) + exp(-t/1e-10)", "t", t),
                       np.sin(2 * np.pi * 1e9 * t) + np.exp(-t / 1e-10), rtol=1e-14)
    xyz = np.array([[1e-9, 2e-9, -3e-9], [0.5, -0.25, 2.0]])
    assert np.allclose(_ev("-y", "x,y,z", xyz.ravel()), -xyz[:, 1])
    assert np.allclose(_ev("x*x + y**2 - z % 1e-9", "x,y,z", xyz.ravel()),
                       xyz[:, 0] ** 2 + xyz[:, 1] ** 2 - np.fmod(xyz[:, 2], 1e-9), rtol=1e-14)
    c = lambda e: float(_ev(e)[0])
    assert c("2 ** 3 ** 2") == 512.0              # right-associative
    assert c("round(2.5)") == 3.0 and c("round(-2.5)") == -2.0  # halves towards +infinity
    assert c("0 || 7") == 7.0 and c("3 && 4") == 4.0 and c("0 && 4") == 0.0
    assert c("1 < 2 ? 10 : 20") == 10.0 and c("(1 >= 2) + 1") == 1.0 and c("!0") == 1.0
    assert c("min(3, -1, 2)") == -1.0 and c("max()") == -math.inf and c("hypot(3, 4)") == 5.0
    assert c("atan2(1, 1)") == math.atan2(1, 1) and c("sign(-0.3)") == -1.0
    assert abs(c("cosh(0.7) - (exp(0.7) + exp(-0.7)) / 2")) < 1e-15
    assert c("Math.E") == math.e and c("SQRT1_2") == math.sqrt(0.5) and math.isnan(c("NaN"))
    for bad, msg in (("sin(t", "expected"), ("foo(t)", "unknown name 'foo'"), ("2 +* 3", "unexpected"),
                     ("atan2(1)", "takes 2"), ("x", "unknown name 'x'")):
        with pytest.raises(ValueError, match=msg):
            _ev(bad)
