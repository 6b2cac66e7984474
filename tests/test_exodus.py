"""On-disk mesh input and output (Exodus II over the netCDF classic formats),
host only.

The netCDF side is checked against scipy.io.netcdf_file, an independent
implementation of the CDF-1 / CDF-2 formats: files scipy writes (in either
format, with coordinates stored either way, several blocks per topology,
node and side sets, nodal variables over two time steps) are read by the
library, and the files the library writes are read back by scipy.  The
side-face tables are pinned by geometry: the faces of every element form a
closed, consistently oriented surface whose normals point outward, and the
nodes of a box's boundary side sets are exactly the nodes on that plane."""
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

from tests import problems as pb

ETYPE = {0: b"HEX8", 1: b"TETRA4", 2: b"WEDGE6", 3: b"PYRAMID5"}
NPE = {0: 8, 1: 4, 2: 6, 3: 5}


@pytest.fixture(scope="module")
def nw():
    return pb.nw()


@pytest.fixture(scope="module")
def mesh(nw):
    coords, types, conn = nw.mixed_box(6, 0.8, 0.1, jitter=0.2, seed=3)  # 73 Hex8, 412 Tet4, 42 Wed6, 460 Pyr5
    return coords, types, conn


def _names(rows, width=33):
    a = np.zeros((len(rows), width), "S1")
    for i, s in enumerate(rows):
        a[i, :len(s)] = np.frombuffer(s.encode(), "S1")
    return a


def _scipy_exodus(path, coords, blocks, node_sets=(), side_sets=(), nodvars=(),
                  times=(), version=2, packed_coord=False, names=True):
    """Exodus II file written with scipy from the Exodus II variable names:
    blocks [(id, name, etype, conn0 [n][npe])], node_sets [(id, name, nodes0)],
    side_sets [(id, name, elem0, side0)], nodvars [(name, values [T][N])]."""
    f = netcdf_file(path, "w", version=version)
    f.title = b"written by scipy"
    f.api_version = np.float32(8.03)
    f.floating_point_word_size = np.int32(8)
    f.createDimension("time_step", None)
    f.createDimension("len_name", 33)
    f.createDimension("num_dim", 3)
    f.createDimension("num_nodes", coords.shape[0])
    f.createDimension("num_elem", sum(b[3].shape[0] for b in blocks))
    f.createDimension("num_el_blk", len(blocks))
    t = f.createVariable("time_whole", "d", ("time_step",))
    if len(times):
        t[0:len(times)] = np.asarray(times)
    ids = f.createVariable("eb_prop1", "i", ("num_el_blk",))
    ids[:] = [b[0] for b in blocks]
    ids.name = b"ID"
    if names:
        f.createVariable("eb_names", "c", ("num_el_blk", "len_name"))[:] = \
            _names([b[1] for b in blocks])
    if packed_coord:
        f.createVariable("coord", "d", ("num_dim", "num_nodes"))[:] = coords.T
    else:
        for d, nm in enumerate(("coordx", "coordy", "coordz")):
            f.createVariable(nm, "d", ("num_nodes",))[:] = coords[:, d]
    for k, (bid, _, etype, c0) in enumerate(blocks, 1):
        f.createDimension("num_el_in_blk%d" % k, c0.shape[0])
        f.createDimension("num_nod_per_el%d" % k, c0.shape[1])
        v = f.createVariable("connect%d" % k, "i", ("num_el_in_blk%d" % k,
                                                    "num_nod_per_el%d" % k))
        v[:] = c0 + 1
        v.elem_type = etype
    for kind, sets in (("ns", node_sets), ("ss", side_sets)):
        if not sets:
            continue
        f.createDimension("num_node_sets" if kind == "ns" else "num_side_sets", len(sets))
        dim = "num_node_sets" if kind == "ns" else "num_side_sets"
        p = f.createVariable(kind + "_prop1", "i", (dim,))
        p[:] = [s[0] for s in sets]
        p.name = b"ID"
        if names:
            f.createVariable(kind + "_names", "c", (dim, "len_name"))[:] = \
                _names([s[1] for s in sets])
        for k, s in enumerate(sets, 1):
            if kind == "ns":
                f.createDimension("num_nod_ns%d" % k, len(s[2]))
                f.createVariable("node_ns%d" % k, "i", ("num_nod_ns%d" % k,))[:] = s[2] + 1
            else:
                f.createDimension("num_side_ss%d" % k, len(s[2]))
                f.createVariable("elem_ss%d" % k, "i", ("num_side_ss%d" % k,))[:] = s[2] + 1
                f.createVariable("side_ss%d" % k, "i", ("num_side_ss%d" % k,))[:] = s[3] + 1
    if nodvars:
        f.createDimension("num_nod_var", len(nodvars))
        f.createVariable("name_nod_var", "c", ("num_nod_var", "len_name"))[:] = \
            _names([v[0] for v in nodvars])
        for k, (_, vals) in enumerate(nodvars, 1):
            v = f.createVariable("vals_nod_var%d" % k, "d", ("time_step", "num_nodes"))
            v[0:vals.shape[0]] = vals
    f.close()


def _blocks_of(types, conn, split=2):
    """The mesh's elements as blocks: each topology split into `split`
    blocks, the blocks interleaved by topology; returns the blocks and the
    element order they imply."""
    blocks, order = [], []
    parts = []
    for t in range(4):
        idx = np.where(types == t)[0]
        for s, chunk in enumerate(np.array_split(idx, split)):
            if chunk.size:
                parts.append((s, t, chunk))
    parts.sort(key=lambda p: (p[0], p[1]))
    for k, (_, t, chunk) in enumerate(parts):
        blocks.append((100 + k, "blk_%d_%d" % (t, k), ETYPE[t], conn[chunk, :NPE[t]]))
        order.append(chunk)
    return blocks, np.concatenate(order)


@pytest.mark.parametrize("version,packed", [(1, False), (2, False), (2, True)])
def test_reads_files_written_by_scipy(nw, mesh, tmp_path, version, packed):
    coords, types, conn = mesh
    n = coords.shape[0]
    blocks, order = _blocks_of(types, conn)
    rng = np.random.default_rng(5)
    top = np.where(coords[:, 2] >= 1 - 1e-12)[0].astype(np.int32)
    some = rng.choice(n, 17, replace=False).astype(np.int32)
    ss_elem = rng.choice(types.size, 9, replace=False).astype(np.int32)
    ss_side = np.array([rng.integers(0, {0: 6, 1: 4, 2: 5, 3: 5}[types[e]])
                        for e in ss_elem], np.int32)
    pos = np.empty(types.size, np.int64)
    pos[order] = np.arange(types.size)
    vals = rng.normal(size=(2, n))
    path = str(tmp_path / "scipy.exo")
    _scipy_exodus(path, coords, blocks,
                  node_sets=[(10, "top", top), (20, "some", some)],
                  side_sets=[(7, "faces", pos[ss_elem], ss_side)],
                  nodvars=[("pressure", vals), ("temperature", 2 * vals)],
                  times=[0.25, 0.75], version=version, packed_coord=packed)
    with nw.ExodusMesh(path) as x:
        assert x.format == version and x.nnode == n and x.nelem == types.size
        assert [b["id"] for b in x.blocks] == [b[0] for b in blocks]
        assert [b["name"] for b in x.blocks] == [b[1] for b in blocks]
        assert [b["topo"] for b in x.blocks] == [int(b[1].split("_")[1]) for b in blocks]
        c, t, cn, blk = x.read_mesh()
        np.testing.assert_array_equal(c, coords)  # bitwise
        np.testing.assert_array_equal(t, types[order])
        np.testing.assert_array_equal(cn, conn[order])
        np.testing.assert_array_equal(np.bincount(blk), [b[3].shape[0] for b in blocks])
        assert [s["name"] for s in x.node_sets] == ["top", "some"]
        assert [s["id"] for s in x.node_sets] == [10, 20]
        np.testing.assert_array_equal(x.node_set("top"), top)
        np.testing.assert_array_equal(x.node_set(1), some)
        e, s = x.side_set("faces")
        np.testing.assert_array_equal(order[e], ss_elem)
        np.testing.assert_array_equal(s, ss_side)
        assert x.nodal_vars == ["pressure", "temperature"] and x.ntime == 2
        v, tm = x.nodal_var("temperature", 1)
        np.testing.assert_array_equal(v, 2 * vals[1])
        assert tm == 0.75
        v, tm = x.nodal_var(0, 0)
        np.testing.assert_array_equal(v, vals[0])
        assert tm == 0.25


def test_files_written_by_the_library_read_by_scipy(nw, mesh, tmp_path):
    coords, types, conn = mesh
    n = coords.shape[0]
    path = str(tmp_path / "lib.exo")
    top = np.where(coords[:, 2] >= 1 - 1e-12)[0].astype(np.int32)
    elem = np.array([0, 5, 9], np.int32)
    side = np.array([0, 1, 2], np.int32)
    p = np.sin(np.arange(n, dtype=float))
    order = nw.exo_write(path, coords, types, conn, node_sets={"top": top, "empty": []},
                         side_sets={"wall": (elem, side)},
                         fields={"pressure": p, "pressure_x2": 2 * p}, time=3.5,
                         title="mixed box")
    np.testing.assert_array_equal(order, np.argsort(types, kind="stable"))
    f = netcdf_file(path, "r", mmap=False)
    try:
        assert f.version_byte == 2
        V = f.variables
        np.testing.assert_array_equal(V["coordx"][:], coords[:, 0])
        np.testing.assert_array_equal(V["coordz"][:], coords[:, 2])
        present = [t for t in range(4) if np.any(types == t)]
        assert f.dimensions["num_el_blk"] == len(present)
        for k, t in enumerate(present, 1):
            cv = V["connect%d" % k]
            assert cv.elem_type == ETYPE[t]
            np.testing.assert_array_equal(cv[:], conn[types == t, :NPE[t]] + 1)
        np.testing.assert_array_equal(V["node_ns1"][:], top + 1)
        assert "node_ns2" not in V  # an empty set has no entries
        np.testing.assert_array_equal(order[V["elem_ss1"][:] - 1], elem)
        np.testing.assert_array_equal(V["side_ss1"][:], side + 1)
        names = [r.tobytes().rstrip(b"\0").decode() for r in V["name_nod_var"][:]]
        assert names == ["pressure", "pressure_x2"]
        np.testing.assert_array_equal(V["vals_nod_var2"][0], 2 * p)
        assert V["time_whole"][0] == 3.5
        assert f.title == b"mixed box"
    finally:
        f.close()
    # and back through the library
    with nw.ExodusMesh(path) as x:
        c, t, cn, _ = x.read_mesh()
        np.testing.assert_array_equal(c, coords)
        np.testing.assert_array_equal(t, types[order])
        np.testing.assert_array_equal(cn, conn[order])
        assert [s["name"] for s in x.node_sets] == ["top", "empty"]
        assert x.node_sets[1]["count"] == 0 and x.node_set("empty").size == 0
        e, s = x.side_set("wall")
        np.testing.assert_array_equal(order[e], elem)
        np.testing.assert_array_equal(s, side)
        np.testing.assert_array_equal(x.nodal_var("pressure_x2")[0], 2 * p)


def _face_area(x):
    """area vector of a planar-or-not polygon by the right-hand rule (the
    triangle fan about its centroid; for a quad this equals half the cross
    product of its diagonals)"""
    c = x.mean(axis=0)
    a = np.zeros(3)
    for k in range(len(x)):
        a += 0.5 * np.cross(x[k] - c, x[(k + 1) % len(x)] - c)
    return a


def test_side_tables_close_and_point_outward(nw, mesh):
    """Every element's sides form a closed, consistently oriented surface:
    each directed edge of the faces appears once and its reverse once, the
    face area vectors sum to zero, and each points away from the element's
    centroid.  A wrong node in a face breaks the closure, a face listed
    clockwise the orientation."""
    coords, types, conn = mesh
    nsides = {0: 6, 1: 4, 2: 5, 3: 5}
    for t in range(4):
        faces = [nw.side_nodes(t, s) for s in range(nsides[t])]
        directed = [(f[k], f[(k + 1) % len(f)]) for f in faces for k in range(len(f))]
        assert len(set(directed)) == len(directed)
        assert all((b, a) in directed for a, b in directed)
        assert sorted({v for f in faces for v in f}) == list(range(NPE[t]))
        with pytest.raises(nw.NwError):
            nw.side_nodes(t, nsides[t])
        for e in np.where(types == t)[0][:50]:
            x = coords[conn[e, :NPE[t]]]
            cen = x.mean(axis=0)
            total = np.zeros(3)
            for f in faces:
                a = _face_area(x[f])
                total += a
                assert np.dot(a, x[f].mean(axis=0) - cen) > 0
            assert np.abs(total).max() < 1e-14


def test_boundary_side_sets_select_the_plane_nodes(nw, mesh):
    """The faces that belong to one element only are the boundary; those on
    z = 0, z = 1, x = 0 give exactly the nodes on that plane."""
    coords, types, conn = mesh
    nsides = {0: 6, 1: 4, 2: 5, 3: 5}
    owner = {}
    for e in range(types.size):
        for s in range(nsides[types[e]]):
            key = tuple(sorted(conn[e, nw.side_nodes(types[e], s)]))
            owner.setdefault(key, []).append((e, s))
    boundary = [v[0] for v in owner.values() if len(v) == 1]
    assert all(len(v) <= 2 for v in owner.values())
    for d, val in ((2, 0.0), (2, 1.0), (0, 0.0)):
        sel = [(e, s) for e, s in boundary
               if np.all(np.abs(coords[conn[e, nw.side_nodes(types[e], s)], d] - val) < 1e-12)]
        elem = np.array([e for e, _ in sel], np.int32)
        side = np.array([s for _, s in sel], np.int32)
        got = nw.sides_to_nodes(types, conn, coords.shape[0], elem, side)
        want = np.where(np.abs(coords[:, d] - val) < 1e-12)[0]
        np.testing.assert_array_equal(got, want)


def test_unsupported_and_malformed_files(nw, mesh, tmp_path):
    coords, types, conn = mesh
    d = str(tmp_path)
    with pytest.raises(nw.NwError, match="cannot open"):
        nw.ExodusMesh(os.path.join(d, "missing.exo"))
    for name, head, msg in (("hdf5.exo", b"\x89HDF\r\n\x1a\n" + bytes(64), "HDF5"),
                            ("cdf5.exo", b"CDF\x05" + bytes(64), "version 5"),
                            ("text.exo", b"hello world" * 10, "not a netCDF")):
        with open(os.path.join(d, name), "wb") as f:
            f.write(head)
        with pytest.raises(nw.NwError, match=msg):
            nw.ExodusMesh(os.path.join(d, name))
    # a truncated file: header intact, data cut short
    path = os.path.join(d, "whole.exo")
    nw.exo_write(path, coords, types, conn)
    raw = open(path, "rb").read()
    cut = os.path.join(d, "cut.exo")
    open(cut, "wb").write(raw[:len(raw) // 2])
    with nw.ExodusMesh(cut) as x:
        with pytest.raises(nw.NwError, match="past the end"):
            x.read_mesh()
    open(cut, "wb").write(raw[:40])
    with pytest.raises(nw.NwError, match="malformed"):
        nw.ExodusMesh(cut)
    # a block of an element type the edge path does not take
    quad = os.path.join(d, "shell.exo")
    _scipy_exodus(quad, coords, [(1, "shells", b"SHELL4", conn[:4, :4]),
                                 (2, "hexes", b"HEX8", conn[types == 0][:3])])
    with nw.ExodusMesh(quad) as x:
        assert [b["topo"] for b in x.blocks] == [-1, 0]
        with pytest.raises(nw.NwError, match="SHELL4"):
            x.read_mesh()
    # a node number outside the mesh
    bad = os.path.join(d, "bad.exo")
    hc, c0 = nw.hex_box(3, 1, 1)
    c0[1, 4] = hc.shape[0]
    _scipy_exodus(bad, hc, [(1, "hexes", b"HEX8", c0)])
    with nw.ExodusMesh(bad) as x:
        with pytest.raises(nw.NwError, match="outside"):
            x.read_mesh()
    with pytest.raises(nw.NwError):
        nw.exo_write(os.path.join(d, "w.exo"), coords, np.full(types.size, 4, np.int32), conn)
    with pytest.raises(nw.NwError, match="out of range"):
        nw.exo_write(os.path.join(d, "w.exo"), coords, types, conn,
                     node_sets={"x": [coords.shape[0]]})
    with pytest.raises(nw.NwError, match="no node set"):
        with nw.ExodusMesh(path) as x:
            x.node_set("nope")
