"""CPU checks of the C ABI library: it loads, it exports every entry point
the headers declare, and its host-side geometry producers agree with the
oracle.  No GPU call is made here."""
import os
import re
import sys

import numpy as np
import pytest

from oracle import orc
from tests import fixtures as fx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "nalu-wind_amd"))
import naluwind_amd as nw  # noqa: E402


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nw_[a-z0-9_]+)\s*\(", src)))


@pytest.mark.parametrize("header", ["nalu_amd.h", "nalu_amd_mesh.h"])
def test_every_declared_symbol_is_exported(header):
    names = _declared(header)
    assert names
    for name in names:
        assert hasattr(nw.lib, name), name
        assert name in nw.EXPORTED, "binding misses " + name


def test_hex_box_matches_fixture_generator():
    c1, k1 = nw.hex_box(3, 4, 5, stretch_z=1.07)
    c2, k2 = fx.hex_box(3, 4, 5, stretch_z=1.07)
    np.testing.assert_array_equal(k1, k2)
    np.testing.assert_allclose(c1, c2, rtol=0, atol=1e-15)


def test_hex_edges_are_sorted_unique_and_oriented():
    coords, conn = nw.hex_box(4, 3, 2)
    edges, ee = nw.hex_edges(conn, coords.shape[0])
    assert (edges[:, 0] < edges[:, 1]).all()
    key = edges[:, 0].astype(np.int64) * coords.shape[0] + edges[:, 1]
    assert (np.diff(key) > 0).all()
    # a box of nx*ny*nz cells has this many edges
    nx, ny, nz = 4, 3, 2
    ne = nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz
    assert edges.shape[0] == ne
    # same edge set as the oracle's first-encounter extraction
    oe, _ = orc.create_edges_hex(conn, coords.shape[0])
    ok = oe[:, 0].astype(np.int64) * coords.shape[0] + oe[:, 1]
    assert set(ok.tolist()) == set(key.tolist())


def test_geometry_producers_match_oracle():
    coords, conn = nw.hex_box(5, 4, 6, stretch_z=1.2, jitter=0.2, seed=9)
    n = coords.shape[0]
    edges, ee = nw.hex_edges(conn, n)
    av = nw.hex_edge_area_vector(conn, coords, ee, edges)
    dv = nw.hex_dual_volume(conn, coords)
    oe, oee = orc.create_edges_hex(conn, n)
    oav = orc.edge_area_vec_hex(conn, coords, oee, oe)
    odv = orc.dual_volume_hex(conn, coords)
    # map oracle edges onto product edges
    pos = {(a, b): i for i, (a, b) in enumerate(edges.tolist())}
    perm = np.array([pos[(a, b)] for a, b in oe.tolist()])
    np.testing.assert_allclose(av[perm], oav, rtol=0, atol=1e-15)
    np.testing.assert_allclose(dv, odv, rtol=0, atol=1e-15)
    phi = np.sin(coords[:, 0]) + coords[:, 1] * coords[:, 2]
    g = nw.nodal_grad_edge(edges, av, dv, phi)
    og = orc.nodal_grad_edge(oe, oav, odv, phi)
    np.testing.assert_allclose(g, og, rtol=1e-12, atol=1e-12)
    rho = 1.0 + coords[:, 0]
    vel = np.stack([coords[:, 1], -coords[:, 0], coords[:, 2] ** 2], 1)
    np.testing.assert_allclose(nw.mdot_edge(edges, av, rho, vel)[perm],
                               orc.mdot_edge(oe, oav, rho, vel), rtol=0,
                               atol=1e-15)


def test_geometry_is_thread_count_independent():
    import subprocess
    code = ("import sys; sys.path.insert(0, %r); import naluwind_amd as nw, numpy as np;"
            "c, k = nw.hex_box(6, 5, 7, 1.1, 0.2, 3); e, ee = nw.hex_edges(k, c.shape[0]);"
            "a = nw.hex_edge_area_vector(k, c, ee, e); d = nw.hex_dual_volume(k, c);"
            "print(a.tobytes().hex()[:64], abs(a).sum().hex(), d.sum().hex())"
            % os.path.join(ROOT, "nalu-wind_amd"))
    outs = []
    for t in ("1", "4"):
        env = dict(os.environ, OMP_NUM_THREADS=t)
        outs.append(subprocess.check_output([sys.executable, "-c", code], env=env))
    assert outs[0] == outs[1]


def test_bad_arguments_are_errors_not_crashes():
    with pytest.raises(nw.NwError):
        nw.hex_box(0, 1, 1)
    with pytest.raises(nw.NwError):
        nw.hex_edges(np.array([[0, 1, 2, 3, 4, 5, 6, 99]], np.int32), 8)
