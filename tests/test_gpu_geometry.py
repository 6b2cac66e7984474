"""GeometryInteriorAlg on the device (nw_mesh_elements +
nw_compute_geometry): edge area vectors and dual volumes of Hex8 boxes,
single pyramids and conforming mixed Hex8/Tet4/Wed6/Pyr5 meshes against the
oracle, the reference unit-cube golds and the reference pyramid's closed
forms, for any edge order and orientation, and after the nodes move."""
import json
import os

import numpy as np
import pytest

from oracle import orc
from tests import fixtures as fx
from tests import problems as pb

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_golds.json")))


@pytest.fixture(scope="module")
def nw():
    return pb.nw()


@pytest.fixture
def ctx(nw):
    c = nw.Context(0)
    yield c
    c.close()


def _device_geometry(nw, ctx, coords, types, conn, edges):
    ctx.mesh_set(coords.shape[0], edges)
    ctx.field_upload(nw.F_COORDINATES, coords)
    ctx.mesh_elements(types, conn)
    ctx.compute_geometry()
    return (ctx.field_download(nw.E_AREA_VECTOR),
            ctx.field_download(nw.F_DUAL_VOLUME))


def test_geometry_unit_cube_gold(nw, ctx):
    """UnitTestGeometryAlg.C:20-100: dual volume 0.125 at every node of the
    unit hex, |A|^2 = 0.0625 on all 12 edges."""
    g = GOLD["geometry_unit_hex"]
    coords, conn, edges, _, _ = fx.unit_hex()
    a, v = _device_geometry(nw, ctx, coords, None, conn, edges)
    np.testing.assert_allclose(v, g["dual_volume"], rtol=0, atol=g["tol"])
    np.testing.assert_allclose((a ** 2).sum(1), g["area_mag_sq"], rtol=0,
                               atol=g["tol"])


@pytest.mark.parametrize("shape", [(6, 5, 4, 1.1, 0.2, 3), (9, 9, 9, 1.05, 0.1, 4)])
def test_geometry_hex_box_matches_oracle(nw, ctx, shape):
    nx, ny, nz, st, jit, seed = shape
    coords, conn = nw.hex_box(nx, ny, nz, st, jit, seed)
    edges, elem_edges = nw.hex_edges(conn, coords.shape[0])
    a, v = _device_geometry(nw, ctx, coords, None, conn, edges)
    ra = orc.edge_area_vec_hex(conn, coords, elem_edges, edges)
    rv = orc.dual_volume_hex(conn, coords)
    scale_a = np.abs(ra).max()
    np.testing.assert_allclose(a, ra, rtol=0, atol=1e-14 * scale_a)
    np.testing.assert_allclose(v, rv, rtol=1e-13, atol=0)
    # the host producers sum the same terms in the same (element) order
    np.testing.assert_allclose(a, nw.hex_edge_area_vector(conn, coords, elem_edges, edges),
                               rtol=0, atol=1e-15 * scale_a)
    np.testing.assert_allclose(v, nw.hex_dual_volume(conn, coords), rtol=1e-14, atol=0)


def test_geometry_pyramids_match_closed_forms_and_oracle(nw, ctx):
    """The reference pyramid (UnitTestUtils.C:226-235) and a distorted one
    sharing its base: dual volumes 32/135 at the base nodes and 52/135 at
    the apex of the first, each surface pair of an apex edge summed onto
    that edge, all against the oracle."""
    ref = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0], [0, 0, 1],
                    [0.3, -0.2, -1.1]], float)
    conn = np.full((2, 8), -1, np.int32)
    conn[0, :5] = [0, 1, 2, 3, 4]
    conn[1, :5] = [0, 3, 2, 1, 5]  # below the base: base reversed
    t = np.array([3, 3], np.int32)
    edges = nw.mixed_edges(t, conn, 6)
    assert edges.shape[0] == 12
    a, v = _device_geometry(nw, ctx, ref, t, conn, edges)
    ra = orc.edge_area_vec_mixed(t, conn, ref, edges)
    rv = orc.dual_volume_mixed(t, conn, ref)
    np.testing.assert_allclose(a, ra, rtol=0, atol=1e-15)
    np.testing.assert_allclose(v, rv, rtol=1e-14, atol=0)
    own = orc.pyr_scv_volume(ref[:5])
    np.testing.assert_allclose(own, [32 / 135] * 4 + [52 / 135], rtol=1e-14)
    np.testing.assert_allclose(v[4], 52 / 135, rtol=1e-14)
    np.testing.assert_allclose(v[5], orc.pyr_scv_volume(ref[[0, 3, 2, 1, 5]])[4],
                               rtol=1e-14)
    assert abs(v.sum() - (4 / 3 + 4 * 1.1 / 3)) < 1e-13


@pytest.mark.parametrize("mix", [(0.6, 0.3), (0.0, 1.0), (0.0, 0.0)])
def test_geometry_mixed_matches_oracle(nw, ctx, mix):
    coords, t, conn = nw.mixed_box(7, mix[0], mix[1], 0.1, seed=5)
    edges = nw.mixed_edges(t, conn, coords.shape[0])
    a, v = _device_geometry(nw, ctx, coords, t, conn, edges)
    ra = orc.edge_area_vec_mixed(t, conn, coords, edges)
    rv = orc.dual_volume_mixed(t, conn, coords)
    np.testing.assert_allclose(a, ra, rtol=0, atol=1e-14 * np.abs(ra).max())
    np.testing.assert_allclose(v, rv, rtol=1e-13, atol=0)
    assert (v > 0).all()


def test_geometry_any_edge_order_and_orientation(nw, ctx):
    """STK hands edges in its own order with node[0] not always the smaller
    id: the area vector follows the edge (sign flips with it)."""
    coords, t, conn = nw.mixed_box(5, 0.5, 0.3, 0.1, seed=8)
    edges = nw.mixed_edges(t, conn, coords.shape[0])
    a0, v0 = _device_geometry(nw, ctx, coords, t, conn, edges)
    rng = np.random.default_rng(1)
    perm = rng.permutation(edges.shape[0])
    flip = rng.random(edges.shape[0]) < 0.5
    e2 = edges[perm].copy()
    e2[flip] = e2[flip][:, ::-1]
    ctx2 = nw.Context(0)
    try:
        a1, v1 = _device_geometry(nw, ctx2, coords, t, conn, e2)
    finally:
        ctx2.close()
    sgn = np.where(flip, -1.0, 1.0)[:, None]
    np.testing.assert_array_equal(a1, a0[perm] * sgn)
    np.testing.assert_array_equal(v1, v0)


def test_geometry_reproducible_and_follows_the_nodes(nw, ctx):
    coords, t, conn = nw.mixed_box(6, 0.6, 0.3, 0.1, seed=2)
    edges = nw.mixed_edges(t, conn, coords.shape[0])
    a0, v0 = _device_geometry(nw, ctx, coords, t, conn, edges)
    ctx.compute_geometry()
    np.testing.assert_array_equal(ctx.field_download(nw.E_AREA_VECTOR), a0)
    np.testing.assert_array_equal(ctx.field_download(nw.F_DUAL_VOLUME), v0)
    # moved nodes: x -> 2 x scales areas by 4 and volumes by 8 exactly
    ctx.field_upload(nw.F_COORDINATES, 2.0 * coords)
    ctx.compute_geometry()
    np.testing.assert_array_equal(ctx.field_download(nw.E_AREA_VECTOR), 4.0 * a0)
    np.testing.assert_array_equal(ctx.field_download(nw.F_DUAL_VOLUME), 8.0 * v0)


def test_geometry_errors(nw, ctx):
    coords, conn = nw.hex_box(2, 2, 2, 1.0, 0.0, 0)
    edges, _ = nw.hex_edges(conn, coords.shape[0])
    ctx.mesh_set(coords.shape[0], edges)
    ctx.field_upload(nw.F_COORDINATES, coords)
    with pytest.raises(RuntimeError):
        ctx.compute_geometry()  # no elements yet
    with pytest.raises(RuntimeError):
        ctx.mesh_elements(None, np.full_like(conn, 99))  # node out of range
    bad = conn.copy()
    bad[0, [1, 6]] = bad[0, [6, 1]]  # a diagonal is not an edge of the mesh
    with pytest.raises(RuntimeError):
        ctx.mesh_elements(None, bad)
    with pytest.raises(RuntimeError):
        ctx.mesh_elements(np.full(conn.shape[0], 4, np.int32), conn)  # no type 4
