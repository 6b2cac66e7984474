"""The edge algorithms around the pressure solve (MdotEdgeAlg,
NodalGradEdgeAlg) through the C ABI against the oracle, on stretched and
jittered Hex8 boxes and a conforming mixed Hex8/Tet4/Wed6/Pyr5 mesh."""
import numpy as np
import pytest

from oracle import orc
from tests import problems as pb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nw():
    return pb.nw()


@pytest.fixture
def ctx(nw):
    c = nw.Context(0)
    yield c
    c.close()


CASES = [dict(nx=7, ny=5, nz=6, stretch=1.1, jitter=0.2, seed=3),
         dict(nx=12, ny=12, nz=12, stretch=1.05, jitter=0.1, seed=4)]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("noc,interp", [(1.0, 1.0), (0.0, 1.0), (1.0, 0.0)])
def test_mdot_matches_oracle(nw, ctx, case, noc, interp):
    P = pb.Problem(**case)
    P.upload(ctx)
    ctx.compute_mdot(noc, interp)
    got = ctx.field_download(nw.E_MASS_FLOW_RATE)
    ref = orc.mdot_edge_alg(P.edges, P.coords, P.vel, P.dpdx, P.rho, P.p,
                            P.udiag, P.areav, noc, interp)
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


def test_mdot_is_the_continuity_flux_on_device(nw, ctx):
    """The device continuity rhs equals the signed sum of the device mdot
    over each node's edges divided by dt/gamma1 (denScale = 1)."""
    P = pb.Problem(9, 8, 7, stretch=1.08, jitter=0.15, seed=5)
    P.upload(ctx)
    ctx.build_graph(1)
    dt = 0.25
    ctx.assemble_continuity(nw.continuity_defaults(dt=dt, gamma1=1.0))
    rhs = ctx.download()[3]
    ctx.compute_mdot(1.0, 1.0)
    md = ctx.field_download(nw.E_MASS_FLOW_RATE)
    ref = np.zeros(P.n)
    np.add.at(ref, P.edges[:, 0], -md / dt)
    np.add.at(ref, P.edges[:, 1], md / dt)
    np.testing.assert_allclose(rhs, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("case", CASES)
def test_nodal_grad_matches_oracle(nw, ctx, case):
    P = pb.Problem(**case)
    P.upload(ctx)
    ctx.build_graph(1)
    for fin, fout, val in ((nw.F_PRESSURE, nw.F_DPDX, P.p),
                           (nw.F_VELOCITY, nw.F_DUDX, P.vel),
                           (nw.F_SCALAR_Q, nw.F_DQDX, P.q)):
        ctx.nodal_grad(fin, fout)
        got = ctx.field_download(fout)
        ref = orc.nodal_grad_edge(P.edges, P.areav, P.dualvol, val)
        np.testing.assert_allclose(got.reshape(ref.shape), ref, rtol=0,
                                   atol=1e-12 * np.abs(ref).max())
        # accumulate: added to what the field holds (a boundary part, here
        # random), as the reference's algorithms add into a zeroed field
        g0 = np.random.default_rng(1).normal(size=ref.shape)
        ctx.field_upload(fout, g0)
        ctx.nodal_grad(fin, fout, accumulate=1)
        np.testing.assert_allclose(ctx.field_download(fout).reshape(ref.shape),
                                   g0 + ref, rtol=0,
                                   atol=1e-12 * np.abs(ref).max() + 1e-15)


def test_nodal_grad_exact_for_linear_field_and_reproducible(nw, ctx):
    """Interior nodes of the uniform box (every sub-control-surface centroid
    at its edge midpoint): the edge gradient of a linear field is exact;
    two calls give identical bits; the momentum (3-dof) graph gives the same
    result."""
    P = pb.Problem(8, 8, 8, stretch=1.0, jitter=0.0, seed=6)
    g = np.array([0.7, -1.3, 2.1])
    P.p = P.coords @ g + 0.4
    P.upload(ctx)
    ctx.build_graph(1)
    ctx.nodal_grad(nw.F_PRESSURE, nw.F_DPDX)
    a = ctx.field_download(nw.F_DPDX).reshape(-1, 3)
    ctx.nodal_grad(nw.F_PRESSURE, nw.F_DPDX)
    b = ctx.field_download(nw.F_DPDX).reshape(-1, 3)
    assert np.array_equal(a, b)
    lo, hi = P.coords.min(0), P.coords.max(0)
    inner = np.all((P.coords > lo + 1e-9) & (P.coords < hi - 1e-9), axis=1)
    np.testing.assert_allclose(a[inner], np.tile(g, (inner.sum(), 1)), rtol=0,
                               atol=1e-11)
    ctx.build_graph(3)
    ctx.nodal_grad(nw.F_PRESSURE, nw.F_DPDX)
    c = ctx.field_download(nw.F_DPDX).reshape(-1, 3)
    assert np.array_equal(a, c)


def test_edge_algs_on_mixed_mesh(nw, ctx):
    P = pb.Problem(6, 6, 6, jitter=0.1, seed=7, mixed=(0.6, 0.3))
    rng = np.random.default_rng(8)
    P.dualvol = rng.uniform(0.5, 2.0, P.n)
    P.upload(ctx)
    ctx.build_graph(1)
    ctx.compute_mdot(1.0, 1.0)
    ref = orc.mdot_edge_alg(P.edges, P.coords, P.vel, P.dpdx, P.rho, P.p,
                            P.udiag, P.areav)
    np.testing.assert_allclose(ctx.field_download(nw.E_MASS_FLOW_RATE), ref,
                               rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    ctx.nodal_grad(nw.F_VELOCITY, nw.F_DUDX)
    ref = orc.nodal_grad_edge(P.edges, P.areav, P.dualvol, P.vel)
    np.testing.assert_allclose(ctx.field_download(nw.F_DUDX).reshape(ref.shape),
                               ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_edge_alg_errors(nw, ctx):
    P = pb.Problem(4, 4, 4, seed=9)
    P.upload(ctx)
    with pytest.raises(nw.NwError) as e:   # no graph yet
        ctx.nodal_grad(nw.F_PRESSURE, nw.F_DPDX)
    assert e.value.code == nw.E_STATE
    ctx.build_graph(1)
    for fin, fout in ((nw.F_PRESSURE, nw.F_DUDX), (nw.F_PRESSURE, nw.F_PRESSURE),
                      (nw.F_DUDX, nw.F_DPDX), (nw.E_AREA_VECTOR, nw.F_DPDX)):
        with pytest.raises(nw.NwError) as e:
            ctx.nodal_grad(fin, fout)
        assert e.value.code == nw.E_ARG


def test_mdot_reference_gold(nw, ctx):
    """UnitTestMdotAlg.C:21-78 through the C ABI: unit-cube Hex8, density 1,
    velocity 10, zero pressure and gradient -> mdot = 2.5 on all 12 edges."""
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                    "reference_golds.json")))["mdot_edge_unit_hex"]
    coords, conn = nw.hex_box(1, 1, 1, 1.0, 0.0, 0)
    n = coords.shape[0]
    edges, elem_edges = nw.hex_edges(conn, n)
    areav = nw.hex_edge_area_vector(conn, coords, elem_edges, edges)
    ctx.mesh_set(n, edges)
    for fid, v in ((nw.F_COORDINATES, coords), (nw.E_AREA_VECTOR, areav),
                   (nw.F_VELOCITY, np.full((n, 3), g["velocity"])),
                   (nw.F_DPDX, np.full((n, 3), g["dpdx"])),
                   (nw.F_DENSITY, np.full(n, g["density"])),
                   (nw.F_PRESSURE, np.full(n, g["pressure"])),
                   (nw.F_UDIAG, np.ones(n))):
        ctx.field_upload(fid, v)
    ctx.compute_mdot(1.0, g["interp_together"])
    md = ctx.field_download(nw.E_MASS_FLOW_RATE)
    assert md.size == g["num_edges"]
    np.testing.assert_allclose(md, g["mdot"], rtol=0, atol=g["tol"])


def test_mdot_in_schedule_order_equals_edge_order(nw):
    """With the edge graph built, MdotEdgeAlg visits the edges in the row
    schedule's order; every edge is still computed once, with the same
    operations, so the result equals the edge-order pass bit for bit (on a
    permuted, half-flipped edge list as STK hands them)."""
    P = pb.Problem(12, 10, 11, stretch=1.08, jitter=0.2, seed=11)
    rng = np.random.default_rng(5)
    perm = rng.permutation(P.ne)
    flip = rng.random(P.ne) < 0.5
    e2 = P.edges[perm].copy()
    e2[flip] = e2[flip][:, ::-1]
    a2 = P.areav[perm] * np.where(flip, -1.0, 1.0)[:, None]
    out = []
    for graph in (False, True):
        c = nw.Context(0)
        try:
            c.mesh_set(P.n, e2)
            for fid, a in [(nw.F_COORDINATES, P.coords), (nw.F_VELOCITY, P.vel),
                           (nw.F_DPDX, P.dpdx), (nw.F_DENSITY, P.rho),
                           (nw.F_PRESSURE, P.p), (nw.F_UDIAG, P.udiag),
                           (nw.E_AREA_VECTOR, a2)]:
                c.field_upload(fid, a)
            if graph:
                c.build_graph(1)
            c.compute_mdot(1.0, 1.0)
            out.append(c.field_download(nw.E_MASS_FLOW_RATE))
        finally:
            c.close()
    assert np.array_equal(out[0], out[1])
    ref = orc.mdot_edge_alg(e2, P.coords, P.vel, P.dpdx, P.rho, P.p, P.udiag, a2,
                            1.0, 1.0)
    np.testing.assert_allclose(out[1], ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
