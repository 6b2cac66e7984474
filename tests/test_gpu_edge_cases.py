"""Edge cases of the assembly path on the device, against the oracle: the
smallest mesh (one edge), nodes with no edges and an empty edge list (rows
holding only their diagonal slot), hub rows longer than the kernels' LDS
slot staging (2,100 neighbours: the rows' values go straight to HBM),
malformed edge lists, and a zero-length edge (non-finite coefficients,
caught by the opt-in finite check)."""
import numpy as np
import pytest

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


def _all_three_match(nw, ctx, P):
    """continuity, scalar (limiter on) and momentum on P against the oracle"""
    P.upload(ctx)
    ctx.build_graph(1)
    o = nw.continuity_defaults()
    ctx.assemble_continuity(o)
    pb.assert_csr_close(ctx.download(), P.oracle_continuity(pb.to_orc_continuity(o)))
    s = nw.scalar_defaults()
    s.use_limiter = 1
    ctx.assemble_scalar(s)
    pb.assert_csr_close(ctx.download(), P.oracle_scalar(pb.to_orc_scalar(s)))
    ctx.build_graph(3)
    m = nw.momentum_defaults(alpha=0.5)
    ctx.assemble_momentum(m)
    pb.assert_csr_close(ctx.download(), P.oracle_momentum(pb.to_orc_momentum(m)))


def test_one_edge(nw, ctx):
    coords = np.array([[0.0, 0.0, 0.0], [0.3, 0.1, -0.2]])
    P = pb.Problem.from_graph(coords, [[0, 1]], [[0.5, 0.2, -0.1]], seed=1)
    _all_three_match(nw, ctx, P)


def test_nodes_without_edges(nw, ctx):
    """Nodes 0, 4 and 7 take part in no edge: their rows hold only the
    diagonal slot, which no edge writes (zero block, zero rhs)."""
    rng = np.random.default_rng(2)
    coords = rng.normal(size=(9, 3))
    edges = np.array([[1, 2], [2, 3], [1, 5], [5, 6], [3, 8], [6, 8]], np.int32)
    areav = (coords[edges[:, 1]] - coords[edges[:, 0]]) * rng.uniform(0.5, 1.5, (6, 1))
    P = pb.Problem.from_graph(coords, edges, areav, seed=2)
    _all_three_match(nw, ctx, P)
    rp, cols, vals, rhs = ctx.download()  # momentum, point rows 3 r + d
    for r in (0, 4, 7):
        for i in range(3 * r, 3 * r + 3):
            np.testing.assert_array_equal(cols[rp[i]:rp[i + 1]], [3 * r, 3 * r + 1, 3 * r + 2])
            assert not np.any(vals[rp[i]:rp[i + 1]]) and rhs[i] == 0


def test_empty_edge_list(nw, ctx):
    rng = np.random.default_rng(3)
    P = pb.Problem.from_graph(rng.normal(size=(5, 3)), np.zeros((0, 2), np.int32),
                              np.zeros((0, 3)), seed=3)
    _all_three_match(nw, ctx, P)
    rp, cols, vals, rhs = ctx.download()
    np.testing.assert_array_equal(rp, 3 * np.arange(16))  # point rows of 3 entries
    assert not vals.any() and not rhs.any()


@pytest.mark.parametrize("nleaf", [2100])
def test_hub_rows_beyond_the_slot_staging(nw, ctx, nleaf):
    """Two hubs joined to each other and to nleaf leaves each (a leaf shared
    by both every third node), plus a chain through the leaves: the hub
    rows hold more slots than the kernels stage in LDS (1,024 for the 1-dof
    systems, 224 for momentum), so their values are written straight to
    HBM; the leaf rows stay short."""
    rng = np.random.default_rng(4)
    n = nleaf + 2
    coords = rng.normal(size=(n, 3))
    coords[0] = [0.0, 0.0, 0.0]
    coords[1] = [0.2, 0.1, 0.0]
    e = [(0, 1)]
    e += [(0, k) for k in range(2, n)]
    e += [(1, k) for k in range(2, n, 3)]
    e += [(k, k + 1) for k in range(2, n - 1)]
    edges = np.array(e, np.int32)
    dx = coords[edges[:, 1]] - coords[edges[:, 0]]
    areav = dx * rng.uniform(0.5, 1.5, (len(e), 1)) + 0.1 * rng.normal(size=dx.shape)
    P = pb.Problem.from_graph(coords, edges, areav, seed=4)
    _all_three_match(nw, ctx, P)
    rp = ctx.download()[0]
    assert rp[1] - rp[0] == 3 * (nleaf + 2)  # hub: self + every other node


def test_malformed_edge_lists_are_errors(nw, ctx):
    with pytest.raises(nw.NwError):
        ctx.mesh_set(4, np.array([[0, 1], [2, 2]], np.int32))  # self edge
    with pytest.raises(nw.NwError):
        ctx.mesh_set(4, np.array([[0, 1], [2, 4]], np.int32))  # node 4 of 4
    with pytest.raises(nw.NwError):
        ctx.mesh_set(4, np.array([[0, 1], [-1, 2]], np.int32))
    ctx.mesh_set(4, np.array([[0, 1], [1, 2], [2, 1]], np.int32))
    with pytest.raises(nw.NwError):
        ctx.build_graph(1)  # the same edge twice (once reversed)


def test_zero_length_edge(nw, ctx):
    """Two coincident nodes: a . dx = 0, so the edge's coefficient is
    infinite in the reference as here; with the finite check on the
    assembly reports it instead of handing it on."""
    coords = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    edges = np.array([[0, 1], [1, 2], [0, 2]], np.int32)
    areav = np.array([[1.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    P = pb.Problem.from_graph(coords, edges, areav, seed=5)
    P.upload(ctx)
    ctx.build_graph(1)
    o = nw.continuity_defaults()
    ctx.assemble_continuity(o)
    vals = ctx.download()[2]
    assert not np.isfinite(vals).all()
    ctx.set_check_finite(True)
    with pytest.raises(nw.NwError, match="non-finite"):
        ctx.assemble_continuity(o)
