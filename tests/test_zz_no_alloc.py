"""The allocation-free sumInto (nw_linsys_sum_into_device_noalloc) against
nw_linsys_sum_into_device, through the C ABI: two calls in a row, the
second smaller, and an over-size call that is refused without applying
anything."""
import numpy as np
import pytest

from tests import problems as pb

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ndof", [1, 3])
def test_noalloc_sum_into_matches_the_allocating_entry(ndof):
    nw = pb.nw()
    P = pb.Problem(9, 7, 8, stretch=1.15, jitter=0.25, seed=4)
    rng = np.random.default_rng(70 + ndof)
    o = nw.continuity_defaults() if ndof == 1 else nw.momentum_defaults(alpha=0.2)
    ctxs = [nw.Context(0), nw.Context(0)]
    bufs = []
    try:
        for c in ctxs:
            P.upload(c)
            c.build_graph(ndof)
            if ndof == 1:
                c.assemble_continuity(o)
            else:
                c.assemble_momentum(o)
        # two calls in a row on the same fixed scratch, the second smaller:
        # what the first left in the scratch must not leak into the second
        for nb in (5000, 700):
            nodes = P.edges[rng.integers(0, P.ne, nb)].astype(np.int32)
            nodes[::3] = nodes[::3, ::-1]
            nr = 2 * ndof
            lhs = rng.normal(size=(nb, nr, nr))
            rhs = rng.normal(size=(nb, nr))
            b = [nw.DeviceBuffer(a) for a in (nodes, lhs, rhs)]
            bufs += b
            ctxs[0].sum_into_device(*b, nb, 2)
            ctxs[1].sum_into_device_noalloc(*b, nb, 2)
            ref, got = ctxs[0].download(), ctxs[1].download()
            for k in range(4):
                assert np.array_equal(got[k], ref[k]), (nb, k)
        # a column absent from the graph: an error, nothing applied
        before = ctxs[1].download()
        far = nw.DeviceBuffer(np.array([[0, P.n - 1]], np.int32))
        one = [nw.DeviceBuffer(np.ones((1, 2 * ndof, 2 * ndof))),
               nw.DeviceBuffer(np.ones((1, 2 * ndof)))]
        bufs += [far] + one
        with pytest.raises(nw.NwError) as ei:
            ctxs[1].sum_into_device_noalloc(far, *one, 1, 2)
        assert ei.value.code == nw.E_NOGRAPHCOL
        # more contributions than the fixed scratch holds: refused before
        # any kernel runs (the arrays are never read), nothing applied
        with pytest.raises(nw.NwError) as ei:
            ctxs[1].sum_into_device_noalloc(far, *one, 4_000_000, 2)
        assert ei.value.code == nw.E_ARG
        after = ctxs[1].download()
        assert all(np.array_equal(x, y) for x, y in zip(before, after))
    finally:
        for d in bufs:
            d.free()
        for c in ctxs:
            c.close()
