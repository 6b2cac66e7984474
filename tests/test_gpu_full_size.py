"""The benchmark configurations at their full sizes, through the C ABI:

  momentum  128^3-cell box stretched 1.05 toward z = 0, +-10 % jitter
            (2,146,689 nodes, 6,390,144 edges), hybrid factor 1, alpha 0.5
  continuity 256^3-cell box with the same z grading (16,974,593 nodes,
            50,725,632 edges), Dirichlet p = 0 on the top face

Each system is compared with the oracle over every entry (row-scaled 1e-12,
the bar of the small-mesh tests) and checked against properties that hold
at any size and do not depend on the oracle:

  continuity  the LHS is a symmetric graph Laplacian: A = A^T, every row and
              column sums to zero; the fluxes are antisymmetric: sum(rhs) = 0
  momentum    the scheme is conservative (relaxation 1): every column of the
              LHS sums to zero (an edge adds -c to one row and +c to the
              other in the same column), and sum(rhs) = 0 per component

The pressure solves of the 256^3 system (SA-AMG, Jacobi and Chebyshev
preconditioning) are checked by their residual, recomputed on the host from
the downloaded system; the device geometry,
NodalGradEdgeAlg and MdotEdgeAlg of the 256^3 box against the oracle, and
the geometry by the tiling of the box and the closure of the area vectors;
the conforming mixed Hex8/Tet4/Wed6/Pyr5 box's continuity system and device
geometry.

The oracle runs on up to 16 host threads (its atomic scatter only reorders
the sums); each test takes 5-30 s on a 16-core host, most of it host mesh
generation and the oracle."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import orc
from tests import problems as pb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nw():
    return pb.nw()


@pytest.fixture(scope="module")
def wl(nw):
    import workloads
    return workloads


@pytest.fixture
def host_threads():
    orc.set_threads(min(16, os.cpu_count() or 1))
    yield
    orc.set_threads(1)


def _sparse(got, n):
    rp, cols, vals, _ = got
    return sp.csr_matrix((vals, cols, rp), shape=(n, n))


def _conservation(S, rhs, ncomp, rows=True, cols=True):
    """Row and column sums of S within 1e-13 of the row / column sums of
    |S|; sum(rhs) per component within 1e-12 of sum(|rhs|)."""
    A = abs(S)
    if rows:
        rs = np.asarray(S.sum(1)).ravel()
        ra = np.asarray(A.sum(1)).ravel()
        assert (np.abs(rs) <= 1e-13 * ra).all(), np.abs(rs / ra).max()
    if cols:
        cs = np.asarray(S.sum(0)).ravel()
        ca = np.asarray(A.sum(0)).ravel()
        assert (np.abs(cs) <= 1e-13 * ca).all(), np.abs(cs / ca).max()
    r = rhs.reshape(-1, ncomp)
    assert (np.abs(r.sum(0)) <= 1e-12 * np.abs(r).sum(0)).all(), r.sum(0)


@pytest.mark.timeout(600)
def test_continuity_256_cubed_full_size(nw, wl, host_threads):
    W = wl.continuity_rf(256)
    assert (W.n, W.ne) == (16974593, 50725632)
    f = W.fields
    A = orc.CsrSystem(W.n, W.edges, 1)
    orc.assemble_continuity(pb.to_orc_continuity(W.opts), W.n, W.edges,
                            W.coords, f[nw.F_VELOCITY], f[nw.F_DPDX],
                            f[nw.F_DENSITY], f[nw.F_PRESSURE], f[nw.F_UDIAG],
                            W.areav, csr=A)
    with nw.Context(0) as ctx:
        W.upload(ctx)
        ctx.build_graph(1)
        ctx.assemble_continuity(W.opts)
        got = ctx.download()
        print("continuity 256^3: %d rows, %d entries" % (W.n, got[1].size), flush=True)
        pb.assert_csr_close(got, A)
        S = _sparse(got, W.n)
        D = S - S.T
        assert D.nnz == 0 or abs(D).max() <= 1e-14 * abs(S).max()
        _conservation(S, got[3], 1)
        del S, D, got
        # the Dirichlet rows of the pressure solve
        wl.apply_pressure_bc(ctx, W)
        A.apply_dirichlet(W.top, 0, 1, np.zeros(W.n), np.zeros(W.n))
        pb.assert_csr_close(ctx.download(), A)


@pytest.mark.timeout(600)
def test_momentum_128_cubed_full_size(nw, wl, host_threads):
    W = wl.momentum_abl(128)
    assert (W.n, W.ne) == (2146689, 6390144)
    f = W.fields
    pf = W.peclet
    ref_pn, ref_pf = orc.momentum_edge_peclet(
        W.edges, W.coords, f[nw.F_VELOCITY], f[nw.F_DENSITY], f[nw.F_VISCOSITY],
        orc.Peclet(pf.form, pf.hybrid, pf.c1, pf.c2), with_number=True)
    with nw.Context(0) as ctx:
        W.upload(ctx)
        ctx.build_graph(3)
        ctx.momentum_edge_peclet(pf)
        pec = ctx.field_download(nw.E_PECLET_FACTOR)
        np.testing.assert_allclose(pec, ref_pf, rtol=1e-13, atol=1e-15)
        # the Peclet number itself, over all 6.4 M edges: to 1e-13 of the
        # magnitudes summed in u.dx (edges where u.dx cancels lose relative
        # accuracy in any arithmetic; the device fuses multiply-adds)
        E = W.edges
        dx = W.coords[E[:, 1]] - W.coords[E[:, 0]]
        us = f[nw.F_VELOCITY][E[:, 0]] + f[nw.F_VELOCITY][E[:, 1]]
        mu, rho = f[nw.F_VISCOSITY], f[nw.F_DENSITY]
        diff = 0.5 * (mu[E[:, 0]] / rho[E[:, 0]] + mu[E[:, 1]] / rho[E[:, 1]])
        scale = 0.5 * np.abs(dx * us).sum(1) / (diff + 1e-16)
        err = np.abs(ctx.field_download(nw.E_PECLET_NUMBER) - ref_pn)
        assert (err <= 1e-13 * scale).all(), (err / scale).max()
        assert np.median(err / np.maximum(ref_pn, 1e-300)) < 1e-15
        # the blend is exercised: the factor spans (0, 1)
        assert pec.min() < 0.1 and pec.max() > 0.9
        A = orc.CsrSystem(W.n, W.edges, 3)
        orc.assemble_momentum(pb.to_orc_momentum(W.opts), W.n, W.edges,
                              W.coords, f[nw.F_VELOCITY], f[nw.F_DUDX],
                              f[nw.F_VISCOSITY], f[nw.F_DENSITY], W.areav,
                              f[nw.E_MASS_FLOW_RATE], ref_pf, f[nw.F_WALL_MASK],
                              csr=A)
        ctx.assemble_momentum(W.opts)
        two = ctx.download()
        print("momentum 128^3: %d rows, %d entries" % (3 * W.n, two[1].size), flush=True)
        pb.assert_csr_close(two, A)
        del A
        assert W.opts.relax == 1.0
        _conservation(_sparse(two, 3 * W.n), two[3], 3, rows=False)
        # the one-pass form (Peclet factor in the assembly kernel) gives the
        # same system
        ctx.assemble_momentum_with_peclet(W.opts, pf)
        one = ctx.download()
        assert np.array_equal(one[0], two[0]) and np.array_equal(one[1], two[1])
        np.testing.assert_allclose(one[2], two[2], rtol=0,
                                   atol=1e-14 * np.abs(two[2]).max())
        np.testing.assert_allclose(one[3], two[3], rtol=0,
                                   atol=1e-14 * np.abs(two[3]).max())
        np.testing.assert_allclose(ctx.field_download(nw.E_PECLET_FACTOR), pec,
                                   rtol=1e-15, atol=1e-300)


@pytest.mark.timeout(600)
def test_pressure_solve_256_cubed_full_size(nw, wl):
    """The pressure solves of the benchmark at 256^3 (GMRES(50) + SA-AMG, tol
    1e-8): FP64 from x0 = 0 (the reference's method), the FP32-inner /
    FP64-refined form from x0 = 0 and from the network's guess (untrained:
    the quadratic extrapolation of the three previous solutions).  Every
    solution's residual is recomputed here in FP64 from the downloaded
    system, and the solutions agree to the accuracy the tolerance implies."""
    with nw.Context(0) as ctx:
        S = wl.PressureSeries(ctx, 256, seed=1234, dt_flow=0.02)
        base = dict(kspace=50, max_iters=2000, tol=1e-8, precond=nw.PREC_AMG,
                    reuse_preconditioner=1)
        fp64 = lambda x0: nw.solver_cfg(mode=nw.SOLVE_FP64, x0=x0, **base)
        fp32 = lambda x0: nw.solver_cfg(mode=nw.SOLVE_FP32_REFINE, inner_tol=1e-4,
                                        x0=x0, **base)
        # steps 0-3 give the guess its history of solutions
        for k in range(4):
            S.step(k)
            assert ctx.solve(fp64(nw.X0_ZERO if k == 0 else nw.X0_SOLUTION)).converged
        net = nw.Mpnn(ctx, 64, 3, seed=7)
        try:
            x_prev = ctx.solution()
            S.step(4)
            rp, cols, vals, b = ctx.download()
            A = sp.csr_matrix((vals, cols, rp), shape=(S.W.n, S.W.n))
            bn = np.linalg.norm(b)
            out = {}
            for name, cfg in [("learned", fp32(nw.X0_LEARNED)),
                              ("fp32_refined", fp32(nw.X0_ZERO)),
                              ("fp64", fp64(nw.X0_ZERO))]:
                ctx.set_solution(x_prev)
                ctx.set_learned_guess(net)
                st = ctx.solve(cfg)
                x = ctx.solution()
                rel = np.linalg.norm(b - A @ x) / bn
                print("256^3 %s: %d iterations, %.1f ms, residual %.3e (host %.3e)" %
                      (name, st.iters, st.ms, st.final_rel, rel), flush=True)
                assert st.converged and st.final_rel <= 1e-8 and rel <= 1.0001e-8
                out[name] = (st, x)
        finally:
            net.close()
        x64 = out["fp64"][1]
        for name in ("fp32_refined", "learned"):
            err = np.linalg.norm(out[name][1] - x64) / np.linalg.norm(x64)
            assert err < 1e-5, (name, err)
        assert out["learned"][0].guess == nw.X0_LEARNED
        assert out["learned"][0].iters < out["fp32_refined"][0].iters


@pytest.mark.timeout(600)
def test_pressure_solve_256_cubed_jacobi_and_chebyshev(nw, wl):
    """The configuration's own preconditioners at 256^3: GMRES(50) right-
    preconditioned by Jacobi and by Chebyshev(8) on D^-1 A, FP64 and
    FP32-inner / FP64-refined, x0 = 0, tol 1e-8, on the continuity system
    with its Dirichlet rows.  Chebyshev converges; its residual, recomputed
    here in FP64 from the downloaded system, meets the tolerance, and the
    FP32-refined solution agrees with the FP64 one.  Jacobi does not
    converge in its 2000 iterations on this graded mesh: the solve says so
    (converged = 0) and the residual it reports is the one recomputed
    here."""
    with nw.Context(0) as ctx:
        W = wl.continuity_rf(256)
        W.upload(ctx)
        ctx.build_graph(1)
        ctx.assemble_continuity(W.opts)
        wl.apply_pressure_bc(ctx, W)
        rp, cols, vals, b = ctx.download()
        A = sp.csr_matrix((vals, cols, rp), shape=(W.n, W.n))
        bn = np.linalg.norm(b)
        out = {}
        for name, pre, maxit in (
                ("chebyshev8", dict(precond=nw.PREC_CHEBYSHEV, cheb_degree=8,
                                    cheb_ratio=30.0), 4000),
                ("jacobi", dict(precond=nw.PREC_JACOBI), 2000)):
            for mode, mname in ((nw.SOLVE_FP64, "fp64"),
                                (nw.SOLVE_FP32_REFINE, "fp32_refined")):
                st = ctx.solve(nw.solver_cfg(kspace=50, max_iters=maxit, tol=1e-8,
                                             mode=mode, inner_tol=1e-5,
                                             x0=nw.X0_ZERO, **pre),
                               raise_on_noconv=False)
                x = ctx.solution()
                rel = np.linalg.norm(b - A @ x) / bn
                print("256^3 %s %s: %d iterations, %.0f ms, converged %d, residual "
                      "%.3e (host %.3e)" % (name, mname, st.iters, st.ms,
                                            st.converged, st.final_rel, rel),
                      flush=True)
                assert st.final_rel == pytest.approx(rel, rel=1e-3, abs=1e-14)
                out[(name, mname)] = (st, x)
        for mname in ("fp64", "fp32_refined"):
            st, _ = out[("chebyshev8", mname)]
            assert st.converged and st.final_rel <= 1e-8
            st, _ = out[("jacobi", mname)]
            assert not st.converged and st.iters == 2000 and st.final_rel > 1e-8
        x64 = out[("chebyshev8", "fp64")][1]
        x32 = out[("chebyshev8", "fp32_refined")][1]
        assert np.linalg.norm(x32 - x64) / np.linalg.norm(x64) < 1e-5


@pytest.mark.timeout(900)
def test_geometry_and_edge_algorithms_256_cubed_full_size(nw, wl, host_threads):
    """GeometryInteriorAlg, NodalGradEdgeAlg and MdotEdgeAlg on the device
    for the 256^3 box (16,777,216 Hex8 elements): against the oracle, and
    the geometry against what holds at any size: the dual volumes tile the
    unit box and the area vectors around every interior node close."""
    W = wl.continuity_rf(256, keep_elements=True)
    edges, elem_edges = nw.hex_edges(W.conn, W.n)
    assert np.array_equal(edges, W.edges)
    ra = orc.edge_area_vec_hex(W.conn, W.coords, elem_edges, edges)
    del elem_edges
    rv = orc.dual_volume_hex(W.conn, W.coords)
    f = W.fields
    x = W.coords
    p = np.sin(3.0 * x[:, 0]) * np.cos(2.0 * x[:, 1]) + x[:, 2] ** 2
    rdp = orc.nodal_grad_edge(edges, ra, rv, p)
    rmd = orc.mdot_edge_alg(edges, x, f[nw.F_VELOCITY], rdp, f[nw.F_DENSITY], p,
                            f[nw.F_UDIAG], ra)
    with nw.Context(0) as ctx:
        ctx.mesh_set(W.n, edges)
        ctx.field_upload(nw.F_COORDINATES, x)
        ctx.mesh_elements(None, W.conn)
        ctx.compute_geometry()
        a = ctx.field_download(nw.E_AREA_VECTOR)
        v = ctx.field_download(nw.F_DUAL_VOLUME)
        print("geometry 256^3: %d elements, %d edges" % (W.conn.shape[0], edges.shape[0]),
              flush=True)
        amax = np.abs(ra).max()
        np.testing.assert_allclose(a, ra, rtol=0, atol=1e-14 * amax)
        np.testing.assert_allclose(v, rv, rtol=1e-13, atol=0)
        assert abs(v.sum() - 1.0) <= 1e-12
        close = np.stack([np.bincount(edges[:, 0], weights=a[:, d], minlength=W.n) -
                          np.bincount(edges[:, 1], weights=a[:, d], minlength=W.n)
                          for d in range(3)], 1)
        lo, hi = x.min(0), x.max(0)
        inner = np.all((x > lo + 1e-12) & (x < hi - 1e-12), axis=1)
        assert np.abs(close[inner]).max() <= 1e-12 * amax
        del close
        # the edge algorithms on the device geometry
        for fid in (nw.F_VELOCITY, nw.F_DENSITY, nw.F_UDIAG):
            ctx.field_upload(fid, f[fid])
        ctx.field_upload(nw.F_PRESSURE, p)
        ctx.build_graph(1)
        ctx.nodal_grad(nw.F_PRESSURE, nw.F_DPDX)
        dp = ctx.field_download(nw.F_DPDX).reshape(rdp.shape)
        np.testing.assert_allclose(dp, rdp, rtol=0, atol=1e-12 * np.abs(rdp).max())
        ctx.compute_mdot(1.0, 1.0)
        md = ctx.field_download(nw.E_MASS_FLOW_RATE)
        np.testing.assert_allclose(md, rmd, rtol=1e-12, atol=1e-12 * np.abs(rmd).max())


@pytest.mark.timeout(900)
def test_mixed_160_cubed_full_size(nw, wl, host_threads):
    """The conforming mixed box of the benchmark (160^3 cells cut 60/30/10 %
    into Hex8 / Kuhn Tet4 / Wed6, transition cells of Pyr5 and Tet4 where
    faces would not match, +-10 % jitter, RCM-numbered: 4,328,769 nodes,
    19,200,110 edges): the continuity system against the oracle over every
    entry, symmetric with zero row sums and sum(rhs) = 0; the device
    geometry against the oracle (to 1e-13 of the largest area vector: the
    two sum an edge's sub-control-surface terms in different element
    orders), its dual volumes tiling the unit box and its area vectors
    closing around every interior node; the AMG pressure solves on it,
    checked by their residual recomputed on the host."""
    W = wl.continuity_mixed(160, keep_elements=True)
    assert (W.n, W.ne) == (4328769, 19200110)
    f = W.fields
    x = W.coords
    A = orc.CsrSystem(W.n, W.edges, 1)
    orc.assemble_continuity(pb.to_orc_continuity(W.opts), W.n, W.edges, x,
                            f[nw.F_VELOCITY], f[nw.F_DPDX], f[nw.F_DENSITY],
                            f[nw.F_PRESSURE], f[nw.F_UDIAG], W.areav, csr=A)
    ra = orc.edge_area_vec_mixed(W.types, W.conn, x, W.edges)
    rv = orc.dual_volume_mixed(W.types, W.conn, x)
    with nw.Context(0) as ctx:
        W.upload(ctx)
        ctx.build_graph(1)
        ctx.assemble_continuity(W.opts)
        got = ctx.download()
        print("mixed 160^3: %d rows, %d entries" % (W.n, got[1].size), flush=True)
        pb.assert_csr_close(got, A)
        S = _sparse(got, W.n)
        D = S - S.T
        assert D.nnz == 0 or abs(D).max() <= 1e-14 * abs(S).max()
        _conservation(S, got[3], 1)
        del S, D, got, A
        # the pressure solve on this mesh (a reverse Cuthill-McKee
        # numbering: the fine level in 16-bit delta form), FP64 and
        # FP32-refined, each residual recomputed here from the downloaded
        # system with its Dirichlet rows
        wl.apply_pressure_bc(ctx, W)
        rp, cols, vals, b = ctx.download()
        M = sp.csr_matrix((vals, cols, rp), shape=(W.n, W.n))
        bn = np.linalg.norm(b)
        for mode in (nw.SOLVE_FP64, nw.SOLVE_FP32_REFINE):
            st = ctx.solve(nw.solver_cfg(kspace=50, max_iters=500, tol=1e-8,
                                         precond=nw.PREC_AMG, mode=mode, inner_tol=1e-4,
                                         reuse_preconditioner=1))
            rel = np.linalg.norm(b - M @ ctx.solution()) / bn
            print("mixed 160^3 solve mode %d: %d iterations, %.1f ms, host residual %.3e"
                  % (mode, st.iters, st.ms, rel), flush=True)
            assert st.converged and rel <= 1.0001e-8
        del M, rp, cols, vals, b
        ctx.mesh_elements(W.types, W.conn)
        ctx.compute_geometry()
        a = ctx.field_download(nw.E_AREA_VECTOR)
        v = ctx.field_download(nw.F_DUAL_VOLUME)
        np.testing.assert_allclose(a, ra, rtol=0, atol=1e-13 * np.abs(ra).max())
        np.testing.assert_allclose(v, rv, rtol=1e-12, atol=0)
        assert (v > 0).all()
        assert abs(v.sum() - 1.0) < 1e-11
        clo = np.zeros_like(x)
        np.add.at(clo, W.edges[:, 0], a)
        np.add.at(clo, W.edges[:, 1], -a)
        inner = np.all((x > 1e-9) & (x < 1 - 1e-9), axis=1)
        assert np.abs(clo[inner]).max() < 1e-12 * np.abs(a).max()
