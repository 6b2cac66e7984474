"""The pressure-projection step around the continuity solve
(LowMachEquationSystem::solve_and_update, src/LowMachEquationSystem.C:
793-819) through the C ABI against the oracle: continuity assembly with the
current p and dp/dx, Dirichlet rows, the solve for dp, p += dp
(solution_update), MdotEdgeAlg and the nodal velocity projection, over a
few steps of a time-evolving flow; and the 256^3 projection workload of
the learned-guess benchmark checked by its residuals."""
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


def test_solution_update_matches_oracle(nw, ctx):
    P = pb.Problem(5, 4, 6, stretch=1.1, jitter=0.1, seed=2)
    P.upload(ctx)
    ctx.build_graph(1)
    rng = np.random.default_rng(3)
    x = rng.normal(size=P.n)
    ctx.set_solution(x)
    # one rounding (fused multiply-add) on the device, two in numpy: within
    # an ulp of the larger product
    ctx.solution_update(0.7, nw.LINSYS_SOLUTION, -1.3, nw.F_PRESSURE)
    np.testing.assert_allclose(ctx.field_download(nw.F_PRESSURE),
                               orc.solution_update(0.7, x, -1.3, P.p), rtol=0,
                               atol=4.5e-16 * (np.abs(0.7 * x) + np.abs(1.3 * P.p)).max())
    ctx.solution_update(2.0, nw.F_DPDX, 0.5, nw.F_DQDX)
    np.testing.assert_allclose(ctx.field_download(nw.F_DQDX),
                               orc.solution_update(2.0, P.dpdx, 0.5, P.dqdx), rtol=0,
                               atol=4.5e-16 * (np.abs(2 * P.dpdx) + np.abs(0.5 * P.dqdx)).max())
    with pytest.raises(nw.NwError):   # 3-wide field, 1-dof system
        ctx.solution_update(1.0, nw.LINSYS_SOLUTION, 1.0, nw.F_VELOCITY)
    with pytest.raises(nw.NwError):   # shapes differ
        ctx.solution_update(1.0, nw.F_PRESSURE, 1.0, nw.F_VELOCITY)


def test_projection_steps_match_oracle(nw, ctx):
    """Four projection steps; after each, p, dp/dx, u and mdot against the
    oracle's own steps (continuity assembly + Dirichlet rows, dense solve,
    solution_update, MdotEdgeAlg, project_nodal_velocity)."""
    import workloads as wl
    P = pb.Problem(7, 6, 8, stretch=1.1, jitter=0.15, seed=9)
    n = P.n
    dt = 0.05
    o = nw.continuity_defaults(dt=dt, gamma1=1.0)
    top = np.where(np.abs(P.coords[:, 2] - 1.0) < 1e-12)[0].astype(np.int32)
    rho, udiag = P.rho, np.full(n, 1.0 / dt)
    p, dpdx = np.zeros(n), np.zeros((n, 3))
    P.p, P.dpdx, P.udiag = p.copy(), dpdx.copy(), udiag
    P.upload(ctx)
    ctx.field_upload(nw.F_PRESSURE_BC, np.zeros(n))
    ctx.build_graph(1)
    modes = wl.fourier_modes(5)
    cfg = nw.solver_cfg(kspace=50, max_iters=4000, tol=1e-13, precond=nw.PREC_JACOBI)
    for k in range(4):
        t = 0.1 * k
        ctx.fourier_velocity(modes, t)
        vel = wl.fourier_velocity(P.coords, 5, t=t)
        np.testing.assert_allclose(ctx.field_download(nw.F_VELOCITY), vel,
                                   rtol=0, atol=1e-12)
        vel = ctx.field_download(nw.F_VELOCITY)
        ctx.assemble_continuity(o)
        ctx.apply_dirichlet(top, nw.F_PRESSURE, nw.F_PRESSURE_BC, 0, 1)
        A = orc.CsrSystem(n, P.edges, 1)
        orc.assemble_continuity(pb.to_orc_continuity(o), n, P.edges, P.coords, vel,
                                dpdx, rho, p, udiag, P.areav, csr=A)
        A.apply_dirichlet(top, 0, 1, p, np.zeros(n))
        got = ctx.download()
        # the matrix to the row-scaled 1e-12 bar; the rhs sums fluxes that
        # carry the pressure (|lhs| |p| each) and cancel to the divergence,
        # so its error is measured against that scale
        pb.assert_csr_close(got[:3] + (A.rhs.copy(),), A)
        rs = np.abs(A.vals).max() * max(np.abs(p).max(), 1.0) * 16 + np.abs(A.rhs).max()
        np.testing.assert_allclose(got[3], A.rhs, rtol=0, atol=1e-12 * rs)
        st = ctx.solve(cfg)
        assert st.converged
        dp = orc.dense_solve(A.dense(), A.rhs)
        ctx.solution_update(1.0, nw.LINSYS_SOLUTION, 1.0, nw.F_PRESSURE)
        ctx.compute_mdot(1.0, 1.0)
        dpdx_device = ctx.field_download(nw.F_DPDX)
        ctx.project_nodal_velocity()
        p = orc.solution_update(1.0, dp, 1.0, p)
        mdot = orc.mdot_edge_alg(P.edges, P.coords, vel, dpdx, rho, p, udiag, P.areav)
        dpdx_new, vel = orc.project_nodal_velocity(P.edges, P.areav, P.dualvol, p,
                                                   dpdx, vel, rho, udiag)
        # uTmp keeps the gradient the projection started from (a copy of the
        # device's own dp/dx, which follows the oracle's to the solve's
        # tolerance)
        np.testing.assert_array_equal(ctx.field_download(nw.F_UTMP), dpdx_device)
        np.testing.assert_allclose(dpdx_device, dpdx, rtol=0,
                                   atol=1e-9 * max(np.abs(dpdx).max(), 1e-300))
        dpdx = dpdx_new
        sc = np.abs(p).max()
        np.testing.assert_allclose(ctx.field_download(nw.F_PRESSURE), p, rtol=0,
                                   atol=1e-9 * sc)
        np.testing.assert_allclose(ctx.field_download(nw.F_DPDX), dpdx, rtol=0,
                                   atol=1e-9 * np.abs(dpdx).max())
        np.testing.assert_allclose(ctx.field_download(nw.F_VELOCITY), vel, rtol=0,
                                   atol=1e-9 * np.abs(vel).max())
        np.testing.assert_allclose(ctx.field_download(nw.E_MASS_FLOW_RATE), mdot,
                                   rtol=0, atol=1e-9 * np.abs(mdot).max())
        # the projected mass flow satisfies the discrete continuity equation
        # the solve imposed: its edge divergence vanishes off the top rows
        div = np.zeros(n)
        np.add.at(div, P.edges[:, 0], -mdot)
        np.add.at(div, P.edges[:, 1], mdot)
        inner = np.ones(n, bool)
        inner[top] = False
        assert np.abs(div[inner]).max() <= 1e-8 * np.abs(mdot).max()
