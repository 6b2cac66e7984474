"""GPU pressure-Poisson solve through the C ABI: FP64 GMRES and the
FP32-inner / FP64-refined form against a dense direct solve and the oracle
GMRES on small continuity systems; stopping rule, error paths."""
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


def _pressure_system(nw, ctx, nx=8, stretch=1.15, jitter=0.2, seed=2):
    """Continuity system with Dirichlet p = 0 on the top face z = 1."""
    P = pb.Problem(nx, nx, nx, stretch=stretch, jitter=jitter, seed=seed)
    P.upload(ctx)
    ctx.build_graph(1)
    o = nw.continuity_defaults()
    ctx.assemble_continuity(o)
    top = np.where(np.abs(P.coords[:, 2] - 1.0) < 1e-12)[0].astype(np.int32)
    ctx.field_upload(nw.F_PRESSURE, np.zeros(P.n))
    ctx.field_upload(nw.F_DIFF_FLUX_COEFF, np.zeros(P.n))
    ctx.apply_dirichlet(top, nw.F_PRESSURE, nw.F_DIFF_FLUX_COEFF, 0, 1)
    A = P.oracle_continuity(pb.to_orc_continuity(o))
    A.apply_dirichlet(top, 0, 1, np.zeros(P.n), np.zeros(P.n))
    return P, A


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_fp64_gmres_matches_direct_solve(nw, ctx, prec):
    P, A = _pressure_system(nw, ctx)
    xd = orc.dense_solve(A.dense(), A.rhs)
    cfg = nw.solver_cfg(kspace=20, max_iters=4000, tol=1e-10, precond=prec,
                        cheb_degree=3, cheb_ratio=20.0)
    st = ctx.solve(cfg)
    x = ctx.solution()
    assert st.converged and st.final_rel <= 1e-10
    # explicit residual recomputed here in FP64
    r = A.rhs - A.spmv(x)
    assert np.linalg.norm(r) <= 1.0001e-10 * np.linalg.norm(A.rhs)
    np.testing.assert_allclose(x, xd, rtol=0, atol=1e-7 * np.abs(xd).max())
    # same method on the oracle reaches the same tolerance
    xo, so = A.gmres(A.rhs, kspace=20, max_iters=4000, tol=1e-10, precond=prec,
                     cheb_degree=3, cheb_ratio=20.0)
    assert so.converged
    np.testing.assert_allclose(x, xo, rtol=0, atol=1e-7 * np.abs(xd).max())


@pytest.mark.parametrize("prec", [1, 2])
def test_fp32_refined_reaches_fp64_tolerance(nw, ctx, prec):
    P, A = _pressure_system(nw, ctx, nx=12)
    base = dict(kspace=30, max_iters=20000, tol=1e-8, precond=prec,
                cheb_degree=3, cheb_ratio=20.0)
    s64 = ctx.solve(nw.solver_cfg(**base))
    x64 = ctx.solution()
    s32 = ctx.solve(nw.solver_cfg(mode=nw.SOLVE_FP32_REFINE, inner_tol=1e-4, **base))
    x32 = ctx.solution()
    assert s64.converged and s32.converged
    assert s32.final_rel <= 1e-8
    r = A.rhs - A.spmv(x32)
    assert np.linalg.norm(r) <= 1.0001e-8 * np.linalg.norm(A.rhs)
    # both solutions agree to the tolerance-implied accuracy
    xd = orc.dense_solve(A.dense(), A.rhs) if P.n <= 3000 else x64
    err = np.linalg.norm(x32 - xd) / np.linalg.norm(xd)
    assert err < 1e-5, err


def test_solve_from_initial_guess(nw, ctx):
    P, A = _pressure_system(nw, ctx)
    cfg = nw.solver_cfg(kspace=20, max_iters=4000, tol=1e-10)
    s0 = ctx.solve(cfg)
    x = ctx.solution()
    ctx.set_solution(x)
    cfg.x0 = nw.X0_SOLUTION
    s1 = ctx.solve(cfg)
    assert s1.converged and s1.iters == 0


def test_non_convergence_is_reported(nw, ctx):
    P, A = _pressure_system(nw, ctx)
    cfg = nw.solver_cfg(kspace=5, max_iters=3, tol=1e-12)
    with pytest.raises(nw.NwError) as ei:
        ctx.solve(cfg)
    assert ei.value.code == nw.E_NOCONV
    st = ctx.solve(cfg, raise_on_noconv=False)
    assert not st.converged and st.iters == 3 and st.final_rel > 1e-12


def test_zero_diagonal_is_an_error(nw, ctx):
    P, A = _pressure_system(nw, ctx)
    ctx.reset_rows(np.array([3], np.int32), 0, 1, 0.0, 0.0)
    with pytest.raises(nw.NwError) as ei:
        ctx.solve(nw.solver_cfg())
    assert ei.value.code == nw.E_NONFINITE


def test_spmv_matches_oracle(nw, ctx):
    """y = A x on library-owned device buffers (two node fields)."""
    P, A = _pressure_system(nw, ctx, nx=10)
    x = np.random.default_rng(1).normal(size=P.n)
    ctx.field_upload(nw.F_PRESSURE, x)
    ctx.spmv_device(ctx.field_ptr(nw.F_PRESSURE), ctx.field_ptr(nw.F_DIFF_FLUX_COEFF))
    ctx.synchronize()
    y = ctx.field_download(nw.F_DIFF_FLUX_COEFF)
    np.testing.assert_allclose(y, A.spmv(x), rtol=1e-13,
                               atol=1e-13 * np.abs(A.vals).max() * np.abs(x).max())


@pytest.mark.parametrize("mode", [0, 1])
def test_amg_gmres_matches_direct_solve(nw, ctx, mode):
    P, A = _pressure_system(nw, ctx, nx=10)
    xd = orc.dense_solve(A.dense(), A.rhs)
    cfg = nw.solver_cfg(kspace=30, max_iters=500, tol=1e-10, precond=nw.PREC_AMG,
                        mode=mode, reuse_preconditioner=0)
    st = ctx.solve(cfg)
    x = ctx.solution()
    assert st.converged and st.precond_levels >= 2
    r = A.rhs - A.spmv(x)
    assert np.linalg.norm(r) <= 1.0001e-10 * np.linalg.norm(A.rhs)
    np.testing.assert_allclose(x, xd, rtol=0, atol=1e-7 * np.abs(xd).max())


@pytest.mark.parametrize("mode", [0, 1])
def test_amg_on_a_system_within_the_coarse_size(nw, ctx, mode):
    """A system of at most 1,000 rows is its own coarsest level: the
    preconditioner is its dense inverse, so GMRES converges at once (in
    FP32-inner form, one FP32 cycle per refinement)."""
    P, A = _pressure_system(nw, ctx, nx=6)  # 343 rows
    xd = orc.dense_solve(A.dense(), A.rhs)
    st = ctx.solve(nw.solver_cfg(kspace=20, max_iters=50, tol=1e-10, precond=nw.PREC_AMG,
                                 mode=mode, reuse_preconditioner=0))
    x = ctx.solution()
    assert st.converged and st.precond_levels == 1 and st.setup_ms > 0
    assert st.iters <= (2 if mode == 0 else 6), st.iters
    r = A.rhs - A.spmv(x)
    assert np.linalg.norm(r) <= 1.0001e-10 * np.linalg.norm(A.rhs)
    np.testing.assert_allclose(x, xd, rtol=0, atol=1e-7 * np.abs(xd).max())


def test_amg_levels_and_fp32_copy(nw, ctx):
    """The hierarchy's level sizes shrink and the FP32 solve made from the
    same setup (its copy narrowed on the device, counted as setup of that
    call) reaches the FP64 tolerance by refinement."""
    P, A = _pressure_system(nw, ctx, nx=24, stretch=1.05, jitter=0.1)
    s64 = ctx.solve(nw.solver_cfg(kspace=50, max_iters=300, tol=1e-9,
                                  precond=nw.PREC_AMG, reuse_preconditioner=0))
    lv = ctx.amg_levels()
    assert len(lv) == s64.precond_levels >= 2
    assert lv[0] == (P.n, A.vals.size)
    assert all(b[0] < a[0] for a, b in zip(lv, lv[1:])) and lv[-1][0] <= 1000
    s32 = ctx.solve(nw.solver_cfg(kspace=50, max_iters=300, tol=1e-9, precond=nw.PREC_AMG,
                                  mode=nw.SOLVE_FP32_REFINE, inner_tol=1e-4,
                                  reuse_preconditioner=1))
    assert s32.converged and 0 < s32.setup_ms < s64.setup_ms
    r = A.rhs - A.spmv(ctx.solution())
    assert np.linalg.norm(r) <= 1.0001e-9 * np.linalg.norm(A.rhs)
    s32b = ctx.solve(nw.solver_cfg(kspace=50, max_iters=300, tol=1e-9, precond=nw.PREC_AMG,
                                   mode=nw.SOLVE_FP32_REFINE, inner_tol=1e-4,
                                   reuse_preconditioner=1))
    assert s32b.setup_ms == 0 and s32b.iters == s32.iters


def test_amg_beats_jacobi_and_is_reused(nw, ctx):
    """On a 32^3 box stretched 1.05 the AMG V-cycle needs a small, bounded
    number of GMRES iterations; a second solve reuses the hierarchy."""
    P, A = _pressure_system(nw, ctx, nx=32, stretch=1.05, jitter=0.0)
    cj = nw.solver_cfg(kspace=50, max_iters=3000, tol=1e-8, precond=nw.PREC_JACOBI)
    sj = ctx.solve(cj, raise_on_noconv=False)
    ca = nw.solver_cfg(kspace=50, max_iters=300, tol=1e-8, precond=nw.PREC_AMG,
                       reuse_preconditioner=0)
    s1 = ctx.solve(ca)
    assert s1.converged and s1.iters <= 40 and s1.iters * 5 < sj.iters
    assert s1.setup_ms > 0
    ca.reuse_preconditioner = 1
    s2 = ctx.solve(ca)
    assert s2.converged and s2.setup_ms == 0 and s2.iters == s1.iters
    r = A.rhs - A.spmv(ctx.solution())
    assert np.linalg.norm(r) <= 1.0001e-8 * np.linalg.norm(A.rhs)


@pytest.mark.parametrize("mode", [0, 1])
def test_mixed_mesh_pressure_solve(nw, ctx, mode):
    """Pressure system of a conforming mixed Hex8/Tet4/Wed6/Pyr5 mesh
    (RCM-numbered): AMG-preconditioned GMRES, FP64 and FP32-inner/FP64-
    refined, against the dense solve."""
    P = pb.Problem(10, 0, 0, jitter=0.1, seed=8, mixed=(0.6, 0.3))
    P.upload(ctx)
    ctx.build_graph(1)
    o = nw.continuity_defaults()
    ctx.assemble_continuity(o)
    top = np.where(P.coords[:, 2] >= 1.0 - 1e-12)[0].astype(np.int32)
    ctx.field_upload(nw.F_PRESSURE, np.zeros(P.n))
    ctx.field_upload(nw.F_DIFF_FLUX_COEFF, np.zeros(P.n))
    ctx.apply_dirichlet(top, nw.F_PRESSURE, nw.F_DIFF_FLUX_COEFF, 0, 1)
    A = P.oracle_continuity(pb.to_orc_continuity(o))
    A.apply_dirichlet(top, 0, 1, np.zeros(P.n), np.zeros(P.n))
    xd = orc.dense_solve(A.dense(), A.rhs)
    st = ctx.solve(nw.solver_cfg(kspace=50, max_iters=500, tol=1e-9,
                                 precond=nw.PREC_AMG, mode=mode, inner_tol=1e-5,
                                 reuse_preconditioner=0))
    x = ctx.solution()
    assert st.converged and st.final_rel <= 1e-9
    r = A.rhs - A.spmv(x)
    assert np.linalg.norm(r) <= 1.0001e-9 * np.linalg.norm(A.rhs)
    np.testing.assert_allclose(x, xd, rtol=0, atol=1e-6 * np.abs(xd).max())
